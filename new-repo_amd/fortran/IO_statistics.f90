!------------------------------------------------------------------------------
!! module: mod_IO_statistics
!------------------------------------------------------------------------------
!! Running-average statistics (src/IO_statistics.f90:40-513) over the device
!! state: the sums live on the device (les_stat_*) and are copied to the host
!! only when a RAV file set is written.  Every running_period steps (and past
!! start_stats) it writes, like the reference, RAV_<jt>_info.txt (nx, ny, nz
!! and the number of accumulated steps), RAV_<jt>_2D.dat (the u* sum, float64
!! (nx, ny)) and RAV_<jt>_3D.dat (26 float64 (nx, ny, nz) sums in the order
!! u v w p u2 v2 w2 u3 v3 w3 u4 v4 w4 uv uw vw txx tyy tzz txy txz tyz dudz
!! dvdz nut cs), then starts the sums again.
!------------------------------------------------------------------------------
module mod_IO_statistics
  use iso_c_binding
  use les_gpu_abi
  use parameters_IO
  implicit none
  private

  public :: io_stat_init, io_stat_finalize, io_stat_averages

contains

  subroutine io_stat_init()
    call les_check(les_stat_init(), 'io_stat_init')
  end subroutine io_stat_init

  subroutine io_stat_finalize()
  end subroutine io_stat_finalize

  subroutine io_stat_averages(jt, jt_total)
    integer, intent(in) :: jt, jt_total
    real(rprec), allocatable :: sums(:), ustar(:)
    integer(c_int) :: iters
    character(512) :: filename

    call les_check(les_stat_accumulate(), 'io_stat_averages')
    if (mod(jt_total, running_period) == 0 .and. jt_total > start_stats) then
       allocate(sums(int(LES_NSTAT) * nx * ny * nz), ustar(nx * ny))
       call les_check(les_stat_download(sums, ustar, iters), 'io_stat_averages')
       write(*,*) '=============================================================================='
       write(*,'(A,I12)') '# writing momentum stats at jt_t=', jt_total
       write(*,*) '=============================================================================='
       write(filename,'(A,I9.9,A)') trim(out_path)//'RAV_', jt_total, '_info.txt'
       open(99, file=filename, status='replace')
       write(99,'(I6,I6,I6,I10)') nx, ny, nz, iters
       close(99)
       write(filename,'(A,I9.9,A)') trim(out_path)//'RAV_', jt_total, '_2D.dat'
       open(99, file=filename, form='unformatted', access='stream', status='replace')
       write(99) ustar
       close(99)
       write(filename,'(A,I9.9,A)') trim(out_path)//'RAV_', jt_total, '_3D.dat'
       open(99, file=filename, form='unformatted', access='stream', status='replace')
       write(99) sums
       close(99)
       deallocate(sums, ustar)
       call les_check(les_stat_reset(), 'io_stat_averages')
    end if
  end subroutine io_stat_averages

end module mod_IO_statistics
