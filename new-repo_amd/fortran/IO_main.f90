!------------------------------------------------------------------------------
!! module: mod_IO_main
!------------------------------------------------------------------------------
!! Field snapshots (src/IO_main.f90:24-114): 2decomp_sys_<jt_total>.dat holds
!! levels 1..nz of u v w p rhsx rhsy rhsz dudx dudy dudz dvdx dvdy dvdz dwdx
!! dwdy dwdz txx txy txz tyy tyz tzz nu_t cs_opt2, each float64 (nx, ny, nz)
!! in Fortran order, one after the other.  The device fields are copied to
!! the host one at a time (FP32 fields of the mixed modes widened).
!------------------------------------------------------------------------------
module mod_IO_main
  use iso_c_binding
  use les_gpu_abi
  use parameters_IO
  use system_variables
  implicit none
  private

  public :: io_init, io_finalize, output_loop

contains

  subroutine io_init()
  end subroutine io_init

  subroutine io_finalize()
  end subroutine io_finalize

  subroutine output_loop(jt, jt_total)
    integer, intent(in) :: jt, jt_total
    integer(c_int), parameter :: ids(24) = [u, v, w, p, rhsx, rhsy, rhsz, dudx, dudy, dudz, dvdx, dvdy, dvdz, &
         dwdx, dwdy, dwdz, txx, txy, txz, tyy, tyz, tzz, nu_t, cs_opt2]
    real(rprec), allocatable :: a(:,:,:)
    character(512) :: filename
    integer :: i

    write(*,*) 'output at jt=', jt
    write(filename,'(A,I9.9,A)') trim(out_path)//'2decomp_sys_', jt_total, '.dat'
    allocate(a(nx, ny, 0:nz+1))
    open(96, file=filename, form='unformatted', access='stream', status='replace')
    do i = 1, size(ids)
       call get_field(ids(i), a)
       write(96) a(:,:,1:nz)
    end do
    close(96)
    deallocate(a)
  end subroutine output_loop

end module mod_IO_main
