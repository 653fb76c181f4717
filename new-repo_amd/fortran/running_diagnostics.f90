!------------------------------------------------------------------------------
!! module: running_diagnositcs
!------------------------------------------------------------------------------
!! Running diagnostics of the device state (src/running_diagnostics.f90:
!! 49-136): CFL, mean kinetic energy, plane-mean u* and rms divergence come
!! from one device reduction (les_running_diagnostics); this module keeps the
!! reference's module and procedure names, its console line every
!! running_diag_promt steps and its 100-line buffer written to
!! ./running_diagnostics.txt (replaced when the first full buffer is written
!! at jt_total = buffer_size, appended otherwise).
!------------------------------------------------------------------------------
module running_diagnositcs
  use iso_c_binding
  use les_gpu_abi
  use parameters_IO
  implicit none
  private

  integer, parameter :: buffer_size = 100
  integer, parameter :: running_diag_promt = 50
  integer, save :: buffer_pos = 0
  integer, save :: buffer_jt(buffer_size)
  real(rprec), save :: buffer_tt(buffer_size), buffer_cfl(buffer_size), buffer_mke(buffer_size)
  real(rprec), save :: buffer_u_star(buffer_size), buffer_divu(buffer_size)

  public :: running_diagnostics_init, compute_running_diagnistics

contains

  subroutine running_diagnostics_init()
    buffer_jt = 0
    buffer_tt = 0._rprec
    buffer_cfl = 0._rprec
    buffer_mke = 0._rprec
    buffer_u_star = 0._rprec
    buffer_divu = 0._rprec
    buffer_pos = 0
  end subroutine running_diagnostics_init

  subroutine compute_running_diagnistics(jt, jt_total, tt, flag_add2buffer)
    integer, intent(in) :: jt, jt_total
    real(rprec), intent(in) :: tt
    logical, intent(in) :: flag_add2buffer
    real(c_double) :: d(4)
    integer :: n

    call les_check(les_running_diagnostics(d), 'compute_running_diagnistics')
    if (jt == 0) then
       write(*,*) '## running diagnostics initilized - flow summary'
       write(*,'(A,I10,A,E11.4,A,E11.4,A,E11.4)') '    jt_t=', jt_total, ', t=', tt, ', CFL=', d(1), ', MKE=', d(2)
    else if (mod(jt, running_diag_promt) == 0) then
       write(*,'(A,I10,A,I10,A,E11.4,A,E11.4,A,E11.4,A,E11.4,A,E11.4)') ' jt=', jt, ', jt_t=', jt_total, &
            ', t=', tt, ', CFL=', d(1), ', MKE=', d(2), ', u_star=', d(3), ', divu=', d(4)
    end if

    if (flag_add2buffer) then
       buffer_pos = buffer_pos + 1
       buffer_jt(buffer_pos) = jt_total
       buffer_tt(buffer_pos) = tt
       buffer_cfl(buffer_pos) = d(1)
       buffer_mke(buffer_pos) = d(2)
       buffer_u_star(buffer_pos) = d(3)
       buffer_divu(buffer_pos) = d(4)
    end if

    if (buffer_pos == buffer_size) then
       if (jt_total == buffer_size) then
          open(98, file='./running_diagnostics.txt', status='replace')
       else
          open(98, file='./running_diagnostics.txt', position='append')
       end if
       do n = 1, buffer_size
          write(98,'(I10,E13.6,E16.9,E16.9,E16.9,E16.9)') buffer_jt(n), buffer_tt(n), buffer_cfl(n), &
               buffer_mke(n), buffer_u_star(n), buffer_divu(n)
       end do
       close(98)
       buffer_pos = 0
    end if
  end subroutine compute_running_diagnistics

end module running_diagnositcs
