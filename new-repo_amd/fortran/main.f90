!------------------------------------------------------------------------------
!! program: main (src/main.f90:16-61) -- the device LES driver
!------------------------------------------------------------------------------
program main
  use LES_engine
  implicit none
  integer(8) :: c0, c1, r
  call system_clock(c0, r)
  call LES_main_engine()
  call system_clock(c1)
  write(*,'(A,F12.3,A)') ' total wall time: ', real(c1 - c0, 8)/real(r, 8), 's'
end program main
