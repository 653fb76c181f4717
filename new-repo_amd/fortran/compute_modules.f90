!------------------------------------------------------------------------------
!! Drop-in operator modules: the reference's compute_* module procedures
!! (same module and procedure names) as thin bodies over the device library.
!! Arguments that were arrays in the reference are device field handles.
!------------------------------------------------------------------------------

!! src/compute_derivatives.f90:94-577
module compute_derivatives
  use iso_c_binding
  use les_gpu_abi
  implicit none
  private
  public :: grad_velocity_filt, ddz_uv, ddz_w, div_stress
contains
  subroutine grad_velocity_filt(f, dfdx, dfdy)
    integer(c_int), intent(in) :: f, dfdx, dfdy
    call les_check(les_grad_velocity_filt(f, dfdx, dfdy), 'grad_velocity_filt')
  end subroutine grad_velocity_filt

  subroutine ddz_uv(dfdz, f)
    integer(c_int), intent(in) :: dfdz, f
    call les_check(les_ddz_uv(dfdz, f), 'ddz_uv')
  end subroutine ddz_uv

  subroutine ddz_w(dfdz, f)
    integer(c_int), intent(in) :: dfdz, f
    call les_check(les_ddz_w(dfdz, f), 'ddz_w')
  end subroutine ddz_w

  !! the pressure BC layers (bottom_wk_press, top_wk_press) stay on the device
  subroutine div_stress(rhs, tx, ty, tz, stag, save_wk_press)
    integer(c_int), intent(in) :: rhs, tx, ty, tz
    logical, intent(in) :: stag
    logical, intent(in), optional :: save_wk_press
    integer(c_int) :: s, wk
    s = 0; if (stag) s = 1
    wk = 0
    if (present(save_wk_press)) then
       if (save_wk_press) wk = 1
    end if
    call les_check(les_div_stress(rhs, tx, ty, tz, s, wk), 'div_stress')
  end subroutine div_stress
end module compute_derivatives

!! src/compute_convective_padd.f90:115-262
module compute_convective_padd
  use iso_c_binding
  use les_gpu_abi
  implicit none
  private
  public :: compute_dealias_convec
contains
  subroutine compute_dealias_convec(cx, cy, cz)
    integer(c_int), intent(in) :: cx, cy, cz
    call les_check(les_compute_dealias_convec(cx, cy, cz), 'compute_dealias_convec')
  end subroutine compute_dealias_convec
end module compute_convective_padd

!! src/compute_pressure.f90:116-242
module compute_pressure
  use iso_c_binding
  use les_gpu_abi
  implicit none
  private
  public :: press_stag_array
contains
  subroutine press_stag_array(tadv1)
    real(c_double), intent(in) :: tadv1
    call les_check(les_press_stag_array(tadv1), 'press_stag_array')
  end subroutine press_stag_array
end module compute_pressure

!! src/testfilter.f90:144-184 (G selected by the filter width: 2 -> G_test,
!! 4 -> G_test_test)
module testfilter
  use iso_c_binding
  use les_gpu_abi
  implicit none
  private
  integer(c_int), parameter, public :: G_test = 2, G_test_test = 4
  public :: test_filter
contains
  subroutine test_filter(f, G)
    integer(c_int), intent(in) :: f, G
    call les_check(les_test_filter(f, G), 'test_filter')
  end subroutine test_filter
end module testfilter

!! src/compute_sgs.f90:105-247 (includes compute_LagScaleDep_coeff)
module compute_sgs
  use iso_c_binding
  use les_gpu_abi
  implicit none
  private
  public :: sgs_stress
contains
  subroutine sgs_stress(jt, jtt)
    integer, intent(in) :: jt, jtt
    call les_check(les_sgs_stress(int(jt, c_int), int(jtt, c_int)), 'sgs_stress')
  end subroutine sgs_stress
end module compute_sgs

!! src/compute_wall_law.f90:107-254: obukhov computes u* and the wall
!! stress and gradient in one device pass; wallstress has nothing left to do.
module compute_wall_law
  use iso_c_binding
  use les_gpu_abi
  implicit none
  private
  public :: obukhov, wallstress
contains
  subroutine obukhov(wl_type)
    character(*), intent(in) :: wl_type
    if (trim(wl_type) /= 'wall_filter' .and. trim(wl_type) /= 'wall_inst') then
       write(*,*) 'ABORT - Invalid wall law type'
       error stop 1
    end if
    call les_check(les_wall_law(), 'obukhov')
  end subroutine obukhov

  subroutine wallstress(wl_type)
    character(*), intent(in) :: wl_type
    if (len_trim(wl_type) == 0) return
  end subroutine wallstress
end module compute_wall_law

!! src/compute_geostrophic.f90:24-45: the geostrophic wind turns with a
!! 72 h period from geostr_start on, held at its 6 h direction after that,
!! keeping its magnitude; the device's Coriolis forcing takes the new
!! (ug, vg) from the next operator on.
module compute_geostrophic
  use iso_c_binding
  use les_gpu_abi
  use parameters_IO
  implicit none
  private
  real(rprec), parameter :: t_period = 72.0_rprec*3600.0_rprec
  real(rprec), parameter :: up_limit = 6.0_rprec*3600.0_rprec
  public :: compute_geostrophic_correct
contains
  subroutine compute_geostrophic_correct(jt_total)
    integer, intent(in) :: jt_total
    real(rprec) :: ug_new, vg_new, curr_time
    curr_time = real((jt_total - geostr_start)*dt_dim, rprec)
    ug_new = sqrt(ug**2 + vg**2)*cos(2.0_rprec*pi*curr_time/t_period)
    vg_new = sqrt(ug**2 + vg**2)*sin(2.0_rprec*pi*curr_time/t_period)
    if (curr_time >= up_limit) then
       ug_new = sqrt(ug**2 + vg**2)*cos((2.0_rprec*pi*up_limit)/t_period)
       vg_new = sqrt(ug**2 + vg**2)*sin((2.0_rprec*pi*up_limit)/t_period)
    end if
    ug = ug_new
    vg = vg_new
    call les_check(les_set_geostrophic(ug, vg), 'compute_geostrophic_correct')
  end subroutine compute_geostrophic_correct
end module compute_geostrophic
