!------------------------------------------------------------------------------
!! module: system_variables
!------------------------------------------------------------------------------
!! The reference's global fields (src/system.f90:21-47) live in device
!! memory; here each name is the handle of its device field, so calls such as
!! grad_velocity_filt(u, dudx, dudy) read as in the reference.  Host mirrors
!! are fetched on demand with get_field / put_field (levels 0..nz+1).
!------------------------------------------------------------------------------
module system_variables
  use iso_c_binding
  use les_gpu_abi
  use parameters_IO, only: rprec, nx, ny, nz
  implicit none

  integer(c_int), parameter, public :: u = LES_U, v = LES_V, w = LES_W
  integer(c_int), parameter, public :: dudx = LES_DUDX, dudy = LES_DUDY, dudz = LES_DUDZ
  integer(c_int), parameter, public :: dvdx = LES_DVDX, dvdy = LES_DVDY, dvdz = LES_DVDZ
  integer(c_int), parameter, public :: dwdx = LES_DWDX, dwdy = LES_DWDY, dwdz = LES_DWDZ
  integer(c_int), parameter, public :: p = LES_P, dpdx = LES_DPDX, dpdy = LES_DPDY, dpdz = LES_DPDZ
  integer(c_int), parameter, public :: rhsx = LES_RHSX, rhsy = LES_RHSY, rhsz = LES_RHSZ
  integer(c_int), parameter, public :: rhsx_f = LES_RHSX_F, rhsy_f = LES_RHSY_F, rhsz_f = LES_RHSZ_F
  integer(c_int), parameter, public :: txx = LES_TXX, txy = LES_TXY, txz = LES_TXZ
  integer(c_int), parameter, public :: tyy = LES_TYY, tyz = LES_TYZ, tzz = LES_TZZ
  integer(c_int), parameter, public :: cs_opt2 = LES_CS_OPT2, nu_t = LES_NU_T
  integer(c_int), parameter, public :: F_LM = LES_F_LM, F_MM = LES_F_MM, F_QN = LES_F_QN, F_NN = LES_F_NN
  integer(c_int), parameter, public :: Beta_mum = LES_BETA

  public :: get_field, put_field

contains

  subroutine get_field(id, a)
    integer(c_int), intent(in) :: id
    real(rprec), intent(out) :: a(nx, ny, 0:nz+1)
    call les_check(les_download_field(id, a), 'get_field')
  end subroutine get_field

  subroutine put_field(id, a)
    integer(c_int), intent(in) :: id
    real(rprec), intent(in) :: a(nx, ny, 0:nz+1)
    call les_check(les_upload_field(id, a), 'put_field')
  end subroutine put_field

end module system_variables
