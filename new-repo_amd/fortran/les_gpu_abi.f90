!------------------------------------------------------------------------------
!! module: les_gpu_abi
!------------------------------------------------------------------------------
!! iso_c_binding interfaces of liblesgpu.so (C header include/les_gpu.h).
!! Every entry point returns an integer status mirroring the reference's
!! decomp_2d_abort codes (0 ok, 1 invalid option, 3 out of memory,
!! 4 double initialisation, 5 device runtime error, 9 NaN, 10 bad shape).
!------------------------------------------------------------------------------
module les_gpu_abi
  use iso_c_binding
  implicit none

  type, bind(C) :: les_params_c
     integer(c_int) :: nx, ny, nz, sgs_model, cs_count, dyn_init, lagran_init, wall_model
     integer(c_int) :: ubc_u_neu, ubc_v_neu, halo_mode, coriolis, precision, judge_every, dealias
     real(c_double) :: lx, ly, lz, dt, z0, pfx, pfy, cs0, coriol, ug, vg, ubc_u, ubc_v, ubc_w
     real(c_double) :: re, learned_zmin
  end type les_params_c

  ! field ids (enum les_field)
  integer(c_int), parameter :: LES_U=0, LES_V=1, LES_W=2
  integer(c_int), parameter :: LES_DUDX=3, LES_DUDY=4, LES_DUDZ=5, LES_DVDX=6, LES_DVDY=7, LES_DVDZ=8
  integer(c_int), parameter :: LES_DWDX=9, LES_DWDY=10, LES_DWDZ=11
  integer(c_int), parameter :: LES_P=12, LES_DPDX=13, LES_DPDY=14, LES_DPDZ=15
  integer(c_int), parameter :: LES_RHSX=16, LES_RHSY=17, LES_RHSZ=18, LES_RHSX_F=19, LES_RHSY_F=20, LES_RHSZ_F=21
  integer(c_int), parameter :: LES_TXX=22, LES_TXY=23, LES_TXZ=24, LES_TYY=25, LES_TYZ=26, LES_TZZ=27
  integer(c_int), parameter :: LES_CS_OPT2=28, LES_NU_T=29, LES_F_LM=30, LES_F_MM=31, LES_F_QN=32, LES_F_NN=33
  integer(c_int), parameter :: LES_BETA=34
  ! layer ids (enum les_layer)
  integer(c_int), parameter :: LES_LBC_U=0, LES_LBC_V=1, LES_USTAR=2, LES_BOTTOM_WK=3, LES_TOP_WK=4, LES_Z0=5

  interface
     integer(c_int) function les_init(p) bind(C, name="les_init")
       import :: c_int, les_params_c
       type(les_params_c), intent(in) :: p
     end function les_init
     integer(c_int) function les_finalize() bind(C, name="les_finalize")
       import :: c_int
     end function les_finalize
     type(c_ptr) function les_last_error() bind(C, name="les_last_error")
       import :: c_ptr
     end function les_last_error
     integer(c_int) function les_set_ic(u, v, w) bind(C, name="les_set_ic")
       import :: c_int, c_double
       real(c_double), intent(in) :: u(*), v(*), w(*)
     end function les_set_ic
     integer(c_int) function les_upload_field(id, a) bind(C, name="les_upload_field")
       import :: c_int, c_double
       integer(c_int), value :: id
       real(c_double), intent(in) :: a(*)
     end function les_upload_field
     integer(c_int) function les_download_field(id, a) bind(C, name="les_download_field")
       import :: c_int, c_double
       integer(c_int), value :: id
       real(c_double), intent(out) :: a(*)
     end function les_download_field
     integer(c_int) function les_upload_layer(id, a) bind(C, name="les_upload_layer")
       import :: c_int, c_double
       integer(c_int), value :: id
       real(c_double), intent(in) :: a(*)
     end function les_upload_layer
     integer(c_int) function les_download_layer(id, a) bind(C, name="les_download_layer")
       import :: c_int, c_double
       integer(c_int), value :: id
       real(c_double), intent(out) :: a(*)
     end function les_download_layer
     integer(c_int) function les_step(n) bind(C, name="les_step")
       import :: c_int
       integer(c_int), value :: n
     end function les_step
     integer(c_int) function les_sync() bind(C, name="les_sync")
       import :: c_int
     end function les_sync
     integer(c_int) function les_set_jt_total(jt) bind(C, name="les_set_jt_total")
       import :: c_int
       integer(c_int), value :: jt
     end function les_set_jt_total
     integer(c_int) function les_check_nan(found) bind(C, name="les_check_nan")
       import :: c_int
       integer(c_int), intent(out) :: found
     end function les_check_nan
     integer(c_int) function les_wall_law() bind(C, name="les_wall_law")
       import :: c_int
     end function les_wall_law
     integer(c_int) function les_compute_vel_grad() bind(C, name="les_compute_vel_grad")
       import :: c_int
     end function les_compute_vel_grad
     integer(c_int) function les_grad_velocity_filt(f, dfdx, dfdy) bind(C, name="les_grad_velocity_filt")
       import :: c_int
       integer(c_int), value :: f, dfdx, dfdy
     end function les_grad_velocity_filt
     integer(c_int) function les_ddz_uv(dfdz, f) bind(C, name="les_ddz_uv")
       import :: c_int
       integer(c_int), value :: dfdz, f
     end function les_ddz_uv
     integer(c_int) function les_ddz_w(dfdz, f) bind(C, name="les_ddz_w")
       import :: c_int
       integer(c_int), value :: dfdz, f
     end function les_ddz_w
     integer(c_int) function les_sgs_stress(jt, jtt) bind(C, name="les_sgs_stress")
       import :: c_int
       integer(c_int), value :: jt, jtt
     end function les_sgs_stress
     integer(c_int) function les_compute_dealias_convec(cx, cy, cz) bind(C, name="les_compute_dealias_convec")
       import :: c_int
       integer(c_int), value :: cx, cy, cz
     end function les_compute_dealias_convec
     integer(c_int) function les_div_stress(rhs, tx, ty, tz, stag, save_wk) bind(C, name="les_div_stress")
       import :: c_int
       integer(c_int), value :: rhs, tx, ty, tz, stag, save_wk
     end function les_div_stress
     integer(c_int) function les_compute_rhs_adv() bind(C, name="les_compute_rhs_adv")
       import :: c_int
     end function les_compute_rhs_adv
     integer(c_int) function les_press_stag_array(tadv1) bind(C, name="les_press_stag_array")
       import :: c_int, c_double
       real(c_double), value :: tadv1
     end function les_press_stag_array
     integer(c_int) function les_test_filter(f, alpha) bind(C, name="les_test_filter")
       import :: c_int
       integer(c_int), value :: f, alpha
     end function les_test_filter
     integer(c_int) function les_ab2(tadv1, tadv2) bind(C, name="les_ab2")
       import :: c_int, c_double
       real(c_double), value :: tadv1, tadv2
     end function les_ab2
     integer(c_int) function les_projection(tadv1) bind(C, name="les_projection")
       import :: c_int, c_double
       real(c_double), value :: tadv1
     end function les_projection
     integer(c_int) function les_stat_init() bind(C, name="les_stat_init")
       import :: c_int
     end function les_stat_init
     integer(c_int) function les_stat_reset() bind(C, name="les_stat_reset")
       import :: c_int
     end function les_stat_reset
     integer(c_int) function les_stat_accumulate() bind(C, name="les_stat_accumulate")
       import :: c_int
     end function les_stat_accumulate
     integer(c_int) function les_stat_download(sums3d, ustar2d, iterations) bind(C, name="les_stat_download")
       import :: c_int, c_double
       real(c_double), intent(out) :: sums3d(*), ustar2d(*)
       integer(c_int), intent(out) :: iterations
     end function les_stat_download
     integer(c_int) function les_stat_profiles(prof) bind(C, name="les_stat_profiles")
       import :: c_int, c_double
       real(c_double), intent(out) :: prof(*)
     end function les_stat_profiles
     integer(c_int) function les_set_geostrophic(ug, vg) bind(C, name="les_set_geostrophic")
       import :: c_int, c_double
       real(c_double), value :: ug, vg
     end function les_set_geostrophic
     integer(c_int) function les_tke_init() bind(C, name="les_tke_init")
       import :: c_int
     end function les_tke_init
     integer(c_int) function les_tke_reset() bind(C, name="les_tke_reset")
       import :: c_int
     end function les_tke_reset
     integer(c_int) function les_tke_accumulate() bind(C, name="les_tke_accumulate")
       import :: c_int
     end function les_tke_accumulate
     integer(c_int) function les_tke_download(sums) bind(C, name="les_tke_download")
       import :: c_int, c_double
       real(c_double), intent(out) :: sums(*)
     end function les_tke_download
     integer(c_int) function les_running_diagnostics(out) bind(C, name="les_running_diagnostics")
       import :: c_int, c_double
       real(c_double), intent(out) :: out(4)
     end function les_running_diagnostics
  end interface

  integer(c_int), parameter :: LES_NSTAT = 26, LES_NPROF = 21, LES_NTKE = 61

contains

  !! stop with the library's message when a call fails (the reference calls
  !! decomp_2d_abort(code, msg) -> MPI_ABORT)
  subroutine les_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: i
    if (rc == 0) return
    call c_f_pointer(les_last_error(), msg, [512])
    text = ' '
    do i = 1, 512
       if (msg(i) == c_null_char) exit
       text(i:i) = msg(i)
    end do
    write(*,'(A,I3,A,A,A,A)') 'ABORT - code ', rc, ' in ', trim(where), ': ', trim(text)
    error stop 1
  end subroutine les_check

end module les_gpu_abi
