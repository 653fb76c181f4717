!------------------------------------------------------------------------------
!! module: parameters_IO
!------------------------------------------------------------------------------
!! Reads the reference's parameters.nlst (same six namelist groups and
!! variable names, src/parameters_IO.f90:151-166) and nondimensionalises it
!! the same way (src/parameters_IO.f90:240-298).  An optional seventh group
!! &les_gpu_param (precision, halo_mode, judge_every, learned_zmin, dealias) selects
!! the device variant; dealias = 'PADD' | 'TRUNC' | 'FSMTHG' | 'NODLSG' replaces the
!! reference's compile-time DSLG choice (src/Makefile:17-40).  read_parameters fills les_params_c for les_init.
!------------------------------------------------------------------------------
module parameters_IO
  use iso_c_binding
  use les_gpu_abi
  implicit none

  integer, parameter, public :: rprec = kind(1.d0)
  real(rprec), parameter, public :: pi = 3.1415926535897932384626433_rprec
  real(rprec), parameter, public :: BOGUS = -1.23456789E20_rprec

  ! simulation_param
  integer, public :: p_on_y = 1, p_on_z = 1, nx, ny, nz, nt
  real(rprec), public :: dt
  logical, public :: dim_dt_flag = .false., restart_flag = .false.
  ! output_param
  character(256), public :: out_path = './'
  logical, public :: output = .false., stat_flag = .false., tke_flag = .false.
  integer, public :: step_out = 1000000, start_stats = 0, running_period = 100
  integer, public :: start_tke = 0, tke_period = 100, checkpoint = 1000000, scheckpoint = 1000000
  ! physical_param
  logical, public :: XY_pi = .true., coriolis_flag = .false., geostr_rot = .false.
  real(rprec), public :: lx = 2._rprec, ly = 2._rprec, lz = 1._rprec, u_scale = 1._rprec, zi = 1._rprec
  real(rprec), public :: lbc_dz0 = 1.e-4_rprec, Re = 1.e5_rprec, pfx = 1._rprec, pfy = 0._rprec
  real(rprec), public :: ug_dim = 0._rprec, vg_dim = 0._rprec
  integer, public :: geostr_start = 100000
  ! sgs_param
  real(rprec), public :: Cs0 = 0.16_rprec
  character(16), public :: sgs_model = 'cst_Smag'
  integer, public :: cs_count = 5, dyn_init = 1
  ! momentum_bc
  character(25), public :: lbc_model = 'std', lbc_special = 'wall_law', ubc_model = 'std', ubc_special = 'std'
  character(16), public :: wall_model = 'wall_filter'
  logical, public :: surface_files = .false.
  ! momentum_ic
  character(16), public :: ic_type = 'files'
  real(rprec), public :: u0 = 0._rprec, v0 = 0._rprec, w0 = 0._rprec, u_star = 1._rprec
  real(rprec), public :: dudz0 = 0._rprec, z_turb = 0.5_rprec, vnf = 1._rprec
  ! les_gpu_param (device variant)
  integer, public :: precision = 0, halo_mode = 0, judge_every = 50
  real(rprec), public :: learned_zmin = 16._rprec
  character(8), public :: dealias = 'PADD'
  ! derived
  real(rprec), public :: dx, dy, dz
  ! the nondimensional geostrophic wind (rotated by compute_geostrophic_correct)
  ! and the dimensional time step (src/parameters_IO.f90:71,83,258-271)
  real(rprec), public :: ug = 0._rprec, vg = 0._rprec, dt_dim
  ! the nondimensional roughness per surface point (set_bc_surface,
  ! src/boundary_condition_mpi.f90:100-130): lbc_dz0 / zi, or
  ! ./input/surface_z0.dat / zi with surface_files
  real(rprec), allocatable, public :: lbc_z0(:,:)
  logical, public :: Lagran_init = .false.

  type(les_params_c), public :: cparams

  public :: read_parameters, write_parameters_restart, read_parameters_restart

contains

  subroutine read_parameters()
    integer :: ios
    namelist / simulation_param / p_on_y, p_on_z, nx, ny, nz, nt, dt, dim_dt_flag, restart_flag
    namelist / output_param / out_path, output, step_out, stat_flag, start_stats, running_period, &
         tke_flag, start_tke, tke_period, checkpoint, scheckpoint
    namelist / physical_param / XY_pi, lx, ly, lz, u_scale, zi, lbc_dz0, Re, &
         pfx, pfy, coriolis_flag, ug_dim, vg_dim, geostr_rot, geostr_start
    namelist / sgs_param / Cs0, sgs_model, cs_count, dyn_init
    namelist / momentum_bc / lbc_model, lbc_special, ubc_model, ubc_special, wall_model, surface_files
    namelist / momentum_ic / ic_type, u0, v0, w0, u_star, dudz0, z_turb, vnf
    namelist / les_gpu_param / precision, halo_mode, judge_every, learned_zmin, dealias

    open(88, file='parameters.nlst', status='OLD', action='READ', delim='APOSTROPHE', iostat=ios)
    if (ios /= 0) call abort22('Could not open parameters.nlst for reading')
    read(88, simulation_param, iostat=ios); if (ios /= 0) call abort22('Error in namelist simulation_param')
    read(88, output_param, iostat=ios);     if (ios /= 0) call abort22('Error in namelist output_param')
    read(88, physical_param, iostat=ios);   if (ios /= 0) call abort22('Error in namelist physical_param')
    read(88, sgs_param, iostat=ios);        if (ios /= 0) call abort22('Error in namelist sgs_param')
    read(88, momentum_bc, iostat=ios);      if (ios /= 0) call abort22('Error in namelist momentum_bc')
    read(88, momentum_ic, iostat=ios);      if (ios /= 0) call abort22('Error in namelist momentum_ic')
    read(88, les_gpu_param, iostat=ios)     ! optional group
    close(88)

    if (XY_pi) then
       lx = lx*pi; ly = ly*pi
    end if
    lx = lx/zi; ly = ly/zi; lz = lz/zi
    if (dim_dt_flag) then
       dt_dim = dt
       dt = dt*u_scale/zi
    else
       dt_dim = dt/u_scale*zi
    end if
    dx = lx/real(nx, rprec); dy = ly/real(ny, rprec); dz = lz/real(nz, rprec)

    cparams%nx = nx; cparams%ny = ny; cparams%nz = nz
    if (trim(sgs_model) == 'dyn_LagScaleDep') then
       cparams%sgs_model = 2
    else if (trim(sgs_model) == 'dyn_Smag') then
       cparams%sgs_model = 1
    else if (trim(sgs_model) == 'dns') then
       cparams%sgs_model = 3
    else if (trim(sgs_model) == 'cst_Smag') then
       cparams%sgs_model = 0
    else
       write(*,*) 'ABORT - Invalid SGS model ', trim(sgs_model)
       error stop 1
    end if
    if (trim(lbc_special) /= 'wall_law') then
       write(*,*) 'ABORT - only lbc_special = wall_law is supported'
       error stop 1
    end if
    cparams%cs_count = cs_count; cparams%dyn_init = dyn_init
    cparams%lagran_init = 0
    if (trim(wall_model) == 'wall_filter') then
       cparams%wall_model = 0
    else if (trim(wall_model) == 'wall_inst') then
       cparams%wall_model = 1
    else
       write(*,*) 'ABORT - Invalid wall law type ', trim(wall_model)
       error stop 1
    end if
    cparams%ubc_u_neu = 1; cparams%ubc_v_neu = 1
    cparams%halo_mode = halo_mode; cparams%precision = precision; cparams%judge_every = judge_every
    select case (trim(dealias))
    case ('PADD');   cparams%dealias = 0
    case ('TRUNC');  cparams%dealias = 1
    case ('FSMTHG'); cparams%dealias = 2
    case ('NODLSG'); cparams%dealias = 3
    case default
       write(*,*) 'ABORT - Invalid dealiasing option ', trim(dealias)
       error stop 1
    end select
    cparams%learned_zmin = learned_zmin
    cparams%re = Re
    cparams%coriolis = 0; cparams%coriol = 0._rprec; cparams%ug = 0._rprec; cparams%vg = 0._rprec
    if (coriolis_flag) then
       cparams%coriolis = 1
       cparams%coriol = 1.067E-04*zi/u_scale
       ug = ug_dim/u_scale; vg = vg_dim/u_scale
       cparams%ug = ug; cparams%vg = vg
    end if
    cparams%lx = lx; cparams%ly = ly; cparams%lz = lz; cparams%dt = dt
    cparams%z0 = lbc_dz0/zi; cparams%pfx = pfx; cparams%pfy = pfy; cparams%cs0 = Cs0
    cparams%ubc_u = 0._rprec; cparams%ubc_v = 0._rprec; cparams%ubc_w = 0._rprec
    if (allocated(lbc_z0)) deallocate(lbc_z0)
    allocate(lbc_z0(nx, ny))
    if (surface_files) then
       block
         real(rprec) :: z0_in(nx, ny)
         integer :: rl, ios
         inquire(iolength=rl) z0_in
         open(98, file='./input/surface_z0.dat', form='unformatted', access='direct', recl=rl, status='old', &
              action='read', iostat=ios)
         if (ios /= 0) then
            write(*,*) 'ABORT - cannot open ./input/surface_z0.dat'
            error stop 1
         end if
         read(98, rec=1) z0_in
         close(98)
         lbc_z0 = z0_in/zi
       end block
    else
       lbc_z0 = lbc_dz0/zi
    end if
  end subroutine read_parameters

  subroutine abort22(msg)
    character(*), intent(in) :: msg
    write(*,'(A,A)') '[ParametresIO error] ', msg
    error stop 22
  end subroutine abort22

  !! parameters_restart.nlst (src/parameters_IO.f90:333-435)
  subroutine write_parameters_restart(jt_total)
    integer, intent(in) :: jt_total
    namelist / restart_param / jt_total_out, dyn_init_out, Lagran_init_out, sgs_model_out
    integer :: jt_total_out, dyn_init_out
    logical :: Lagran_init_out
    character(16) :: sgs_model_out
    jt_total_out = jt_total; dyn_init_out = dyn_init; Lagran_init_out = Lagran_init; sgs_model_out = sgs_model
    open(89, file=trim(out_path)//'parameters_restart.nlst', status='replace', action='write', delim='APOSTROPHE')
    write(89, restart_param)
    close(89)
  end subroutine write_parameters_restart

  subroutine read_parameters_restart(jt_total)
    integer, intent(out) :: jt_total
    namelist / restart_param / jt_total_out, dyn_init_out, Lagran_init_out, sgs_model_out
    integer :: jt_total_out, dyn_init_out, ios
    logical :: Lagran_init_out
    character(16) :: sgs_model_out
    open(89, file='parameters_restart.nlst', status='old', action='read', delim='APOSTROPHE', iostat=ios)
    if (ios /= 0) call abort22('Could not open parameters_restart.nlst')
    read(89, restart_param, iostat=ios)
    if (ios /= 0) call abort22('Error in namelist restart_param')
    close(89)
    jt_total = jt_total_out
    Lagran_init = Lagran_init_out
    if (Lagran_init) cparams%lagran_init = 1
  end subroutine read_parameters_restart

end module parameters_IO
