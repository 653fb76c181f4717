!------------------------------------------------------------------------------
!! module: LES_engine
!------------------------------------------------------------------------------
!! Host time loop of the device LES, following LES_main_engine
!! (src/LES_engine.f90:77-438): read parameters.nlst, initial fields
!! (module initial_conditions) or a checkpoint, then per step the wall
!! law, velocity gradients, SGS stress, right-hand side, AB2 advance,
!! pressure and projection, with the reference's running diagnostics
!! (running_diagnostics.txt), NaN check, running-average statistics (RAV
!! files, stat_flag), field snapshots (2decomp_sys files, output),
!! checkpoint files and profiling.txt.  All field data
!! stays on the device between steps.
!------------------------------------------------------------------------------
module LES_engine
  use iso_c_binding
  use les_gpu_abi
  use parameters_IO
  use system_variables
  use compute_derivatives
  use compute_convective_padd
  use compute_pressure
  use compute_sgs
  use compute_wall_law
  use compute_geostrophic
  use initial_conditions
  use running_diagnositcs
  use mod_IO_statistics
  use tke_budget1D
  use mod_IO_main
  implicit none
  private
  public :: LES_main_engine

contains

  real(rprec) function wtime(sync)
    logical, intent(in) :: sync
    integer(8) :: c, r
    if (sync) call les_check(les_sync(), 'les_sync')
    call system_clock(c, r)
    wtime = real(c, rprec)/real(r, rprec)
  end function wtime

  subroutine LES_main_engine()
    integer, parameter :: NaN_check = 20
    logical, parameter :: profiling = .true.
    integer :: jt, jt_total
    real(rprec) :: tt, tadv1, tadv2
    real(rprec) :: time_grad, time_stress, time_rhs, time_press, time_loop, time_tmp, time_sub

    time_grad = 0; time_stress = 0; time_rhs = 0; time_press = 0; time_loop = 0
    write(*,*) '## Starting main engine (device)'
    call read_parameters()
    jt = 0; jt_total = 0; tt = 0._rprec
    if (restart_flag) then
       call read_parameters_restart(jt_total)
       tt = jt_total*dt
    end if
    if (tke_flag) call tke_budget_init()
    call les_check(les_init(cparams), 'les_init')
    ! a roughness map (surface_files): the wall law's z0 per surface point
    if (surface_files) call les_check(les_upload_layer(LES_Z0, lbc_z0), 'set_bc_surface')
    call running_diagnostics_init()
    if (output) call io_init()
    if (stat_flag) call io_stat_init()
    if (.not. restart_flag) then
       call set_initial_conditions()
    else
       call checkpoint_in(jt_total)
    end if
    call les_check(les_set_jt_total(int(jt_total, c_int)), 'les_set_jt_total')
    if (output) call output_loop(0, jt_total)
    call compute_running_diagnistics(jt, jt_total, tt, .false.)
    write(*,*) '## Flow initialized - ready for simulation'

    do jt = 1, nt
       time_tmp = wtime(.true.)
       tt = tt + dt
       jt_total = jt_total + 1

       ! geostrophic wind correction
       if (geostr_rot .and. jt_total >= geostr_start) call compute_geostrophic_correct(jt_total)

       ! friction velocity and wall stress at k = 1
       call obukhov(wall_model)
       call wallstress(wall_model)

       if (profiling) time_sub = wtime(.true.)
       call compute_vel_grad()
       if (profiling) time_grad = time_grad + wtime(.true.) - time_sub

       if (profiling) time_sub = wtime(.true.)
       call sgs_stress(jt, jt_total)
       if (profiling) time_stress = time_stress + wtime(.true.) - time_sub

       if (profiling) time_sub = wtime(.true.)
       call compute_RHS_adv()
       if (profiling) time_rhs = time_rhs + wtime(.true.) - time_sub

       if (jt_total == 1) then
          tadv1 = 1.0_rprec
          tadv2 = 0.0_rprec
       else
          tadv1 = 1.5_rprec
          tadv2 = 1.0_rprec - tadv1
       end if
       call les_check(les_ab2(tadv1, tadv2), 'advection step')

       if (profiling) time_sub = wtime(.true.)
       call press_stag_array(tadv1)
       if (profiling) time_press = time_press + wtime(.true.) - time_sub

       call les_check(les_projection(tadv1), 'projection step')
       if (profiling) time_loop = time_loop + wtime(.true.) - time_tmp

       call compute_running_diagnistics(jt, jt_total, tt, .true.)
       if (mod(jt, NaN_check) == 0) call check_NaN(jt)
       if (stat_flag .and. jt_total > start_stats) call io_stat_averages(jt, jt_total)
       if (tke_flag .and. jt_total > start_tke) call compute_tke_budget(jt)
       if (output .and. mod(jt, step_out) == 0) call output_loop(jt, jt_total)
       if (mod(jt, scheckpoint) == 0) call scheckpoint_out(jt_total)
       if (mod(jt_total, checkpoint) == 0) call checkpoint_out(jt_total)
    end do

    call les_check(les_set_jt_total(int(jt_total, c_int)), 'les_set_jt_total')
    if (trim(sgs_model) == 'dyn_LagScaleDep' .and. cs_count > 0) then
       if ((jt_total/cs_count)*cs_count >= max(dyn_init, 1)) Lagran_init = .true.
    end if
    if (mod(jt_total, checkpoint) /= 0) call checkpoint_out(jt_total)
    call write_parameters_restart(jt_total)
    if (stat_flag) call io_stat_finalize()
    if (tke_flag) call tke_budget_finalize()
    if (output) call io_finalize()

    write(*,'(A,F10.4,A)') ' wall time for time loop: ', time_loop, 's'
    open(99, file='./profiling.txt', status='replace')
    write(99,'(I15.0)') p_on_y, p_on_z, nx, ny, nz, nt
    write(99,'(F15.7)') time_grad, time_stress, time_rhs, time_press, time_loop, 0._rprec
    close(99)
    call les_check(les_finalize(), 'les_finalize')
  end subroutine LES_main_engine

  !! src/LES_engine.f90:449-498: grad_velocity_filt x3, set_bc, ddz_uv/ddz_w
  subroutine compute_vel_grad()
    call les_check(les_compute_vel_grad(), 'compute_vel_grad')
  end subroutine compute_vel_grad

  !! src/LES_engine.f90:509-572: rhs_f <- rhs, convection, div tau, forcing
  subroutine compute_RHS_adv()
    call les_check(les_compute_rhs_adv(), 'compute_RHS_adv')
  end subroutine compute_RHS_adv

  !! src/LES_engine.f90:583-616
  subroutine check_NaN(jt)
    integer, intent(in) :: jt
    integer(c_int) :: found, rc
    rc = les_check_nan(found)
    if (found /= 0) then
       write(*,'(A,I6)') 'ABORT - got a NaN at jt=', jt
       error stop 9
    end if
    call les_check(rc, 'check_NaN')
  end subroutine check_NaN

  !! checkpoint file (src/IO_checkpoint.f90:119-183): levels 1..nz+1 of
  !! u, v, w, rhsx, rhsy, rhsz, cs_opt2 and, once the Lagrangian averages
  !! exist, F_LM, F_MM, F_QN, F_NN; float64, Fortran order.
  subroutine checkpoint_out(jt)
    integer, intent(in) :: jt
    character(128) :: filename
    write(filename,'(A,I9.9)') 'checkpoint_', jt
    write(*,'(A,A)') '# writing checkpoint file: ', trim(filename)
    call write_checkpoint_file(filename, jt)
  end subroutine checkpoint_out

  !! the same content under the fixed name 'scheckpoint', every scheckpoint
  !! steps of this run (src/IO_checkpoint.f90:100-107; src/LES_engine.f90:387-389)
  subroutine scheckpoint_out(jt_total)
    integer, intent(in) :: jt_total
    call write_checkpoint_file('scheckpoint', jt_total)
  end subroutine scheckpoint_out

  subroutine write_checkpoint_file(filename, jt)
    character(*), intent(in) :: filename
    integer, intent(in) :: jt
    real(rprec), allocatable :: a(:,:,:)
    integer(c_int) :: ids(11)
    integer :: i, n
    ids = [u, v, w, rhsx, rhsy, rhsz, cs_opt2, F_LM, F_MM, F_QN, F_NN]
    n = 7
    if (trim(sgs_model) == 'dyn_LagScaleDep' .and. cs_count > 0) then
       if ((jt/cs_count)*cs_count >= max(dyn_init, 1)) Lagran_init = .true.
    end if
    if (Lagran_init) n = 11
    allocate(a(nx, ny, 0:nz+1))
    open(97, file=trim(out_path)//trim(filename), form='unformatted', access='stream', status='replace')
    do i = 1, n
       call get_field(ids(i), a)
       write(97) a(:,:,1:nz+1)
    end do
    close(97)
    deallocate(a)
  end subroutine write_checkpoint_file

  subroutine checkpoint_in(jt)
    integer, intent(in) :: jt
    character(128) :: filename
    real(rprec), allocatable :: a(:,:,:)
    integer(c_int) :: ids(11)
    integer :: i, n, ios
    write(filename,'(A,I9.9)') 'checkpoint_', jt
    write(*,'(A,A)') '# reading checkpoint file: ', trim(filename)
    ids = [u, v, w, rhsx, rhsy, rhsz, cs_opt2, F_LM, F_MM, F_QN, F_NN]
    n = 7
    if (Lagran_init) n = 11
    allocate(a(nx, ny, 0:nz+1))
    open(97, file=trim(filename), form='unformatted', access='stream', status='old', action='read', iostat=ios)
    if (ios /= 0) then
       write(*,*) '[Restart error] cannot open checkpoint file'
       error stop 1
    end if
    do i = 1, n
       a = 0._rprec
       read(97) a(:,:,1:nz+1)
       if (ids(i) == w) a(:,:,0) = BOGUS
       call put_field(ids(i), a)
    end do
    close(97)
    deallocate(a)
  end subroutine checkpoint_in

end module LES_engine
