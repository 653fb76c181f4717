!------------------------------------------------------------------------------
!! module: tke_budget1D
!------------------------------------------------------------------------------
!! TKE budget profiles (src/tke_budget1D.f90:64-468) over the device state:
!! the 61 per-level sums of plane averages live on the device (les_tke_*)
!! and come to the host every tke_period steps, when each quantity's
!! record ii of out_path/ta1_profiles/<name>.out (direct access, one record
!! of nz float64 values, the reference's recl=2*nz in 4-byte units) gets the
!! sum / tke_period, as in the reference; then the sums start again and ii
!! advances.  As there, a record holds the sums of the steps since the last
!! one (those past start_tke) divided by tke_period.
!------------------------------------------------------------------------------
module tke_budget1D
  use iso_c_binding
  use les_gpu_abi
  use parameters_IO
  implicit none
  private

  integer, save :: ii = 1
  logical, save :: device_ready = .false.
  character(512), save :: oPath

  ! the reference's avg_var names, in the order of the device's sums
  character(len=8), parameter :: names(LES_NTKE) = [character(len=8) :: &
       'u', 'v', 'w', 'dudx', 'dudy', 'dudz', 'dvdx', 'dvdy', 'dvdz', 'dwdx', 'dwdy', 'dwdz', &
       'uu', 'vv', 'ww', 'uv', 'uw', 'vw', 'txx', 'tyy', 'tzz', 'txy', 'txz', 'tyz', 'cs', &
       'uuw', 'vvw', 'www', 'uww', 'wtzz', 'utxz', 'vtyz', 'dxx', 'dyy', 'dzz', 'dxy', 'dxz', 'dyz', &
       'dudx2', 'dudy2', 'dudz2', 'dvdx2', 'dvdy2', 'dvdz2', 'dwdx2', 'dwdy2', 'dwdz2', &
       'dudxdwdx', 'dudydwdy', 'dudzdwdz', 'p', 'dpdx', 'dpdy', 'dpdz', 'pw', 'pu', &
       'pdudx', 'pdvdy', 'pdwdz', 'pdwdx', 'pdudz']

  public :: tke_budget_init, tke_budget_finalize, compute_tke_budget

contains

  ! Creates the output folder (before the library touches the GPU: the
  ! folder is made by a child process); the device sums are set up by the
  ! first compute_tke_budget.
  subroutine tke_budget_init()
    integer :: stat
    ii = 1
    device_ready = .false.
    oPath = trim(out_path)//'ta1_profiles/'
    write(*,*) 'Creating folder for tke budget output...'
    call execute_command_line('mkdir -p '//trim(oPath), exitstat=stat)
    if (stat == 0) then
       write(*,*) 'Created output folder!'
    else
       write(*,*) 'Output already exists!'
       write(*,*) 'Result=', stat
    end if
  end subroutine tke_budget_init

  subroutine tke_budget_finalize()
  end subroutine tke_budget_finalize

  subroutine compute_tke_budget(jt)
    integer, intent(in) :: jt
    real(rprec), allocatable :: sums(:, :)
    real(rprec) :: n
    integer :: q, rl
    character(600) :: fname

    if (.not. device_ready) then
       call les_check(les_tke_init(), 'tke_budget_init')
       device_ready = .true.
    end if
    call les_check(les_tke_accumulate(), 'compute_tke_budget')
    if (mod(jt, tke_period) == 0) then
       allocate(sums(nz, LES_NTKE))
       call les_check(les_tke_download(sums), 'compute_tke_budget')
       n = real(tke_period, rprec)
       inquire(iolength=rl) sums(:, 1)
       do q = 1, LES_NTKE
          fname = trim(oPath)//trim(names(q))//'.out'
          open(11, file=trim(fname), access='direct', form='unformatted', recl=rl)
          write(11, rec=ii) sums(:, q)/n
          close(11)
       end do
       deallocate(sums)
       call les_check(les_tke_reset(), 'compute_tke_budget')
       ii = ii + 1
    end if
  end subroutine compute_tke_budget

end module tke_budget1D
