!------------------------------------------------------------------------------
!! module: initial_conditions
!------------------------------------------------------------------------------
!! set_initial_conditions (src/initial_conditions.f90:46-208) on the host:
!! the interior levels 1..nz of u, v, w for ic_type 'constant', the
!! Coriolis start (coriolis_flag, any other type), 'log_law', 'poiseuille'
!! or 'files', then (except for 'files') the seeded noise of add_noise1
!! (:265-306), handed to les_set_ic, which applies the boundary conditions
!! and the BOGUS plane.  'do_nothing' leaves the device fields as les_init
!! set them.  The noise comes from Knuth's subtractive generator in the
!! form of Numerical Recipes' ran3 (src/ran3.f90), restarted for every level
!! from the seed -80 - k - rnd (rnd = 4, 8, 12 for u, v, w), so the fields
!! equal the reference's bit for bit.
!------------------------------------------------------------------------------
module initial_conditions
  use iso_c_binding
  use les_gpu_abi
  use parameters_IO
  implicit none
  private

  real(rprec), parameter :: KvonK = 0.4_rprec

  ! subtractive generator state: 55 lagged values and two cursors
  integer, parameter :: RG_BIG = 1000000000, RG_SEED = 161803398
  type :: sub_gen
     integer :: lag(55)
     integer :: cur = 0, ahead = 31
  end type sub_gen

  public :: set_initial_conditions, sub_gen, gen_start, gen_next

contains

  ! (Re)start from a seed s: lag(55) = (RG_SEED - |s|) mod RG_BIG, the other
  ! 54 values spread by the index map 21 i mod 55 as successive differences,
  ! then four passes of lag(i) -= lag(1 + (i + 30) mod 55) modulo RG_BIG.
  subroutine gen_start(g, s)
    type(sub_gen), intent(out) :: g
    integer, intent(in) :: s
    integer :: prev, cur, i, idx, pass
    cur = mod(RG_SEED - abs(s), RG_BIG)
    g%lag(55) = cur
    prev = 1
    do i = 1, 54
       idx = mod(21*i, 55)
       g%lag(idx) = prev
       prev = cur - prev
       if (prev < 0) prev = prev + RG_BIG
       cur = g%lag(idx)
    end do
    do pass = 1, 4
       do i = 1, 55
          g%lag(i) = g%lag(i) - g%lag(1 + mod(i + 30, 55))
          if (g%lag(i) < 0) g%lag(i) = g%lag(i) + RG_BIG
       end do
    end do
    g%cur = 0
    g%ahead = 31
  end subroutine gen_start

  ! next value in [0, 1): lag(cur) -= lag(ahead) modulo RG_BIG, scaled
  function gen_next(g) result(r)
    type(sub_gen), intent(inout) :: g
    real(rprec) :: r
    integer :: d
    g%cur = g%cur + 1
    if (g%cur == 56) g%cur = 1
    g%ahead = g%ahead + 1
    if (g%ahead == 56) g%ahead = 1
    d = g%lag(g%cur) - g%lag(g%ahead)
    if (d < 0) d = d + RG_BIG
    g%lag(g%cur) = d
    r = d*(1.d0/RG_BIG)
  end function gen_next

  ! add_noise1 on the interior levels of x(nx, ny, nz): below zturb every
  ! point gets 3 vnf u*^2 (2 r - 1) (lz - 0.9 z) / lz, the generator restarted
  ! per level
  subroutine add_noise(x, rnd, zturb)
    real(rprec), intent(inout) :: x(:,:,:)
    integer, intent(in) :: rnd
    real(rprec), intent(in) :: zturb
    type(sub_gen) :: g
    integer :: i, j, k
    real(rprec) :: z, noise, vf
    if (.not. (u_star > 0.0_rprec)) return
    vf = 3.0_rprec*vnf*u_star**2
    do k = 1, nz
       z = (k - 0.5_rprec)*dz
       if (z <= zturb) then
          call gen_start(g, -80 - k - rnd)
          do j = 1, ny
             do i = 1, nx
                noise = gen_next(g)*2.0_rprec - 1.0_rprec
                x(i, j, k) = x(i, j, k) + vf*(noise*(lz - z*0.9_rprec)/lz)
             end do
          end do
       end if
    end do
  end subroutine add_noise

  subroutine set_initial_conditions()
    real(rprec), allocatable :: u(:,:,:), v(:,:,:), w(:,:,:)
    real(rprec) :: norm, dir_x, dir_y, z
    integer :: k

    if (trim(ic_type) == 'do_nothing' .and. .not. coriolis_flag) return
    allocate(u(nx, ny, nz), v(nx, ny, nz), w(nx, ny, nz))
    u = 0._rprec; v = 0._rprec; w = 0._rprec
    norm = (pfx**2 + pfy**2)**0.5_rprec
    if (norm > 0.0_rprec) then
       dir_x = pfx/norm
       dir_y = pfy/norm
    else
       dir_x = 1.0_rprec
       dir_y = 0.0_rprec
    end if

    if (trim(ic_type) == 'constant') then
       u = u0; v = v0; w = w0
    else if (coriolis_flag) then
       do k = 1, nz
          z = (k - 0.5_rprec)*dz
          u(:,:,k) = ug
          if (z <= 0.5_rprec) v(:,:,k) = vg
       end do
    else if (trim(ic_type) == 'log_law') then
       do k = 1, nz
          z = (k - 0.5_rprec)*dz
          if (z >= 0._rprec) then
             u(:,:,k) = dir_x*((1._rprec/KvonK)*log(z/lbc_z0))*u_star
             v(:,:,k) = dir_y*((1._rprec/KvonK)*log(z/lbc_z0))*u_star
          end if
       end do
    else if (trim(ic_type) == 'poiseuille') then
       do k = 1, nz
          z = (k - 0.5_rprec)*dz
          u(:,:,k) = Re*z*(1.0_rprec - 0.5_rprec*z)
       end do
    else if (trim(ic_type) == 'files') then
       call read_raw('vel_u.dat', u)
       call read_raw('vel_v.dat', v)
       call read_raw('vel_w.dat', w)
    else
       write(*,*) 'ABORT - Invalid initial flow condition type ', trim(ic_type)
       error stop 1
    end if

    if (trim(ic_type) /= 'files') then
       call add_noise(u, 4, z_turb*lz)
       call add_noise(v, 8, z_turb*lz)
       call add_noise(w, 12, z_turb*lz)
    end if
    call les_check(les_set_ic(u, v, w), 'set_initial_conditions')
    deallocate(u, v, w)
  end subroutine set_initial_conditions

  subroutine read_raw(name, a)
    character(*), intent(in) :: name
    real(rprec), intent(out) :: a(:,:,:)
    integer :: ios
    open(98, file=name, form='unformatted', access='stream', status='old', action='read', iostat=ios)
    if (ios /= 0) then
       write(*,*) 'ABORT - cannot open ', trim(name)
       error stop 1
    end if
    read(98) a
    close(98)
  end subroutine read_raw

end module initial_conditions
