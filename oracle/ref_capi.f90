! C-ABI wrappers around the upstream OpenSC Fortran library, used ONLY as a
! test oracle and CPU baseline (never shipped, never on the product path).
!
! The upstream modules are compiled unchanged from their own sources by
! oracle/Makefile into oracle/_ref/libopensc_ref.so together with this file.
! Every wrapper passes ALL optional arguments of the upstream routine, because
! the upstream code evaluates `present(x) .and. .not. x` without
! short-circuiting (code/open_spacecharge_mod.f90:67,70,92,95,98,118,121).
!
! Arrays are Fortran column-major, x fastest, lower bound 1.
module osc_ref_capi
use iso_c_binding
use open_spacecharge_mod
use open_spacecharge_core_mod
use fft_interface_mod, only: ccfft3d
use test_mod, only: get_mesh_quantities, depose_rho_scalar
implicit none
contains

subroutine set_mesh(mesh, nlo, nhi, npad, mn, mx, delta, gamma)
  type(mesh3d_struct), intent(inout) :: mesh
  integer(c_int), intent(in) :: nlo(3), nhi(3), npad(3)
  real(c_double), intent(in) :: mn(3), mx(3), delta(3), gamma
  mesh%nlo = nlo; mesh%nhi = nhi; mesh%npad = npad
  mesh%min = mn; mesh%max = mx; mesh%delta = delta; mesh%gamma = gamma
end subroutine

subroutine alloc_mesh(mesh)
  type(mesh3d_struct), intent(inout) :: mesh
  integer :: lo(3), hi(3)
  lo = mesh%nlo; hi = mesh%nhi
  if (.not. allocated(mesh%rho)) allocate(mesh%rho(lo(1):hi(1),lo(2):hi(2),lo(3):hi(3)))
  if (.not. allocated(mesh%phi)) allocate(mesh%phi(lo(1):hi(1),lo(2):hi(2),lo(3):hi(3)))
  if (.not. allocated(mesh%efield)) allocate(mesh%efield(lo(1):hi(1),lo(2):hi(2),lo(3):hi(3),3))
  if (.not. allocated(mesh%bfield)) allocate(mesh%bfield(lo(1):hi(1),lo(2):hi(2),lo(3):hi(3),3))
  mesh%phi = 0; mesh%efield = 0; mesh%bfield = 0
end subroutine

! deposit_particles (code/open_spacecharge_mod.f90:156-288).
! qmode: 1 = per-particle qa, 2 = total_charge, 0 = neither (upstream prints an error).
subroutine ref_deposit_particles(n, x, y, z, q, qmode, total_charge, resize, &
    nlo, nhi, npad, mn, mx, delta, gamma, charge, rho) bind(C, name='ref_deposit_particles')
  integer(c_int64_t), value :: n
  real(c_double), intent(in) :: x(n), y(n), z(n), q(n)
  integer(c_int), value :: qmode, resize
  real(c_double), value :: total_charge, gamma
  integer(c_int), intent(in) :: nlo(3), nhi(3), npad(3)
  real(c_double), intent(inout) :: mn(3), mx(3), delta(3)
  real(c_double), intent(out) :: charge
  real(c_double), intent(inout) :: rho(*)
  type(mesh3d_struct) :: mesh
  logical :: rs
  integer :: ntot
  call set_mesh(mesh, nlo, nhi, npad, mn, mx, delta, gamma)
  mesh%charge = 0
  rs = resize /= 0
  if (qmode == 1) then
    call deposit_particles(x, y, z, mesh, qa=q, resize_mesh=rs)
  else if (qmode == 2) then
    call deposit_particles(x, y, z, mesh, total_charge=total_charge, resize_mesh=rs)
  else
    call deposit_particles(x, y, z, mesh, resize_mesh=rs)
  end if
  mn = mesh%min; mx = mesh%max; delta = mesh%delta; charge = mesh%charge
  ntot = size(mesh%rho)
  rho(1:ntot) = reshape(mesh%rho, [ntot])
end subroutine

! get_mesh_quantities (code/test_mod.f90:288-358)
subroutine ref_get_mesh_quantities(n, x, y, z, nlo, nhi, isetzmin, delta, umin) &
    bind(C, name='ref_get_mesh_quantities')
  integer(c_int), value :: n, isetzmin
  real(c_double), intent(in) :: x(n), y(n), z(n)
  integer(c_int), intent(in) :: nlo(3), nhi(3)
  real(c_double), intent(out) :: delta(3), umin(3)
  real(c_double), allocatable :: lost(:), xa(:), ya(:), za(:)
  integer :: lo(3), hi(3)
  allocate(lost(n)); lost = 0
  allocate(xa(n), ya(n), za(n)); xa = x; ya = y; za = z
  lo = nlo; hi = nhi
  call get_mesh_quantities(xa, ya, za, lost, delta, umin, lo, hi, lo, hi, 7, n, n, isetzmin)
end subroutine

! depose_rho_scalar (code/test_mod.f90:410-474)
subroutine ref_depose_rho_scalar(n, x, y, z, nlo, nhi, chrgpermacro, delta, umin, rho, ifail) &
    bind(C, name='ref_depose_rho_scalar')
  integer(c_int), value :: n
  real(c_double), intent(in) :: x(n), y(n), z(n)
  integer(c_int), intent(in) :: nlo(3), nhi(3)
  real(c_double), value :: chrgpermacro
  real(c_double), intent(in) :: delta(3), umin(3)
  real(c_double), intent(out) :: rho(nlo(1):nhi(1), nlo(2):nhi(2), nlo(3):nhi(3))
  integer(c_int), intent(out) :: ifail
  real(c_double), allocatable :: lost(:)
  integer :: lo(3), nf
  allocate(lost(n)); lost = 0
  lo = nlo
  call depose_rho_scalar(x, y, z, lost, rho, chrgpermacro, lo, delta, umin, n, lo, nf)
  ifail = nf
end subroutine

! space_charge_3d (code/open_spacecharge_mod.f90:431-489)
subroutine ref_space_charge_3d(nlo, nhi, npad, mn, mx, delta, gamma, rho, offset, &
    at_cathode, calc_bfield, efield, bfield) bind(C, name='ref_space_charge_3d')
  integer(c_int), intent(in) :: nlo(3), nhi(3), npad(3)
  real(c_double), intent(in) :: mn(3), mx(3), delta(3)
  real(c_double), value :: gamma
  real(c_double), intent(in) :: rho(*), offset(3)
  integer(c_int), value :: at_cathode, calc_bfield
  real(c_double), intent(out) :: efield(*), bfield(*)
  type(mesh3d_struct) :: mesh
  integer :: nt
  call set_mesh(mesh, nlo, nhi, npad, mn, mx, delta, gamma)
  call alloc_mesh(mesh)
  nt = size(mesh%rho)
  mesh%rho = reshape(rho(1:nt), shape(mesh%rho))
  call space_charge_3d(mesh, offset=offset, at_cathode=(at_cathode /= 0), calc_bfield=(calc_bfield /= 0))
  efield(1:3*nt) = reshape(mesh%efield, [3*nt])
  bfield(1:3*nt) = reshape(mesh%bfield, [3*nt])
end subroutine

! osc_freespace_solver2 (code/open_spacecharge_mod.f90:535-597), efield and phi both present
subroutine ref_freespace_solver2(nx, ny, nz, rho, gamma, delta, offset, efield, phi) &
    bind(C, name='ref_freespace_solver2')
  integer(c_int), value :: nx, ny, nz
  real(c_double), intent(in) :: rho(nx, ny, nz), delta(3), offset(3)
  real(c_double), value :: gamma
  real(c_double), intent(out) :: efield(nx, ny, nz, 3), phi(nx, ny, nz)
  call osc_freespace_solver2(rho, gamma, delta, efield=efield, phi=phi, offset=offset)
end subroutine

! osc_get_cgrn_freespace (code/open_spacecharge_mod.f90:633-707)
subroutine ref_get_cgrn_freespace(isize, jsize, ksize, cgrn, delta, gamma, icomp, offset) &
    bind(C, name='ref_get_cgrn_freespace')
  integer(c_int), value :: isize, jsize, ksize, icomp
  complex(c_double_complex), intent(out) :: cgrn(isize, jsize, ksize)
  real(c_double), intent(in) :: delta(3), offset(3)
  real(c_double), value :: gamma
  call osc_get_cgrn_freespace(cgrn, delta, gamma, icomp, offset=offset)
end subroutine

! ccfft3d (code/fft_interface_mod.f90:96-143), in place
subroutine ref_ccfft3d(n1, n2, n3, a, idir) bind(C, name='ref_ccfft3d')
  integer(c_int), value :: n1, n2, n3
  complex(c_double_complex), intent(inout) :: a(n1, n2, n3)
  integer(c_int), intent(in) :: idir(3)
  integer :: d(3)
  d = idir
  call ccfft3d(a, a, d, n1, n2, n3, 0)
end subroutine

! osc_getgrnfree (code/open_spacecharge_core_mod.f90:454-497): FFT'd legacy Green function
subroutine ref_getgrnfree(nlo, nhi, npad, cgrn, delta, gamma, icomp, igfflag) &
    bind(C, name='ref_getgrnfree')
  integer(c_int), intent(in) :: nlo(3), nhi(3), npad(3)
  complex(c_double_complex), intent(out) :: cgrn(*)
  real(c_double), intent(in) :: delta(3)
  real(c_double), value :: gamma
  integer(c_int), value :: icomp, igfflag
  call grn_helper(cgrn, nlo - nhi, nhi - nlo + npad)
contains
  subroutine grn_helper(g, lo, hi)
    integer :: lo(3), hi(3)
    complex(c_double_complex) :: g(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3))
    real(c_double) :: dl(3), gm
    integer :: pd(3), ic, ig
    dl = delta; gm = gamma; pd = npad; ic = icomp; ig = igfflag
    call osc_getgrnfree(g, dl, gm, lo(1), lo(2), lo(3), pd, ic, ig)
  end subroutine
end subroutine

! osc_getgrnimageshift (code/open_spacecharge_core_mod.f90:1135-1178)
subroutine ref_getgrnimageshift(nlo, nhi, npad, cgrn, delta, gamma, zshft, icomp, igfflag) &
    bind(C, name='ref_getgrnimageshift')
  integer(c_int), intent(in) :: nlo(3), nhi(3), npad(3)
  complex(c_double_complex), intent(out) :: cgrn(*)
  real(c_double), intent(in) :: delta(3)
  real(c_double), value :: gamma, zshft
  integer(c_int), value :: icomp, igfflag
  call grn_helper(cgrn, nlo - nhi, nhi - nlo + npad)
contains
  subroutine grn_helper(g, lo, hi)
    integer :: lo(3), hi(3)
    complex(c_double_complex) :: g(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3))
    real(c_double) :: dl(3), gm, zs
    integer :: pd(3), ic, ig
    dl = delta; gm = gamma; pd = npad; ic = icomp; ig = igfflag; zs = zshft
    call osc_getgrnimageshift(g, dl, gm, zs, lo(1), lo(2), lo(3), pd, ic, ig)
  end subroutine
end subroutine

! osc_getgrnimageconvcorr (code/open_spacecharge_core_mod.f90:1185-1230)
subroutine ref_getgrnimageconvcorr(nlo, nhi, npad, cgrn, umin, delta, gamma, icomp, igfflag) &
    bind(C, name='ref_getgrnimageconvcorr')
  integer(c_int), intent(in) :: nlo(3), nhi(3), npad(3)
  complex(c_double_complex), intent(out) :: cgrn(*)
  real(c_double), intent(in) :: umin(3), delta(3)
  real(c_double), value :: gamma
  integer(c_int), value :: icomp, igfflag
  integer :: lo(3), hi(3)
  lo(1:2) = nlo(1:2) - nhi(1:2); hi(1:2) = nhi(1:2) - nlo(1:2) + npad(1:2)
  lo(3) = 2*nlo(3); hi(3) = 2*nhi(3) + npad(3)
  call grn_helper(cgrn, lo, hi)
contains
  subroutine grn_helper(g, lo, hi)
    integer :: lo(3), hi(3)
    complex(c_double_complex) :: g(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3))
    real(c_double) :: dl(3), gm, um(3)
    integer :: pd(3), ic, ig, nl(3)
    dl = delta; gm = gamma; pd = npad; ic = icomp; ig = igfflag; um = umin; nl = nlo
    call osc_getgrnimageconvcorr(g, um, dl, gm, lo(1), lo(2), lo(3), nl, pd, ic, ig)
  end subroutine
end subroutine

! space_charge_freespace (code/open_spacecharge_mod.f90:60-75)
subroutine ref_space_charge_freespace(nlo, nhi, npad, mn, mx, delta, gamma, rho, &
    direct, igf, phi, efield, bfield) bind(C, name='ref_space_charge_freespace')
  integer(c_int), intent(in) :: nlo(3), nhi(3), npad(3)
  real(c_double), intent(in) :: mn(3), mx(3), delta(3)
  real(c_double), value :: gamma
  real(c_double), intent(in) :: rho(*)
  integer(c_int), value :: direct, igf
  real(c_double), intent(inout) :: phi(*), efield(*), bfield(*)
  type(mesh3d_struct) :: mesh
  integer :: nt
  call set_mesh(mesh, nlo, nhi, npad, mn, mx, delta, gamma)
  call alloc_mesh(mesh)
  nt = size(mesh%rho)
  mesh%rho = reshape(rho(1:nt), shape(mesh%rho))
  mesh%phi = reshape(phi(1:nt), shape(mesh%phi))
  mesh%efield = reshape(efield(1:3*nt), shape(mesh%efield))
  mesh%bfield = reshape(bfield(1:3*nt), shape(mesh%bfield))
  call space_charge_freespace(mesh, direct_field_calc=(direct /= 0), integrated_green_function=(igf /= 0))
  phi(1:nt) = reshape(mesh%phi, [nt])
  efield(1:3*nt) = reshape(mesh%efield, [3*nt])
  bfield(1:3*nt) = reshape(mesh%bfield, [3*nt])
end subroutine

! space_charge_cathodeimages (code/open_spacecharge_mod.f90:83-103); method 2 mutates rho
subroutine ref_space_charge_cathodeimages(nlo, nhi, npad, mn, mx, delta, gamma, rho, &
    direct, igf, method, phi, efield, bfield) bind(C, name='ref_space_charge_cathodeimages')
  integer(c_int), intent(in) :: nlo(3), nhi(3), npad(3)
  real(c_double), intent(in) :: mn(3), mx(3), delta(3)
  real(c_double), value :: gamma
  real(c_double), intent(inout) :: rho(*)
  integer(c_int), value :: direct, igf, method
  real(c_double), intent(inout) :: phi(*), efield(*), bfield(*)
  type(mesh3d_struct) :: mesh
  integer :: nt, im
  call set_mesh(mesh, nlo, nhi, npad, mn, mx, delta, gamma)
  call alloc_mesh(mesh)
  nt = size(mesh%rho)
  mesh%rho = reshape(rho(1:nt), shape(mesh%rho))
  mesh%phi = reshape(phi(1:nt), shape(mesh%phi))
  mesh%efield = reshape(efield(1:3*nt), shape(mesh%efield))
  mesh%bfield = reshape(bfield(1:3*nt), shape(mesh%bfield))
  im = method
  call space_charge_cathodeimages(mesh, direct_field_calc=(direct /= 0), &
       integrated_green_function=(igf /= 0), image_method=im)
  rho(1:nt) = reshape(mesh%rho, [nt])
  phi(1:nt) = reshape(mesh%phi, [nt])
  efield(1:3*nt) = reshape(mesh%efield, [3*nt])
  bfield(1:3*nt) = reshape(mesh%bfield, [3*nt])
end subroutine

! osc_alloc_rectpipe_arrays + osc_getgrnpipe (code/open_spacecharge_core_mod.f90:876-898, 632-715)
! and optionally osc_write_rectpipe_grn (:851-870) into ./rectpipegrn.dat
subroutine ref_pipe_green(nlo, nhi, npad, gamma, apipe, bpipe, delta, umin, umax, write_file) &
    bind(C, name='ref_pipe_green')
  integer(c_int), intent(in) :: nlo(3), nhi(3), npad(3)
  real(c_double), value :: gamma, apipe, bpipe
  real(c_double), intent(in) :: delta(3), umin(3), umax(3)
  integer(c_int), value :: write_file
  integer :: lo(3), hi(3), pd(3)
  real(c_double) :: g, a, b, dl(3), um(3), ux(3)
  lo = nlo; hi = nhi; pd = npad; g = gamma; a = apipe; b = bpipe; dl = delta; um = umin; ux = umax
  call osc_alloc_rectpipe_arrays(lo, hi, pd)
  call osc_getgrnpipe(g, a, b, dl, um, pd)
  if (write_file /= 0) call osc_write_rectpipe_grn(a, b, dl, um, ux, lo, hi, g)
end subroutine

! osc_alloc_rectpipe_arrays + osc_read_rectpipe_grn (reads ./rectpipegrn.dat)
subroutine ref_pipe_read(nlo, nhi, npad) bind(C, name='ref_pipe_read')
  integer(c_int), intent(in) :: nlo(3), nhi(3), npad(3)
  integer :: lo(3), hi(3), pd(3)
  lo = nlo; hi = nhi; pd = npad
  call osc_alloc_rectpipe_arrays(lo, hi, pd)
  call osc_read_rectpipe_grn
end subroutine

! space_charge_rectpipe (code/open_spacecharge_mod.f90:110-127); Green from ref_pipe_green
subroutine ref_space_charge_rectpipe(nlo, nhi, npad, mn, mx, delta, gamma, rho, apipe, bpipe, &
    direct, igf, phi, efield, bfield) bind(C, name='ref_space_charge_rectpipe')
  integer(c_int), intent(in) :: nlo(3), nhi(3), npad(3)
  real(c_double), intent(in) :: mn(3), mx(3), delta(3)
  real(c_double), value :: gamma, apipe, bpipe
  real(c_double), intent(in) :: rho(*)
  integer(c_int), value :: direct, igf
  real(c_double), intent(inout) :: phi(*), efield(*), bfield(*)
  type(mesh3d_struct) :: mesh
  integer :: nt
  real(c_double) :: a, b
  call set_mesh(mesh, nlo, nhi, npad, mn, mx, delta, gamma)
  call alloc_mesh(mesh)
  nt = size(mesh%rho)
  mesh%rho = reshape(rho(1:nt), shape(mesh%rho))
  mesh%phi = reshape(phi(1:nt), shape(mesh%phi))
  mesh%efield = reshape(efield(1:3*nt), shape(mesh%efield))
  mesh%bfield = reshape(bfield(1:3*nt), shape(mesh%bfield))
  a = apipe; b = bpipe
  call space_charge_rectpipe(mesh, a, b, direct_field_calc=(direct /= 0), &
       integrated_green_function=(igf /= 0))
  phi(1:nt) = reshape(mesh%phi, [nt])
  efield(1:3*nt) = reshape(mesh%efield, [3*nt])
  bfield(1:3*nt) = reshape(mesh%bfield, [3*nt])
end subroutine

! interpolate_field (code/open_spacecharge_mod.f90:311-390) for n points; E, B are (3, n)
subroutine ref_interpolate_field(nlo, nhi, mn, mx, delta, efield, bfield, n, x, y, z, e, b) &
    bind(C, name='ref_interpolate_field')
  integer(c_int), intent(in) :: nlo(3), nhi(3)
  real(c_double), intent(in) :: mn(3), mx(3), delta(3)
  real(c_double), intent(in) :: efield(*), bfield(*)
  integer(c_int64_t), value :: n
  real(c_double), intent(in) :: x(n), y(n), z(n)
  real(c_double), intent(out) :: e(3, n), b(3, n)
  type(mesh3d_struct) :: mesh
  integer(c_int64_t) :: i
  integer :: nt, pd(3)
  pd = 1
  call set_mesh(mesh, nlo, nhi, pd, mn, mx, delta, 1.0_c_double)
  call alloc_mesh(mesh)
  nt = size(mesh%rho)
  mesh%efield = reshape(efield(1:3*nt), shape(mesh%efield))
  mesh%bfield = reshape(bfield(1:3*nt), shape(mesh%bfield))
  do i = 1, n
    call interpolate_field(x(i), y(i), z(i), mesh, E=e(:, i), B=b(:, i))
  end do
end subroutine

end module osc_ref_capi
