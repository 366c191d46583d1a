! Test driver of the Fortran drop-in module open_spacecharge_mod (GPU).
! It makes the calls a Fortran caller of the upstream module makes
! (code/test_opensc.f90:139-220 pattern): deposit_particles ->
! space_charge_3d with B -> interpolate_field (batch and per point), the
! legacy free-space solver, and the pipe Green functions written to and read
! back from rectpipegrn.dat followed by space_charge_rectpipe.
!
! usage: drive_shim IN OUT NX NY NZ
!   IN : stream file, int64 n, then x(n), y(n), z(n) float64
!   OUT: stream file of float64 blocks in this order:
!        min(3) max(3) delta(3) charge, rho, efield, bfield, E(3,n), B(3,n),
!        E of the first 4 points by per-point calls (3 x 4),
!        E(3,m), B(3,m) of the first m = min(n, 10000) points by per-point
!        calls, the seconds those m calls took (1), E and B (3 + 3) of one
!        point past the upper x edge (clamped; upstream prints messages),
!        legacy direct efield, pipe phi, pipe efield
program drive_shim
use, intrinsic :: iso_fortran_env
use open_spacecharge_mod
implicit none
integer, parameter :: dp = REAL64
type(mesh3d_struct) :: mesh, pmesh
character(len=512) :: fin, fout, arg
integer :: nx, ny, nz, u, i
integer(int64) :: n, ncl, m1, t0, t1, rate
real(dp), allocatable :: x(:), y(:), z(:), E(:,:), B(:,:), Ep(:,:), Bp(:,:)
real(dp) :: e1(3), b1(3), apipe, bpipe, delta(3), umin(3), umax(3), xo

call get_command_argument(1, fin)
call get_command_argument(2, fout)
call get_command_argument(3, arg); read(arg, *) nx
call get_command_argument(4, arg); read(arg, *) ny
call get_command_argument(5, arg); read(arg, *) nz

open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
read(u) n
allocate(x(n), y(n), z(n), E(3, n), B(3, n))
read(u) x, y, z
close(u)

mesh%nhi = [nx, ny, nz]
mesh%gamma = 5.0e6_dp / 0.51099891e6_dp

! error path: neither qa nor total_charge prints and returns (mesh resized)
call deposit_particles(x, y, z, mesh)

call deposit_particles(x, y, z, mesh, total_charge=1.0e-9_dp)
call space_charge_3d(mesh, calc_bfield=.true.)
call interpolate_field_batch(x, y, z, mesh, E, B, ncl)

open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
write(u) mesh%min, mesh%max, mesh%delta, mesh%charge
write(u) mesh%rho, mesh%efield, mesh%bfield, E, B
do i = 1, 4
  call interpolate_field(x(i), y(i), z(i), mesh, E=e1)
  write(u) e1
enddo
! many per-point calls, as upstream callers make them in their particle loops
m1 = min(n, 10000_int64)
allocate(Ep(3, m1), Bp(3, m1))
call system_clock(t0, rate)
do i = 1, int(m1)
  call interpolate_field(x(i), y(i), z(i), mesh, E=Ep(:, i), B=Bp(:, i))
enddo
call system_clock(t1)
write(u) Ep, Bp, real(t1 - t0, dp) / real(rate, dp)
! one point past the upper x edge: clamped to the edge cell
xo = mesh%max(1) + 2 * mesh%delta(1)
call interpolate_field(xo, y(1), z(1), mesh, E=e1, B=b1)
write(u) e1, b1

! legacy free-space solver, direct field
call space_charge_freespace(mesh, direct_field_calc=.true., integrated_green_function=.true.)
write(u) mesh%efield

! rectangular pipe, harness mesh rule (code/test_opensc.f90:139-140)
apipe = 10.0e-3_dp
bpipe = 6.0e-3_dp
pmesh%nhi = [nx, ny, nz]
pmesh%gamma = mesh%gamma
umin = [-apipe / 2, -bpipe / 2, -6 * 3.0e-3_dp]
umax = -umin
delta = (umax - umin) / (pmesh%nhi - pmesh%nlo)
pmesh%min = umin
pmesh%max = umax
pmesh%delta = delta
call deposit_particles(x, y, z, pmesh, total_charge=1.0e-9_dp, resize_mesh=.false.)
call osc_alloc_rectpipe_arrays(pmesh%nlo, pmesh%nhi, pmesh%npad)
call osc_getgrnpipe(pmesh%gamma, apipe, bpipe, delta, umin, pmesh%npad)
call osc_write_rectpipe_grn(apipe, bpipe, delta, umin, umax, pmesh%nlo, pmesh%nhi, pmesh%gamma)
call osc_read_rectpipe_grn
call space_charge_rectpipe(pmesh, apipe, bpipe, direct_field_calc=.false.)
write(u) pmesh%phi, pmesh%efield
close(u)
call osc_finalize()
print *, 'drive_shim ok', n, ncl
end program drive_shim
