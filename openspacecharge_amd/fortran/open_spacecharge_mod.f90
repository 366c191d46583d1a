! Drop-in replacement of the upstream module open_spacecharge_mod
! (code/open_spacecharge_mod.f90) that computes on an AMD Instinct GPU
! (gfx950) through the C-ABI of include/opensc_hip.h.
!
! Same module name, same mesh3d_struct (components, defaults), same routine
! names, argument lists and optional arguments as upstream, plus the
! open_spacecharge_core_mod routines that callers use directly for the pipe
! Green functions (code/test_opensc.f90:206-213).  A caller rebuilds against
! this module and links libopensc_f.so + libopensc_hip.so instead of the
! upstream objects.
!
! Error behaviour follows upstream: the same messages; `stop` where upstream
! stops (FFT length, missing pipe Green functions) and print + return where
! it returns.  Absent optional logicals are handled (upstream's
! `present(x) .and. .not. x` is not short-circuited by every compiler).
! With no usable GPU the first call stops with a message: there is no CPU
! path.
!
! Extensions (no upstream counterpart): osc_set_device, osc_set_precision,
! osc_finalize, interpolate_field_batch (many points per call) and
! space_charge_step (deposit -> space_charge_3d -> gather in one call).
module open_spacecharge_mod

use, intrinsic :: iso_fortran_env
use, intrinsic :: iso_c_binding
use opensc_hip

implicit none

integer, parameter, private :: dp = REAL64

type mesh3d_struct
  integer :: nlo(3) = [ 1,  1,  1]       ! lowest grid index in x, y, z
  integer :: nhi(3) = [64, 64, 64]       ! highest grid index in x, y, z
  integer :: npad(3) = [ 1,  1,  1]      ! padding of the cyclic convolution (legacy solvers)
  real(dp) :: min(3)                     ! minimum in each dimension
  real(dp) :: max(3)                     ! maximum in each dimension
  real(dp) :: delta(3)                   ! grid spacing
  real(dp) :: gamma                      ! relativistic gamma
  real(dp) :: charge                     ! total charge on the mesh
  real(dp), allocatable, dimension(:,:,:) :: rho
  real(dp), allocatable, dimension(:,:,:) :: phi
  real(dp), allocatable, dimension(:,:,:,:) :: efield
  real(dp), allocatable, dimension(:,:,:,:) :: bfield
end type

type(c_ptr), save, private :: ctx_ = c_null_ptr
integer, save, private :: device_ = 0
integer, save, private :: precision_ = 0

private :: ctx, to_c, from_c, last_error, fail, ensure_arrays, flag

contains

!------------------------------------------------------------------------
! context (one per process, created on first use)
!------------------------------------------------------------------------
function ctx() result(h)
  type(c_ptr) :: h
  integer(c_int) :: rc
  if (.not. c_associated(ctx_)) then
    rc = c_osc_ctx_create(int(device_, c_int), ctx_)
    if (rc /= OSC_OK) then
      write(*, *) 'open_spacecharge_mod: no usable GPU (osc_ctx_create returned', rc, ')'
      stop 1
    endif
    if (precision_ /= 0) rc = c_osc_ctx_set_precision(ctx_, int(precision_, c_int))
  endif
  h = ctx_
end function

! Selects the GPU used by the next context (call before the first solve).
subroutine osc_set_device(device)
  integer, intent(in) :: device
  device_ = device
end subroutine

! fp32 = .true.: FP32 convolution against the FP64-built Green spectrum.
subroutine osc_set_precision(fp32)
  logical, intent(in) :: fp32
  integer(c_int) :: rc
  precision_ = merge(1, 0, fp32)
  if (c_associated(ctx_)) then
    rc = c_osc_ctx_set_precision(ctx_, int(precision_, c_int))
  endif
end subroutine

! Releases the GPU context (device buffers, cached Green spectra).
subroutine osc_finalize()
  integer(c_int) :: rc
  if (c_associated(ctx_)) rc = c_osc_ctx_destroy(ctx_)
  ctx_ = c_null_ptr
end subroutine

function last_error() result(msg)
  character(len=:), allocatable :: msg
  type(c_ptr) :: p
  character(kind=c_char), pointer :: s(:)
  integer :: i, n
  p = c_osc_ctx_last_error(ctx_)
  if (.not. c_associated(p)) then
    msg = ''
    return
  endif
  n = int(c_strlen(p))
  call c_f_pointer(p, s, [n])
  allocate(character(len=n) :: msg)
  do i = 1, n
    msg(i:i) = s(i)
  enddo
end function

! any other failure: print the library's message and stop
subroutine fail(rc, where)
  integer(c_int), intent(in) :: rc
  character(len=*), intent(in) :: where
  if (rc == OSC_ERR_FFT_LENGTH) then   ! fft_interface_mod.f90:158-162
    write(6, *) 'Error return from FFT package due to transform length.'
    write(6, *) 'Try increasing the padding by 1 in each dimension and re-run'
    stop
  endif
  write(*, *) where, ': ', last_error(), ' (status ', rc, ')'
  stop 1
end subroutine

function flag(present_, value_, default_) result(r)
  logical, intent(in) :: present_, value_, default_
  integer(c_int) :: r
  logical :: v
  v = default_
  if (present_) v = value_
  r = merge(1_c_int, 0_c_int, v)
end function

function to_c(m) result(c)
  type(mesh3d_struct), intent(in) :: m
  type(osc_mesh_c) :: c
  c%nlo = m%nlo
  c%nhi = m%nhi
  c%npad = m%npad
  c%min = m%min
  c%max = m%max
  c%delta = m%delta
  c%gamma = m%gamma
  c%charge = m%charge
end function

subroutine from_c(c, m)
  type(osc_mesh_c), intent(in) :: c
  type(mesh3d_struct), intent(inout) :: m
  m%min = c%min
  m%max = c%max
  m%delta = c%delta
  m%charge = c%charge
end subroutine

! (re)allocation rule of deposit_particles (open_spacecharge_mod.f90:193-219)
subroutine ensure_arrays(m)
  type(mesh3d_struct), intent(inout) :: m
  integer :: i
  logical :: good
  if (allocated(m%rho)) then
    good = .true.
    do i = 1, 3
      if (lbound(m%rho, i) /= m%nlo(i)) good = .false.
      if (ubound(m%rho, i) /= m%nhi(i)) good = .false.
    enddo
    if (.not. good) then
      deallocate(m%rho)
      if (allocated(m%phi)) deallocate(m%phi)
      if (allocated(m%efield)) deallocate(m%efield)
      if (allocated(m%bfield)) deallocate(m%bfield)
    endif
  endif
  if (.not. allocated(m%rho)) allocate(m%rho(m%nlo(1):m%nhi(1), m%nlo(2):m%nhi(2), m%nlo(3):m%nhi(3)))
  if (.not. allocated(m%phi)) allocate(m%phi(m%nlo(1):m%nhi(1), m%nlo(2):m%nhi(2), m%nlo(3):m%nhi(3)))
  if (.not. allocated(m%efield)) &
    allocate(m%efield(m%nlo(1):m%nhi(1), m%nlo(2):m%nhi(2), m%nlo(3):m%nhi(3), 3))
  if (.not. allocated(m%bfield)) &
    allocate(m%bfield(m%nlo(1):m%nhi(1), m%nlo(2):m%nhi(2), m%nlo(3):m%nhi(3), 3))
end subroutine

!------------------------------------------------------------------------
! print_mesh3d (open_spacecharge_mod.f90:36-52)
!------------------------------------------------------------------------
subroutine print_mesh3d(mesh3d)
  type(mesh3d_struct) :: mesh3d
  print *, '------------------------'
  print *, 'Mesh: '
  print *, 'nlo: ', mesh3d%nlo
  print *, 'nhi: ', mesh3d%nhi
  print *, 'min: ', mesh3d%min
  print *, 'max: ', mesh3d%max
  print *, 'delta: ', mesh3d%delta
  print *, 'gamma: ', mesh3d%gamma
  print *, 'charge: ', mesh3d%charge
  if (allocated(mesh3d%rho)) print *, 'rho allocated'
  if (allocated(mesh3d%phi)) print *, 'phi allocated'
  if (allocated(mesh3d%efield)) print *, 'efield allocated'
  if (allocated(mesh3d%bfield)) print *, 'bfield allocated'
  print *, '------------------------'
end subroutine

!------------------------------------------------------------------------
! deposit_particles (open_spacecharge_mod.f90:156-288)
!------------------------------------------------------------------------
subroutine deposit_particles(xa, ya, za, mesh3d, qa, total_charge, resize_mesh)
  real(dp), target, contiguous :: xa(:), ya(:), za(:)
  type(mesh3d_struct), target :: mesh3d
  real(dp), optional, target, contiguous :: qa(:)
  real(dp), optional :: total_charge
  logical, optional :: resize_mesh
  type(osc_mesh_c) :: m
  type(c_ptr) :: pq, pt
  real(dp), target :: tc
  integer(c_int) :: rc
  integer(c_int64_t) :: ndrop
  logical :: resize

  resize = .true.
  if (present(resize_mesh)) resize = resize_mesh
  call ensure_arrays(mesh3d)
  m = to_c(mesh3d)
  pq = c_null_ptr
  pt = c_null_ptr
  if (present(qa)) then
    pq = c_loc(qa)
  elseif (present(total_charge)) then
    tc = total_charge
    pt = c_loc(tc)
  endif
  ndrop = 0
  rc = c_osc_deposit_particles(ctx(), int(size(xa), c_int64_t), c_loc(xa), c_loc(ya), c_loc(za), &
                               pq, pt, merge(1_c_int, 0_c_int, resize), m, c_loc(mesh3d%rho), ndrop)
  call from_c(m, mesh3d)   ! the resize happens before the charge check, as upstream
  if (rc == OSC_ERR_NO_CHARGE) then
    write(*, *) 'Error, must specify qa or total_charge'
    return
  endif
  if (rc /= OSC_OK) call fail(rc, 'deposit_particles')
  if (ndrop > 0) write(*, *) 'deposit_particles: ', ndrop, ' particles off the mesh were not deposited'
end subroutine deposit_particles

!------------------------------------------------------------------------
! interpolate_field (open_spacecharge_mod.f90:311-390), one point, on the
! host.  Callers interpolate one point per call inside their particle loops
! (code/test_opensc.f90:237-295) and mesh3d%efield/bfield are host arrays,
! so the 8-node sum is taken here, in upstream's operation order (index and
! weight arithmetic, clamping, left-to-right products and sum; this file is
! compiled without FMA contraction), which is also the order of the device
! gather: the result is bit-identical to interpolate_field_batch.  Upstream
! prints ab, de, gh after the j message before computing them (values left
! from the previous call); the same values are printed here.
!------------------------------------------------------------------------
subroutine interpolate_field(x, y, z, mesh3d, E, B)
  type(mesh3d_struct) :: mesh3d
  real(dp) :: x, y, z
  real(dp), optional :: E(3), B(3)
  real(dp), save :: ab = 0, de = 0, gh = 0
  real(dp) :: hxi, hyi, hzi
  integer :: ip, jp, kp, ip1, jp1, kp1
  logical :: clamped

  clamped = .false.
  hxi = 1.d0 / mesh3d%delta(1); hyi = 1.d0 / mesh3d%delta(2); hzi = 1.d0 / mesh3d%delta(3)
  ip = floor((x - mesh3d%min(1)) * hxi + 1)
  jp = floor((y - mesh3d%min(2)) * hyi + 1)
  kp = floor((z - mesh3d%min(3)) * hzi + 1)
  if (ip < 1 .or. ip > mesh3d%nhi(1) - 1) then
    clamped = .true.
    write(6, *) 'ierror: ip=', ip, (x - mesh3d%min(1)), (x - mesh3d%min(1)) / mesh3d%delta(1)
    ip = merge(1, mesh3d%nhi(1) - 1, ip < 1)
  endif
  if (jp < 1 .or. jp > mesh3d%nhi(2) - 1) then
    clamped = .true.
    write(6, *) 'jerror: jp=', jp, (y - mesh3d%min(2)), (y - mesh3d%min(2)) / mesh3d%delta(2)
    write(6, *) ab, de, gh
    jp = merge(1, mesh3d%nhi(2) - 1, jp < 1)
  endif
  if (kp < 1 .or. kp > mesh3d%nhi(3) - 1) then
    clamped = .true.
    write(6, *) 'kerror:  kp=', kp
    write(6, *) 'z=', z
    write(6, *) 'mesh3d%min(3)=', mesh3d%min(3)
    kp = merge(1, mesh3d%nhi(3) - 1, kp < 1)
  endif
  ab = ((mesh3d%min(1) - x) + ip * mesh3d%delta(1)) * hxi
  de = ((mesh3d%min(2) - y) + jp * mesh3d%delta(2)) * hyi
  gh = ((mesh3d%min(3) - z) + kp * mesh3d%delta(3)) * hzi
  if (clamped) write(6, *) ab, de, gh
  ip1 = ip + 1
  jp1 = jp + 1
  kp1 = kp + 1
  if (present(E)) then
    E = mesh3d%efield(ip, jp,  kp,  :)*ab*de*gh                 &
      + mesh3d%efield(ip, jp1, kp,  :)*ab*(1.-de)*gh            &
      + mesh3d%efield(ip, jp1, kp1, :)*ab*(1.-de)*(1.-gh)       &
      + mesh3d%efield(ip, jp,  kp1, :)*ab*de*(1.-gh)            &
      + mesh3d%efield(ip1,jp,  kp1, :)*(1.-ab)*de*(1.-gh)       &
      + mesh3d%efield(ip1,jp1, kp1, :)*(1.-ab)*(1.-de)*(1.-gh)  &
      + mesh3d%efield(ip1,jp1, kp,  :)*(1.-ab)*(1.-de)*gh       &
      + mesh3d%efield(ip1,jp,  kp,  :)*(1.-ab)*de*gh
  endif
  if (present(B)) then
    B = mesh3d%bfield(ip, jp,  kp,  :)*ab*de*gh                 &
      + mesh3d%bfield(ip, jp1, kp,  :)*ab*(1.-de)*gh            &
      + mesh3d%bfield(ip, jp1, kp1, :)*ab*(1.-de)*(1.-gh)       &
      + mesh3d%bfield(ip, jp,  kp1, :)*ab*de*(1.-gh)            &
      + mesh3d%bfield(ip1,jp,  kp1, :)*(1.-ab)*de*(1.-gh)       &
      + mesh3d%bfield(ip1,jp1, kp1, :)*(1.-ab)*(1.-de)*(1.-gh)  &
      + mesh3d%bfield(ip1,jp1, kp,  :)*(1.-ab)*(1.-de)*gh       &
      + mesh3d%bfield(ip1,jp,  kp,  :)*(1.-ab)*de*gh
  endif
end subroutine

! Many points per call: E(3,n), B(3,n) (either optional); n_clamped counts
! points off the mesh (clamped to the edge cell, no messages).
subroutine interpolate_field_batch(x, y, z, mesh3d, E, B, n_clamped)
  real(dp), target, contiguous :: x(:), y(:), z(:)
  type(mesh3d_struct), target :: mesh3d
  real(dp), optional, target, contiguous :: E(:,:), B(:,:)
  integer(int64), optional :: n_clamped
  type(osc_mesh_c) :: m
  type(c_ptr) :: pe, pb, fe, fb
  integer(c_int) :: rc
  integer(c_int64_t) :: ncl

  m = to_c(mesh3d)
  pe = c_null_ptr; pb = c_null_ptr; fe = c_null_ptr; fb = c_null_ptr
  if (present(E)) then
    pe = c_loc(E)
    fe = c_loc(mesh3d%efield)
  endif
  if (present(B)) then
    pb = c_loc(B)
    fb = c_loc(mesh3d%bfield)
  endif
  ncl = 0
  rc = c_osc_interpolate_field(ctx(), m, fe, fb, int(size(x), c_int64_t), c_loc(x), c_loc(y), &
                               c_loc(z), pe, pb, c_null_ptr, ncl)
  if (rc /= OSC_OK) call fail(rc, 'interpolate_field_batch')
  if (present(n_clamped)) n_clamped = ncl
end subroutine

!------------------------------------------------------------------------
! space_charge_3d (open_spacecharge_mod.f90:431-489)
!------------------------------------------------------------------------
subroutine space_charge_3d(mesh3d, offset, at_cathode, calc_bfield)
  type(mesh3d_struct), target :: mesh3d
  real(dp), optional :: offset(3)
  logical, optional :: at_cathode, calc_bfield
  type(osc_mesh_c) :: m
  real(dp), target :: off(3)
  integer(c_int) :: rc, cath, bcalc
  type(c_ptr) :: pb

  off = 0
  if (present(offset)) off = offset
  cath = 0
  if (present(at_cathode)) cath = flag(.true., at_cathode, .false.)
  bcalc = 0
  if (present(calc_bfield)) bcalc = flag(.true., calc_bfield, .false.)
  call ensure_arrays(mesh3d)
  m = to_c(mesh3d)
  pb = c_null_ptr
  if (bcalc == 1) pb = c_loc(mesh3d%bfield)
  rc = c_osc_space_charge_3d(ctx(), m, c_loc(mesh3d%rho), c_loc(off), cath, bcalc, &
                             c_loc(mesh3d%efield), pb)
  if (rc /= OSC_OK) call fail(rc, 'space_charge_3d')
end subroutine

! One tracking step: deposit_particles, space_charge_3d and the gather of E
! (and B with calc_bfield) at the same particles, E(3,n), B(3,n).
subroutine space_charge_step(xa, ya, za, mesh3d, E, B, qa, total_charge, resize_mesh, offset, &
                             at_cathode, calc_bfield)
  real(dp), target, contiguous :: xa(:), ya(:), za(:)
  type(mesh3d_struct), target :: mesh3d
  real(dp), target, contiguous :: E(:,:)
  real(dp), optional, target, contiguous :: B(:,:)
  real(dp), optional, target, contiguous :: qa(:)
  real(dp), optional :: total_charge
  logical, optional :: resize_mesh, at_cathode, calc_bfield
  real(dp), optional :: offset(3)
  type(osc_mesh_c) :: m
  type(c_ptr) :: pq, pt, pb, pB_out
  real(dp), target :: tc, off(3)
  integer(c_int) :: rc, cath, bcalc, resize
  integer(c_int64_t) :: ndrop, ncl

  resize = 1
  if (present(resize_mesh)) resize = flag(.true., resize_mesh, .true.)
  off = 0
  if (present(offset)) off = offset
  cath = 0
  if (present(at_cathode)) cath = flag(.true., at_cathode, .false.)
  bcalc = 0
  if (present(calc_bfield)) bcalc = flag(.true., calc_bfield, .false.)
  call ensure_arrays(mesh3d)
  m = to_c(mesh3d)
  pq = c_null_ptr
  pt = c_null_ptr
  if (present(qa)) then
    pq = c_loc(qa)
  elseif (present(total_charge)) then
    tc = total_charge
    pt = c_loc(tc)
  endif
  pb = c_null_ptr
  pB_out = c_null_ptr
  if (bcalc == 1) then
    pb = c_loc(mesh3d%bfield)
    if (present(B)) pB_out = c_loc(B)
  endif
  rc = c_osc_space_charge_step(ctx(), int(size(xa), c_int64_t), c_loc(xa), c_loc(ya), c_loc(za), &
                               pq, pt, resize, c_loc(off), cath, &
                               bcalc, m, c_loc(mesh3d%rho), c_loc(mesh3d%efield), pb, c_loc(E), &
                               pB_out, ndrop, ncl)
  call from_c(m, mesh3d)
  if (rc == OSC_ERR_NO_CHARGE) then
    write(*, *) 'Error, must specify qa or total_charge'
    return
  endif
  if (rc /= OSC_OK) call fail(rc, 'space_charge_step')
end subroutine

!------------------------------------------------------------------------
! osc_freespace_solver2 (open_spacecharge_mod.f90:535-597)
!------------------------------------------------------------------------
subroutine osc_freespace_solver2(rho, gamma, delta, efield, phi, offset)
  real(dp), intent(in), target, contiguous, dimension(:,:,:) :: rho
  real(dp), intent(in) :: gamma, delta(3)
  real(dp), optional, intent(out), target, contiguous, dimension(:,:,:,:) :: efield
  real(dp), optional, intent(out), target, contiguous, dimension(:,:,:) :: phi
  real(dp), intent(in), optional :: offset(3)
  real(dp), target :: off(3)
  type(c_ptr) :: pe, pp
  integer(c_int) :: rc
  off = 0
  if (present(offset)) off = offset
  pe = c_null_ptr
  pp = c_null_ptr
  if (present(efield)) pe = c_loc(efield)
  if (present(phi)) pp = c_loc(phi)
  rc = c_osc_freespace_solver2(ctx(), int(size(rho, 1), c_int), int(size(rho, 2), c_int), &
                               int(size(rho, 3), c_int), c_loc(rho), gamma, delta, pe, pp, c_loc(off))
  if (rc /= OSC_OK) call fail(rc, 'osc_freespace_solver2')
end subroutine osc_freespace_solver2

!------------------------------------------------------------------------
! osc_get_cgrn_freespace (open_spacecharge_mod.f90:633-707)
!------------------------------------------------------------------------
subroutine osc_get_cgrn_freespace(cgrn, delta, gamma, icomp, offset)
  complex(dp), intent(out), target, contiguous, dimension(:,:,:) :: cgrn
  real(dp), intent(in), dimension(3) :: delta
  integer, intent(in) :: icomp
  real(dp), intent(in) :: gamma
  real(dp), intent(in), optional :: offset(3)
  real(dp), target :: off(3)
  integer(c_int) :: rc
  off = 0
  if (present(offset)) off = offset
  rc = c_osc_get_cgrn_freespace(ctx(), c_loc(cgrn), int(size(cgrn, 1), c_int), &
                                int(size(cgrn, 2), c_int), int(size(cgrn, 3), c_int), delta, gamma, &
                                int(icomp, c_int), c_loc(off))
  if (rc /= OSC_OK) call fail(rc, 'osc_get_cgrn_freespace')
end subroutine

!------------------------------------------------------------------------
! space_charge_freespace (open_spacecharge_mod.f90:60-75)
!------------------------------------------------------------------------
subroutine space_charge_freespace(mesh3d, direct_field_calc, integrated_green_function)
  type(mesh3d_struct), target :: mesh3d
  logical, optional :: direct_field_calc, integrated_green_function
  integer(c_int) :: rc, direct, igf
  direct = 1
  if (present(direct_field_calc)) direct = flag(.true., direct_field_calc, .true.)
  igf = 1
  if (present(integrated_green_function)) igf = flag(.true., integrated_green_function, .true.)
  call ensure_arrays(mesh3d)
  rc = c_osc_space_charge_freespace(ctx(), to_c(mesh3d), c_loc(mesh3d%rho), direct, igf, &
                                    c_loc(mesh3d%phi), c_loc(mesh3d%efield), c_loc(mesh3d%bfield))
  if (rc == OSC_ERR_DIRECT_NEEDS_IGF) then   ! core:58-62
    print *, 'Error: Direct field calc must use integrated Green function'
    print *, 'Aborting...'
    return
  endif
  if (rc /= OSC_OK) call fail(rc, 'space_charge_freespace')
end subroutine

!------------------------------------------------------------------------
! space_charge_cathodeimages (open_spacecharge_mod.f90:83-103); image
! method 2 flips mesh3d%rho in z, as upstream (core:1047-1057)
!------------------------------------------------------------------------
subroutine space_charge_cathodeimages(mesh3d, direct_field_calc, integrated_green_function, &
                                      image_method)
  type(mesh3d_struct), target :: mesh3d
  logical, optional :: direct_field_calc, integrated_green_function
  integer, optional :: image_method
  integer(c_int) :: rc, direct, igf, imethod
  direct = 1
  if (present(direct_field_calc)) direct = flag(.true., direct_field_calc, .true.)
  igf = 1
  if (present(integrated_green_function)) igf = flag(.true., integrated_green_function, .true.)
  imethod = 1
  if (present(image_method)) then
    if (image_method == 2) imethod = 2
  endif
  call ensure_arrays(mesh3d)
  rc = c_osc_space_charge_cathodeimages(ctx(), to_c(mesh3d), c_loc(mesh3d%rho), direct, igf, &
                                        imethod, c_loc(mesh3d%phi), c_loc(mesh3d%efield), &
                                        c_loc(mesh3d%bfield))
  if (rc == OSC_ERR_DIRECT_NEEDS_IGF) then   ! core:935-939
    print *, 'Error: Direct field calc must use integrated Green function'
    print *, 'Aborting...'
    return
  endif
  if (rc /= OSC_OK) call fail(rc, 'space_charge_cathodeimages')
end subroutine

!------------------------------------------------------------------------
! space_charge_rectpipe (open_spacecharge_mod.f90:110-127)
!------------------------------------------------------------------------
subroutine space_charge_rectpipe(mesh3d, apipe, bpipe, direct_field_calc, integrated_green_function)
  type(mesh3d_struct), target :: mesh3d
  real(dp) :: apipe, bpipe
  logical, optional :: direct_field_calc, integrated_green_function
  integer(c_int) :: rc, direct, igf
  direct = 1
  if (present(direct_field_calc)) direct = flag(.true., direct_field_calc, .true.)
  igf = 1
  if (present(integrated_green_function)) igf = flag(.true., integrated_green_function, .true.)
  call ensure_arrays(mesh3d)
  rc = c_osc_space_charge_rectpipe(ctx(), to_c(mesh3d), apipe, bpipe, direct, igf, &
                                   c_loc(mesh3d%rho), c_loc(mesh3d%phi), c_loc(mesh3d%efield), &
                                   c_loc(mesh3d%bfield))
  if (rc == OSC_ERR_PIPE_DIRECT) then   ! core:608-610
    write(6, *) '**********ERROR********* rectpipe implementation presently requires directfieldcalc=0'
    return
  endif
  if (rc == OSC_ERR_NO_GREEN) then      ! core:578-581
    write(6, *) 'something wrong. rectpipe green functions should have been computed during initialization'
    stop
  endif
  if (rc /= OSC_OK) call fail(rc, 'space_charge_rectpipe')
end subroutine

!------------------------------------------------------------------------
! pipe Green functions (open_spacecharge_core_mod.f90:632-715, 825-898)
!------------------------------------------------------------------------
subroutine osc_alloc_rectpipe_arrays(nlo, nhi, npad)
  integer, intent(in), dimension(3) :: nlo, nhi, npad
  integer(c_int) :: rc
  rc = c_osc_alloc_rectpipe_arrays(ctx(), int(nlo, c_int32_t), int(nhi, c_int32_t), &
                                   int(npad, c_int32_t))
  if (rc /= OSC_OK) call fail(rc, 'osc_alloc_rectpipe_arrays')
end subroutine

subroutine osc_getgrnpipe(gam, a, b, delta, umin, npad)
  real(dp), dimension(3) :: delta, umin
  integer, dimension(3) :: npad
  real(dp) :: gam, a, b
  integer(c_int) :: rc
  rc = c_osc_getgrnpipe(ctx(), gam, a, b, delta, umin, int(npad, c_int32_t))
  if (rc /= OSC_OK) call fail(rc, 'osc_getgrnpipe')
end subroutine

subroutine osc_read_rectpipe_grn
  real(c_double) :: hr(12)
  integer(c_int32_t) :: hi(9)
  integer(c_int) :: rc
  rc = c_osc_read_rectpipe_grn(ctx(), 'rectpipegrn.dat' // c_null_char, hr, hi)
  if (rc /= OSC_OK) call fail(rc, 'osc_read_rectpipe_grn')
  write(6, *) 'Done reading rectpipe transformed Green functions with these parameters:'
  write(6, *) 'apipe,bpipe=', hr(1), hr(2)
  write(6, *) 'dx,dy,dz=', hr(3), hr(4), hr(5)
  write(6, *) 'xmin,ymin,zmin=', hr(6), hr(7), hr(8)
  write(6, *) 'xmax,ymax,zmax=', hr(9), hr(10), hr(11)
  write(6, *) 'ilo,jlo,klo=', hi(1), hi(2), hi(3)
  write(6, *) 'ihi,jhi,khi=', hi(4), hi(5), hi(6)
  write(6, *) 'convolution iperiod,jperiod,kperiod=', hi(7), hi(8), hi(9)
  write(6, *) 'gamma=', hr(12)
end subroutine

subroutine osc_write_rectpipe_grn(apipe, bpipe, delta, umin, umax, nlo, nhi, gamma)
  real(dp) :: apipe, bpipe, gamma
  real(dp), dimension(3) :: delta, umin, umax
  integer, dimension(3) :: nlo, nhi
  integer(c_int) :: rc
  rc = c_osc_write_rectpipe_grn(ctx(), 'rectpipegrn.dat' // c_null_char, apipe, bpipe, delta, umin, &
                                umax, int(nlo, c_int32_t), int(nhi, c_int32_t), gamma)
  if (rc /= OSC_OK) call fail(rc, 'osc_write_rectpipe_grn')
  write(6, *) 'done writing rectpipe transformed Green functions'
end subroutine

end module open_spacecharge_mod
