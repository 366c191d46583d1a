! ISO_C_BINDING interface to the C-ABI of include/opensc_hip.h
! (libopensc_hip.so).  Every pointer argument is passed as type(c_ptr) by
! value so that absent optional arrays become NULL; grids are Fortran
! column-major as the C side expects.
module opensc_hip
use, intrinsic :: iso_c_binding
implicit none
private

integer(c_int), parameter, public :: OSC_OK = 0, OSC_ERR_ARG = 1, OSC_ERR_NO_CHARGE = 2, &
  OSC_ERR_FFT_LENGTH = 3, OSC_ERR_HIP = 4, OSC_ERR_NOMEM = 5, OSC_ERR_DIRECT_NEEDS_IGF = 6, &
  OSC_ERR_PIPE_DIRECT = 7, OSC_ERR_NO_GREEN = 8, OSC_ERR_IO = 9, OSC_ERR_NO_DEVICE = 10
integer(c_int), parameter, public :: OSC_PREC_FP64 = 0, OSC_PREC_FP32 = 1

! osc_mesh: mesh3d_struct without its arrays
type, bind(C), public :: osc_mesh_c
  integer(c_int32_t) :: nlo(3), nhi(3), npad(3)
  real(c_double) :: min(3), max(3), delta(3)
  real(c_double) :: gamma, charge
end type

public :: c_osc_ctx_create, c_osc_ctx_destroy, c_osc_ctx_last_error, c_osc_ctx_set_precision, &
  c_osc_ctx_set_green_cache, c_osc_deposit_particles, c_osc_space_charge_step, &
  c_osc_space_charge_3d, c_osc_freespace_solver2, c_osc_get_cgrn_freespace, &
  c_osc_space_charge_freespace, c_osc_space_charge_cathodeimages, c_osc_alloc_rectpipe_arrays, &
  c_osc_getgrnpipe, c_osc_write_rectpipe_grn, c_osc_read_rectpipe_grn, &
  c_osc_space_charge_rectpipe, c_osc_interpolate_field, c_strlen

interface
  integer(c_int) function c_osc_ctx_create(device, ctx) bind(C, name='osc_ctx_create')
    import :: c_int, c_ptr
    integer(c_int), value :: device
    type(c_ptr) :: ctx
  end function
  integer(c_int) function c_osc_ctx_destroy(ctx) bind(C, name='osc_ctx_destroy')
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
  end function
  type(c_ptr) function c_osc_ctx_last_error(ctx) bind(C, name='osc_ctx_last_error')
    import :: c_ptr
    type(c_ptr), value :: ctx
  end function
  integer(c_int) function c_osc_ctx_set_precision(ctx, prec) bind(C, name='osc_ctx_set_precision')
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int), value :: prec
  end function
  integer(c_int) function c_osc_ctx_set_green_cache(ctx, enable, max_bytes) &
      bind(C, name='osc_ctx_set_green_cache')
    import :: c_int, c_ptr, c_int64_t
    type(c_ptr), value :: ctx
    integer(c_int), value :: enable
    integer(c_int64_t), value :: max_bytes
  end function
  integer(c_int) function c_osc_deposit_particles(ctx, n, xa, ya, za, qa, total_charge, resize, &
      mesh, rho, n_dropped) bind(C, name='osc_deposit_particles')
    import :: c_int, c_ptr, c_int64_t, osc_mesh_c
    type(c_ptr), value :: ctx
    integer(c_int64_t), value :: n
    type(c_ptr), value :: xa, ya, za, qa, total_charge
    integer(c_int), value :: resize
    type(osc_mesh_c) :: mesh
    type(c_ptr), value :: rho
    integer(c_int64_t) :: n_dropped
  end function
  integer(c_int) function c_osc_space_charge_step(ctx, n, xa, ya, za, qa, total_charge, resize, &
      offset, at_cathode, calc_bfield, mesh, rho, efield, bfield, e, b, n_dropped, n_clamped) &
      bind(C, name='osc_space_charge_step')
    import :: c_int, c_ptr, c_int64_t, osc_mesh_c
    type(c_ptr), value :: ctx
    integer(c_int64_t), value :: n
    type(c_ptr), value :: xa, ya, za, qa, total_charge
    integer(c_int), value :: resize
    type(c_ptr), value :: offset
    integer(c_int), value :: at_cathode, calc_bfield
    type(osc_mesh_c) :: mesh
    type(c_ptr), value :: rho, efield, bfield, e, b
    integer(c_int64_t) :: n_dropped, n_clamped
  end function
  integer(c_int) function c_osc_space_charge_3d(ctx, mesh, rho, offset, at_cathode, calc_bfield, &
      efield, bfield) bind(C, name='osc_space_charge_3d')
    import :: c_int, c_ptr, osc_mesh_c
    type(c_ptr), value :: ctx
    type(osc_mesh_c), intent(in) :: mesh
    type(c_ptr), value :: rho, offset
    integer(c_int), value :: at_cathode, calc_bfield
    type(c_ptr), value :: efield, bfield
  end function
  integer(c_int) function c_osc_freespace_solver2(ctx, nx, ny, nz, rho, gamma, delta, efield, &
      phi, offset) bind(C, name='osc_freespace_solver2')
    import :: c_int, c_ptr, c_double
    type(c_ptr), value :: ctx
    integer(c_int), value :: nx, ny, nz
    type(c_ptr), value :: rho
    real(c_double), value :: gamma
    real(c_double), intent(in) :: delta(3)
    type(c_ptr), value :: efield, phi, offset
  end function
  integer(c_int) function c_osc_get_cgrn_freespace(ctx, cgrn, isize, jsize, ksize, delta, gamma, &
      icomp, offset) bind(C, name='osc_get_cgrn_freespace')
    import :: c_int, c_ptr, c_double
    type(c_ptr), value :: ctx, cgrn
    integer(c_int), value :: isize, jsize, ksize
    real(c_double), intent(in) :: delta(3)
    real(c_double), value :: gamma
    integer(c_int), value :: icomp
    type(c_ptr), value :: offset
  end function
  integer(c_int) function c_osc_space_charge_freespace(ctx, mesh, rho, direct, igf, phi, efield, &
      bfield) bind(C, name='osc_space_charge_freespace')
    import :: c_int, c_ptr, osc_mesh_c
    type(c_ptr), value :: ctx
    type(osc_mesh_c), intent(in) :: mesh
    type(c_ptr), value :: rho
    integer(c_int), value :: direct, igf
    type(c_ptr), value :: phi, efield, bfield
  end function
  integer(c_int) function c_osc_space_charge_cathodeimages(ctx, mesh, rho, direct, igf, method, &
      phi, efield, bfield) bind(C, name='osc_space_charge_cathodeimages')
    import :: c_int, c_ptr, osc_mesh_c
    type(c_ptr), value :: ctx
    type(osc_mesh_c), intent(in) :: mesh
    type(c_ptr), value :: rho
    integer(c_int), value :: direct, igf, method
    type(c_ptr), value :: phi, efield, bfield
  end function
  integer(c_int) function c_osc_alloc_rectpipe_arrays(ctx, nlo, nhi, npad) &
      bind(C, name='osc_alloc_rectpipe_arrays')
    import :: c_int, c_ptr, c_int32_t
    type(c_ptr), value :: ctx
    integer(c_int32_t), intent(in) :: nlo(3), nhi(3), npad(3)
  end function
  integer(c_int) function c_osc_getgrnpipe(ctx, gamma, apipe, bpipe, delta, umin, npad) &
      bind(C, name='osc_getgrnpipe')
    import :: c_int, c_ptr, c_double, c_int32_t
    type(c_ptr), value :: ctx
    real(c_double), value :: gamma, apipe, bpipe
    real(c_double), intent(in) :: delta(3), umin(3)
    integer(c_int32_t), intent(in) :: npad(3)
  end function
  integer(c_int) function c_osc_write_rectpipe_grn(ctx, path, apipe, bpipe, delta, umin, umax, &
      nlo, nhi, gamma) bind(C, name='osc_write_rectpipe_grn')
    import :: c_int, c_ptr, c_double, c_int32_t, c_char
    type(c_ptr), value :: ctx
    character(kind=c_char), intent(in) :: path(*)
    real(c_double), value :: apipe, bpipe
    real(c_double), intent(in) :: delta(3), umin(3), umax(3)
    integer(c_int32_t), intent(in) :: nlo(3), nhi(3)
    real(c_double), value :: gamma
  end function
  integer(c_int) function c_osc_read_rectpipe_grn(ctx, path, hdr_real, hdr_int) &
      bind(C, name='osc_read_rectpipe_grn')
    import :: c_int, c_ptr, c_double, c_int32_t, c_char
    type(c_ptr), value :: ctx
    character(kind=c_char), intent(in) :: path(*)
    real(c_double) :: hdr_real(12)
    integer(c_int32_t) :: hdr_int(9)
  end function
  integer(c_int) function c_osc_space_charge_rectpipe(ctx, mesh, apipe, bpipe, direct, igf, rho, &
      phi, efield, bfield) bind(C, name='osc_space_charge_rectpipe')
    import :: c_int, c_ptr, c_double, osc_mesh_c
    type(c_ptr), value :: ctx
    type(osc_mesh_c), intent(in) :: mesh
    real(c_double), value :: apipe, bpipe
    integer(c_int), value :: direct, igf
    type(c_ptr), value :: rho, phi, efield, bfield
  end function
  integer(c_int) function c_osc_interpolate_field(ctx, mesh, efield, bfield, n, x, y, z, e, b, &
      clamped, n_clamped) bind(C, name='osc_interpolate_field')
    import :: c_int, c_ptr, c_int64_t, osc_mesh_c
    type(c_ptr), value :: ctx
    type(osc_mesh_c), intent(in) :: mesh
    type(c_ptr), value :: efield, bfield
    integer(c_int64_t), value :: n
    type(c_ptr), value :: x, y, z, e, b, clamped
    integer(c_int64_t) :: n_clamped
  end function
  integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
    import :: c_size_t, c_ptr
    type(c_ptr), value :: s
  end function
end interface

end module opensc_hip
