"""MI355X-native OpenSC space-charge library.

The public API mirrors the upstream Fortran module ``open_spacecharge_mod``;
the compute path is the HIP library ``libopensc_hip.so`` (C-ABI in
include/opensc_hip.h).  There is no CPU fallback: without the built library
every call raises ``LibraryNotBuilt``.
"""
from ._capi import (OSC_PREC_FP32, OSC_PREC_FP64, Context, LibraryNotBuilt, OSCError)  # noqa: F401
from .api import Server  # noqa: F401
from .api import (Mesh3D, OpenSCStop, ccfft3d, context, deposit_particles,  # noqa: F401
                  depose_rho_scalar, get_mesh_quantities, interpolate_field, osc_alloc_rectpipe_arrays,
                  osc_freespace_solver2, osc_get_cgrn_freespace, osc_getgrnfree,
                  osc_getgrnimageconvcorr, osc_getgrnimageshift, osc_getgrnpipe,
                  osc_read_rectpipe_grn, osc_write_rectpipe_grn, print_mesh3d, set_context,
                  space_charge_3d, space_charge_cathodeimages, space_charge_freespace,
                  space_charge_rectpipe, space_charge_step)

__version__ = "0.1.0"
