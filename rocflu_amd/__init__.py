"""MI355X-native explicit residual/update path of RocfluMP.

Face fluxes (Roe, HLLC, AUSM+), WLSQ cell gradients with WENO-XYZ or
Venkatakrishnan modification, the per-cell residual gather and the RK stage update
run as HIP kernels for gfx950 behind the C-ABI in include/rocflu_amd.h.
"""
from .mesh import box, Mesh, HEX, TET6, BC_OUTFLOW, BC_SLIPWALL  # noqa: F401
from .solver import (GpuRegion, Numerics, n_stages, DISCR_ROE, DISCR_HLLC,  # noqa: F401
                     DISCR_AUSMPLUS, RECONST_NONE, RECONST_WENO_XYZ, RECONST_LIM_VENKAT,
                     RK4_CLASSICAL, RK3_WRAY, RK1_EULER, RK3_SSP, PREC_FP64, PREC_MIXED)

__version__ = "0.1.0"
