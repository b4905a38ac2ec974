"""cueva: MI355X (gfx950) implementation of one CUEVA 2D resistive-RMHD step.

Host-side Python mirror of the reference's per-stage subroutine interface
(main_cueva_code.f95:47-356) over the C ABI of include/cueva.h.
"""
from .abi import (  # noqa: F401
    BC_COPY, BC_PERIODIC, C2P_CARDANO, C2P_LEARNED, CONS_NAMES, FP64, GLOBAL_NAMES, HLL, HLL_MCL, HLLC, HLLC_MCL, HLLE, HLLE_MCL, LAX, MIXED,
    MLP_NPACK, MLP_NPARAM, PRIM_NAMES, CuevaError, Mlp, Params, Solver, default_params, derived, kh_init, load,
    plane_shape,
)
from .problems import kh_params, kh_problem  # noqa: F401
from .restart import read_restart, read_restart_text, restart_file_name  # noqa: F401
