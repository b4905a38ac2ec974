"""ctypes binding of the CPU restatement (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, as the checker of the HIP path.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


class Cfg(ctypes.Structure):
    _fields_ = [("imax", ctypes.c_int32), ("jmax", ctypes.c_int32), ("lmax", ctypes.c_int32),
                ("flux_solver", ctypes.c_int32), ("rec_order", ctypes.c_int32), ("bc_x", ctypes.c_int32),
                ("bc_y", ctypes.c_int32), ("pad0", ctypes.c_int32), ("dt", ctypes.c_double),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("sigma", ctypes.c_double),
                ("kappa_psi", ctypes.c_double), ("kappa_phi", ctypes.c_double), ("gamma", ctypes.c_double),
                ("wb", ctypes.c_double * 8), ("wbt", ctypes.c_double * 8), ("Ab", ctypes.c_double * 64),
                ("Abt", ctypes.c_double * 64)]


_P = ctypes.POINTER(ctypes.c_double)
_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("liboracle.so not built; run make -C oracle")
        lib = ctypes.CDLL(LIB)
        lib.oracle_default_cfg.argtypes = [ctypes.POINTER(Cfg), ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        lib.oracle_create.restype = ctypes.c_void_p
        lib.oracle_create.argtypes = [ctypes.POINTER(Cfg)]
        lib.oracle_destroy.argtypes = [ctypes.c_void_p]
        lib.oracle_set_prims.argtypes = [ctypes.c_void_p, _P]
        lib.oracle_get_fields.argtypes = [ctypes.c_void_p, _P, _P]
        lib.oracle_get_counters.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]
        lib.oracle_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.oracle_get_rhs.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _P]
        lib.oracle_mp_line.argtypes = [ctypes.c_int, _P, ctypes.c_int, _P, _P, ctypes.c_int]
        lib.oracle_hllc_vel.restype = ctypes.c_double
        lib.oracle_hllc_vel.argtypes = [ctypes.c_double] * 5
        lib.oracle_c2p_point.argtypes = [ctypes.c_double, _P, _P, ctypes.POINTER(ctypes.c_int)]
        lib.oracle_set_hllc_seed_mode.argtypes = [ctypes.c_int]
        lib.oracle_set_primitive_rec.argtypes = [ctypes.c_int]
        lib.oracle_set_libm_mode.argtypes = [ctypes.c_int]
        lib.oracle_globals.argtypes = [ctypes.c_void_p, _P]
        _lib = lib
    return _lib


def _dp(a):
    return a.ctypes.data_as(_P)


class Oracle:
    """CPU restatement of the step for one grid; mirrors cueva.Solver's data calls."""

    def __init__(self, params, libm=None):
        """params: a cueva.Params (the same parameters the HIP path gets).
        libm: "c" (C library, as the reference build) or "cr" (correctly rounded);
        default follows params.crmath."""
        lib = load()
        self.lib = lib
        self.c = Cfg()
        order = {0: 5, 1: 7, 2: 9}[params.rec_mp]
        rc = lib.oracle_default_cfg(ctypes.byref(self.c), params.imax, params.jmax, params.longx, params.longy,
                                    2 if params.order == 2 else 4, params.flux_solver, order, params.bc_x,
                                    params.bc_y, params.sigma, params.cfl)
        if rc:
            raise ValueError("oracle config rejected (%d)" % rc)
        lib.oracle_set_hllc_seed_mode(int(params.hllc_seed))
        lib.oracle_set_primitive_rec(int(params.primitive_rec))
        if libm is None:
            libm = "cr" if params.crmath else "c"
        lib.oracle_set_libm_mode(1 if libm == "cr" else 0)
        self.s = lib.oracle_create(ctypes.byref(self.c))
        self.shape = (params.jmax + 13, params.imax + 13)

    def close(self):
        if self.s:
            self.lib.oracle_destroy(self.s)
            self.s = None

    def __del__(self):
        self.close()

    def upload(self, prims):
        prims = np.ascontiguousarray(prims, dtype=np.float64)
        self.lib.oracle_set_prims(self.s, _dp(prims))

    def step(self, n=1):
        rc = self.lib.oracle_step(self.s, int(n))
        if rc:
            raise RuntimeError("oracle stopped with status %d" % rc)

    def download(self):
        prims = np.zeros((14,) + self.shape)
        cons = np.zeros((5,) + self.shape)
        self.lib.oracle_get_fields(self.s, _dp(prims), _dp(cons))
        return prims, cons

    def counters(self):
        c = (ctypes.c_int64 * 3)()
        self.lib.oracle_get_counters(self.s, c)
        return list(c)

    def globals(self):
        """37 diagnostics of 22_writer.f95:818-991, sequential as the reference sums them."""
        out = np.zeros(37)
        self.lib.oracle_globals(self.s, _dp(out))
        return out
