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
                ("Abt", ctypes.c_double * 64), ("primitive_rec", ctypes.c_int32), ("hllc_seed", ctypes.c_int32),
                ("libm_mode", ctypes.c_int32), ("slop", ctypes.c_int32), ("limit", ctypes.c_int32),
                ("mirk", ctypes.c_int32), ("rec_prim", ctypes.c_int32), ("pad1", ctypes.c_int32)]


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
        lib.oracle_scheme_tables.argtypes = [ctypes.POINTER(Cfg)]
        lib.oracle_create.restype = ctypes.c_void_p
        lib.oracle_create.argtypes = [ctypes.POINTER(Cfg)]
        lib.oracle_destroy.argtypes = [ctypes.c_void_p]
        lib.oracle_set_prims.argtypes = [ctypes.c_void_p, _P]
        lib.oracle_get_fields.argtypes = [ctypes.c_void_p, _P, _P]
        lib.oracle_get_counters.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]
        lib.oracle_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.oracle_get_rhs.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _P]
        lib.oracle_get_face.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _P]
        lib.oracle_mp_line.argtypes = [ctypes.c_int, _P, ctypes.c_int, _P, _P, ctypes.c_int]
        lib.oracle_hllc_vel.restype = ctypes.c_double
        lib.oracle_hllc_vel.argtypes = [ctypes.c_double] * 5
        lib.oracle_c2p_point.argtypes = [ctypes.c_double, ctypes.c_double, _P, _P, ctypes.POINTER(ctypes.c_int), _P]
        lib.oracle_set_cons.argtypes = [ctypes.c_void_p, _P]
        lib.oracle_recover.argtypes = [ctypes.c_void_p]
        lib.oracle_last_stop.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]
        lib.oracle_face_events.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]
        lib.oracle_get_face_events.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32), _P, ctypes.c_int]
        lib.oracle_get_fixups.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int32), _P,
                                          ctypes.c_int]
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
                                    int(params.order), params.flux_solver, order, params.bc_x,
                                    params.bc_y, params.sigma, params.cfl)
        if rc:
            raise ValueError("oracle config rejected (%d)" % rc)
        self.c.hllc_seed = int(params.hllc_seed)
        self.c.primitive_rec = int(params.primitive_rec)
        self.c.slop = int(getattr(params, "slop", 2) or 2)
        self.c.limit = int(getattr(params, "limit", 2) or 2)
        self.c.mirk = int(getattr(params, "mirk", 0))
        self.c.rec_prim = int(getattr(params, "rec_prim", 0))
        if lib.oracle_scheme_tables(ctypes.byref(self.c)):
            raise ValueError("oracle: order %d with MIRK = %d is not a scheme of the reference"
                             % (params.order, self.c.mirk))
        if libm is None:
            libm = "cr" if params.crmath else "c"
        self.c.libm_mode = 1 if libm == "cr" else 0
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

    def upload_cons(self, cons):
        cons = np.ascontiguousarray(cons, dtype=np.float64)
        self.lib.oracle_set_cons(self.s, _dp(cons))

    def recover(self):
        """varprimitivecardano alone (l_swap = 1); returns 0 or the stop status."""
        return int(self.lib.oracle_recover(self.s))

    def last_stop(self):
        """(step within the last step() call, stage, i, j) of the last nonzero status."""
        out = (ctypes.c_int32 * 4)()
        self.lib.oracle_last_stop(self.s, out)
        return tuple(out)

    def face_events(self):
        """(superluminal face fallbacks, HLLC face-density fixes) since creation."""
        out = (ctypes.c_int64 * 2)()
        self.lib.oracle_face_events(self.s, out)
        return tuple(out)

    def rhs(self, l, which):
        """Stage right-hand side `which` (0..16) of stage l (0-based) of the last step."""
        out = np.zeros(self.shape)
        self.lib.oracle_get_rhs(self.s, int(l), int(which), _dp(out))
        return out

    def face(self, direction, slot):
        """Face flux `slot` of the last face pass (direction 0: x faces, 1: y faces)."""
        out = np.zeros(self.shape)
        self.lib.oracle_get_face(self.s, int(direction), int(slot), _dp(out))
        return out

    def face_event_list(self):
        """(kind, l, i, j) int32 [n, 4] and values [n] of every face event since creation."""
        n = self.lib.oracle_get_face_events(self.s, None, None, 0)
        k = np.zeros((max(n, 1), 4), dtype=np.int32)
        v = np.zeros(max(n, 1))
        self.lib.oracle_get_face_events(self.s, k.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _dp(v), n)
        return k[:n], v[:n]

    def step_status(self, n=1):
        """Like step() but returns the status instead of raising."""
        return int(self.lib.oracle_step(self.s, int(n)))

    def fixups(self, kind):
        """Fix-up records of the last recovery, kind 0 rho / 1 p / 2 drift:
        (ij int32 [n, 2], values [n, 2]) in the reference's loop order."""
        n = self.lib.oracle_get_fixups(self.s, kind, None, None, 0)
        ij = np.zeros((max(n, 1), 2), dtype=np.int32)
        vals = np.zeros((max(n, 1), 2))
        self.lib.oracle_get_fixups(self.s, kind, ij.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _dp(vals), n)
        return ij[:n], vals[:n]

    def counters(self):
        c = (ctypes.c_int64 * 3)()
        self.lib.oracle_get_counters(self.s, c)
        return list(c)

    def globals(self):
        """37 diagnostics of 22_writer.f95:818-991, sequential as the reference sums them."""
        out = np.zeros(37)
        self.lib.oracle_globals(self.s, _dp(out))
        return out
