"""ctypes binding of the C ABI in include/cueva.h.

This is the Python host side of the drop-in boundary: every method maps to
one ``extern "C"`` entry point, which replaces one subroutine of the
reference time step (main_cueva_code.f95:47-356).  Loading fails loudly when
the HIP library has not been built: there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "lib", "libcueva.so")

HLL, HLLE, HLLC = 5, 6, 7
HLL_MCL = 2  # HLL with the second-order monotonised-central limiter (16_hll_flow_mcl.f95)
LAX = 1  # Lax-Friedrichs flux splitting with the same limiter (15_lax_flow.f95)
HLLE_MCL = 3  # HLLE with the same limiter (16_hlle_flow_mcl.f95)
HLLC_MCL = 4  # HLLC with the same limiter (16_hllc_flow_mcl.f95)
# output order of cueva_reduce_globals (the reference's units 80..217)
GLOBAL_NAMES = ("sum_psi", "sum_phi", "sum_ex", "sum_ey", "sum_ez", "sum_exey", "sum_exez", "sum_eyez",
                "sum_bx", "sum_by", "sum_bz", "sum_bxby", "sum_bxbz", "sum_bybz", "sum_vxvyvz", "sum_bxbybz",
                "sum_vx", "sum_vy", "sum_vz", "sum_vxvy", "sum_vxvz", "sum_vyvz", "sum_q", "sum_p", "sum_rho",
                "ekin_c", "ekin_r", "etot", "efluid_r", "emag", "eint_c", "eint_r", "lorentz", "sum_D",
                "sum_D_med", "max_val_bx", "max_val_phi")
BC_COPY, BC_PERIODIC = 0, 1
FP64, MIXED = 0, 1
C2P_CARDANO, C2P_LEARNED = 0, 1
BND_PRIM, BND_CONS, BND_ELEC, BND_FLUX = 0, 1, 2, 3

# field order of the 14 host planes (restart order, 22_writer.f95:1116-1172)
PRIM_NAMES = ("psi", "phi", "Bx", "By", "Bz", "Vx", "Vy", "Vz", "Ex", "Ey", "Ez", "q", "rho", "p")
CONS_NAMES = ("D", "tau", "Sx", "Sy", "Sz")

STATUS = {
    0: "ok", 1: "invalid argument", 2: "HIP error", 3: "out of device memory",
    7: "get_hllc_vel: no convergence", 8: "HLLC contact speed is NaN", 9: "HLLE wave speed is NaN",
    20: "(not returned)", 21: "Cardano: no admissible resolvent root",
    22: "Cardano: det < 0", 23: "Cardano: both roots below 1", 24: "Cardano: no root selected",
    25: "superluminal velocity, magnetisation below 10", 26: "drift velocity still superluminal",
}


class CuevaError(RuntimeError):
    def __init__(self, code, where=""):
        self.code = code
        super().__init__("cueva %s failed: %s (status %d)" % (where, STATUS.get(code, "?"), code))


class Params(ctypes.Structure):
    _fields_ = [
        ("imax", ctypes.c_int32), ("jmax", ctypes.c_int32), ("order", ctypes.c_int32),
        ("flux_solver", ctypes.c_int32), ("rec_mp", ctypes.c_int32),
        ("bc_x", ctypes.c_int32), ("bc_y", ctypes.c_int32), ("precision", ctypes.c_int32),
        ("c2p_mode", ctypes.c_int32), ("hllc_seed", ctypes.c_int32), ("crmath", ctypes.c_int32),
        ("primitive_rec", ctypes.c_int32), ("slop", ctypes.c_int32), ("limit", ctypes.c_int32),
        ("mirk", ctypes.c_int32), ("rec_prim", ctypes.c_int32),
        ("longx", ctypes.c_double), ("longy", ctypes.c_double), ("cfl", ctypes.c_double),
        ("sigma", ctypes.c_double), ("gamma", ctypes.c_double), ("dt", ctypes.c_double),
        ("reserved_d", ctypes.c_double * 6),
    ]


# every symbol include/cueva.h declares, with (restype, argtypes)
_P = ctypes.POINTER
_c = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double
SIGNATURES = {
    "cueva_abi_version": (_i, []),
    "cueva_params_default": (_i, [_P(Params)]),
    "cueva_derived": (_i, [_P(Params), _P(_d)]),
    "cueva_create": (_i, [_P(Params), _P(_c)]),
    "cueva_destroy": (_i, [_c]),
    "cueva_sync": (_i, [_c]),
    "cueva_device_bytes": (_i, [_c, _P(ctypes.c_int64)]),
    "cueva_upload_prims": (_i, [_c, _P(_d)]),
    "cueva_download": (_i, [_c, _P(_d), _P(_d)]),
    "cueva_boundary": (_i, [_c, _i, _i]),
    "cueva_recons_var_primitive": (_i, [_c]),
    "cueva_startconvar": (_i, [_c]),
    "cueva_cleanvarint": (_i, [_c]),
    "cueva_swap1": (_i, [_c]),
    "cueva_c2p_stage": (_i, [_c, _i]),
    "cueva_interconvar": (_i, [_c, _i]),
    "cueva_sumint": (_i, [_c, _i, _i]),
    "cueva_varint": (_i, [_c, _i]),
    "cueva_electricvarint": (_i, [_c, _i]),
    "cueva_flowinit": (_i, [_c, _i]),
    "cueva_flux": (_i, [_c, _i]),
    "cueva_fluxandsource": (_i, [_c, _i]),
    "cueva_sumfin": (_i, [_c, _i]),
    "cueva_varfin": (_i, [_c]),
    "cueva_electricfield": (_i, [_c]),
    "cueva_c2p": (_i, [_c, _i, _P(ctypes.c_int32), _P(ctypes.c_int32)]),
    "cueva_step": (_i, [_c, _i]),
    "cueva_step_counters": (_i, [_c, _P(ctypes.c_int64), _P(ctypes.c_int64)]),
    "cueva_last_error": (_i, [_c, _P(ctypes.c_int32)]),
    "cueva_set_step_index": (_i, [_c, ctypes.c_int64]),
    "cueva_fixup_records": (_i, [_c, _i, _P(ctypes.c_int32), _P(_d), ctypes.c_int64, _P(ctypes.c_int64)]),
    "cueva_reduce_globals": (_i, [_c, _P(_d)]),
    "cueva_max_speed": (_i, [_c, _P(_d)]),
    "cueva_reduce_globals_noalloc": (_i, [_P(Params), _c, _c, _c, _c]),
    "cueva_kh_init": (_i, [_i, _i, _d, _d, _i, _d, _d, _d, ctypes.c_uint64, _P(_d)]),
    "cueva_upload_cons": (_i, [_c, _P(_d)]),
    "cueva_download_plane": (_i, [_c, _i, _i, _P(_d)]),
    "cueva_plane_range": (_i, [_c, _i, _i, _P(_d)]),
    "cueva_set_profiling": (_i, [_c, _i]),
    "cueva_kernel_times": (_i, [_c, _P(_d), _P(ctypes.c_int64)]),
    "cueva_time_steps": (_i, [_c, _i, _P(_d)]),
    "cueva_set_c2p_params": (_i, [_c, _P(ctypes.c_float)]),
    "cueva_cpaw_init": (_i, [_i, _i, _d, _d, _d, _d, _d, _d, _P(_d), _P(_d)]),
    "cueva_mlp_create": (_i, [_P(_c)]),
    "cueva_mlp_destroy": (_i, [_c]),
    "cueva_mlp_set_params": (_i, [_c, _P(ctypes.c_float)]),
    "cueva_mlp_get_params": (_i, [_c, _P(ctypes.c_float)]),
    "cueva_mlp_set_data": (_i, [_c, _P(ctypes.c_float), _P(ctypes.c_float), ctypes.c_int64]),
    "cueva_mlp_grad": (_i, [_c, ctypes.c_int64, ctypes.c_int64, _P(ctypes.c_float), _P(_d)]),
    "cueva_mlp_train": (_i, [_c, ctypes.c_int64, _i, ctypes.c_float, _P(_d)]),
    "cueva_mlp_forward": (_i, [_c, _P(ctypes.c_float), ctypes.c_int64, _P(ctypes.c_float)]),
}

KERNEL_CLASSES = ("stage", "faces", "fluxsource", "final", "ghosts", "other")

_lib = None


def load(path=LIB_PATH):
    """Load libcueva.so (built by cueva.build.build_hip)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libcueva.so not built (%s); run __graft_entry__.build()" % path)
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(_P(_d))


def check(code, where):
    if code != 0:
        raise CuevaError(code, where)


def default_params(**kw):
    lib = load()
    p = Params()
    check(lib.cueva_params_default(ctypes.byref(p)), "params_default")
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def derived(p):
    out = (ctypes.c_double * 5)()
    check(load().cueva_derived(ctypes.byref(p), out), "derived")
    return dict(zip(("dx", "dy", "dt", "kappa_psi", "kappa_phi"), list(out)))


def plane_shape(p):
    return (p.jmax + 13, p.imax + 13)


def kh_init(imax, jmax, lx, ly, rotated, v0=0.5, bx0=0.01, bz0=0.0, seed=20261202):
    """Seeded KH primitives, shape (14, jmax+13, imax+13) (cueva_kh_init)."""
    out = np.zeros((14, jmax + 13, imax + 13))
    check(load().cueva_kh_init(imax, jmax, lx, ly, int(rotated), v0, bx0, bz0, seed, _dp(out)), "kh_init")
    return out


MLP_NPARAM, MLP_NPACK = 1250, 1260


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def cpaw_init(imax, jmax, lx=1.0, gamma=4.0 / 3.0, rho0=1.0, p0=0.1, b0=1.0, eta=0.1):
    """Circularly polarised Alfven wave primitives and the Alfven speed (cueva_cpaw_init)."""
    out = np.zeros((14, jmax + 13, imax + 13))
    va = ctypes.c_double()
    check(load().cueva_cpaw_init(imax, jmax, lx, gamma, rho0, p0, b0, eta, _dp(out), ctypes.byref(va)),
          "cpaw_init")
    return out, va.value


class Mlp:
    """Device copy of the learned-recovery network with its trainer (cueva_mlp)."""

    def __init__(self, packed=None):
        self.lib = load()
        self.m = ctypes.c_void_p()
        check(self.lib.cueva_mlp_create(ctypes.byref(self.m)), "mlp_create")
        if packed is not None:
            self.set_params(packed)

    def close(self):
        if self.m:
            self.lib.cueva_mlp_destroy(self.m)
            self.m = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, packed):
        packed = np.ascontiguousarray(packed, dtype=np.float32)
        assert packed.shape == (MLP_NPACK,)
        check(self.lib.cueva_mlp_set_params(self.m, _fp(packed)), "mlp_set_params")

    def params(self):
        out = np.zeros(MLP_NPACK, dtype=np.float32)
        check(self.lib.cueva_mlp_get_params(self.m, _fp(out)), "mlp_get_params")
        return out

    def set_data(self, x, y):
        self._x = np.ascontiguousarray(x, dtype=np.float32)
        self._y = np.ascontiguousarray(y, dtype=np.float32)
        assert self._x.shape[1:] == (3,) and self._y.shape == (self._x.shape[0], 2)
        check(self.lib.cueva_mlp_set_data(self.m, _fp(self._x), _fp(self._y), self._x.shape[0]), "mlp_set_data")

    def grad(self, off, cnt):
        g = np.zeros(MLP_NPARAM, dtype=np.float32)
        loss = ctypes.c_double()
        check(self.lib.cueva_mlp_grad(self.m, int(off), int(cnt), _fp(g), ctypes.byref(loss)), "mlp_grad")
        return g, loss.value

    def train(self, batch, steps, lr):
        loss = ctypes.c_double()
        check(self.lib.cueva_mlp_train(self.m, int(batch), int(steps), float(lr), ctypes.byref(loss)), "mlp_train")
        return loss.value

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], 2), dtype=np.float32)
        check(self.lib.cueva_mlp_forward(self.m, _fp(x), x.shape[0], _fp(y)), "mlp_forward")
        return y


class Solver:
    """One device-resident grid (cueva_ctx)."""

    def __init__(self, params):
        self.lib = load()
        self.p = params
        self.ctx = ctypes.c_void_p()
        check(self.lib.cueva_create(ctypes.byref(params), ctypes.byref(self.ctx)), "create")
        self.shape = plane_shape(params)

    def close(self):
        if self.ctx:
            self.lib.cueva_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- data
    def upload(self, prims):
        prims = np.ascontiguousarray(prims, dtype=np.float64)
        assert prims.shape == (14,) + self.shape, prims.shape
        check(self.lib.cueva_upload_prims(self.ctx, _dp(prims)), "upload_prims")

    def download(self, cons=True):
        prims = np.zeros((14,) + self.shape)
        c5 = np.zeros((5,) + self.shape) if cons else None
        check(self.lib.cueva_download(self.ctx, _dp(prims), _dp(c5) if cons else None), "download")
        return (prims, c5) if cons else prims

    def device_bytes(self):
        b = ctypes.c_int64()
        check(self.lib.cueva_device_bytes(self.ctx, ctypes.byref(b)), "device_bytes")
        return b.value

    def sync(self):
        check(self.lib.cueva_sync(self.ctx), "sync")

    def upload_cons(self, cons):
        cons = np.ascontiguousarray(cons, dtype=np.float64)
        assert cons.shape == (5,) + self.shape, cons.shape
        check(self.lib.cueva_upload_cons(self.ctx, _dp(cons)), "upload_cons")

    def plane(self, group, index):
        out = np.zeros(self.shape)
        check(self.lib.cueva_download_plane(self.ctx, group, index, _dp(out)), "download_plane")
        return out

    def plane_range(self, group, index):
        """(min, max) of one device plane over the interior."""
        out = np.zeros(2)
        check(self.lib.cueva_plane_range(self.ctx, group, index, _dp(out)), "plane_range")
        return float(out[0]), float(out[1])

    # -- timing
    def set_profiling(self, on):
        check(self.lib.cueva_set_profiling(self.ctx, int(bool(on))), "set_profiling")

    def kernel_times(self):
        ms = (ctypes.c_double * 6)()
        n = (ctypes.c_int64 * 6)()
        check(self.lib.cueva_kernel_times(self.ctx, ms, n), "kernel_times")
        return {k: (ms[i], n[i]) for i, k in enumerate(KERNEL_CLASSES)}

    def time_steps(self, n):
        ms = ctypes.c_double()
        check(self.lib.cueva_time_steps(self.ctx, int(n), ctypes.byref(ms)), "time_steps")
        return ms.value

    # -- learned primitive recovery
    def set_c2p_params(self, packed):
        packed = np.ascontiguousarray(packed, dtype=np.float32)
        assert packed.shape == (MLP_NPACK,), packed.shape
        check(self.lib.cueva_set_c2p_params(self.ctx, _fp(packed)), "set_c2p_params")

    # -- diagnostics
    def globals(self):
        """The 37 global diagnostics of 22_writer.f95:818-991 (GLOBAL_NAMES order)."""
        out = np.zeros(len(GLOBAL_NAMES))
        check(self.lib.cueva_reduce_globals(self.ctx, _dp(out)), "reduce_globals")
        return dict(zip(GLOBAL_NAMES, out.tolist()))

    def max_speed(self):
        v = ctypes.c_double()
        check(self.lib.cueva_max_speed(self.ctx, ctypes.byref(v)), "max_speed")
        return v.value

    # -- fused step
    def step(self, n=1):
        check(self.lib.cueva_step(self.ctx, int(n)), "step")

    def step_counters(self):
        c = (ctypes.c_int64 * 3)()
        fb = ctypes.c_int64()
        check(self.lib.cueva_step_counters(self.ctx, c, ctypes.byref(fb)), "step_counters")
        return list(c), fb.value

    def c2p(self):
        """varprimitivecardano alone (l_swap = 1): (status, counters, fallback)
        -- the status is returned, not raised, so that stops can be tested."""
        cnt = (ctypes.c_int32 * 3)()
        fb = ctypes.c_int32()
        rc = self.lib.cueva_c2p(self.ctx, 1, cnt, ctypes.byref(fb))
        return rc, list(cnt), fb.value

    def step_status(self, n=1):
        """Like step() but returns the status instead of raising."""
        return int(self.lib.cueva_step(self.ctx, int(n)))

    def last_error(self):
        """(status, (h, l, i, j)) of the first failure (0 and zeros if none)."""
        w = (ctypes.c_int32 * 4)()
        rc = self.lib.cueva_last_error(self.ctx, w)
        return rc, tuple(w)

    def set_step_index(self, h):
        check(self.lib.cueva_set_step_index(self.ctx, int(h)), "set_step_index")

    def fixup_records(self, kind):
        """Fix-up records of one kind (0 rho, 1 p, 2 drift) in the reference's
        order: (hij int32 [n, 3], values [n, 2])."""
        n = ctypes.c_int64()
        check(self.lib.cueva_fixup_records(self.ctx, int(kind), None, None, 0, ctypes.byref(n)), "fixup_records")
        hij = np.zeros((max(n.value, 1), 3), dtype=np.int32)
        vals = np.zeros((max(n.value, 1), 2))
        check(self.lib.cueva_fixup_records(self.ctx, int(kind), hij.ctypes.data_as(_P(ctypes.c_int32)), _dp(vals),
                                           n.value, ctypes.byref(n)), "fixup_records")
        return hij[:n.value], vals[:n.value]

    # -- per-subroutine entry points, in the reference's call order
    def step_by_routines(self):
        """One step as the reference driver sequences it (main_cueva_code.f95:53-328);
        returns the fix-up counters of the primitive recovery."""
        L = self.lib
        x = self.ctx
        lmax = int(self.p.order)  # stages = order
        check(L.cueva_boundary(x, BND_PRIM, 1), "boundary")
        if self.p.primitive_rec:  # main_cueva_code.f95:170-175
            check(L.cueva_recons_var_primitive(x), "recons_var_primitive")
            check(L.cueva_boundary(x, BND_PRIM, 1), "boundary")
        check(L.cueva_startconvar(x), "startconvar")
        check(L.cueva_cleanvarint(x), "cleanvarint")
        check(L.cueva_swap1(x), "swap1")
        for l in range(1, lmax + 1):
            for m in range(1, l):
                check(L.cueva_sumint(x, l, m), "sumint")
            check(L.cueva_varint(x, l), "varint")
            check(L.cueva_electricvarint(x, l), "electricvarint")
            if self.p.rec_prim:  # main_cueva_code.f95:220-228
                check(L.cueva_c2p_stage(x, l), "varprimitivecardano (l_swap = 0)")
                check(L.cueva_interconvar(x, l), "interconvar")
            check(L.cueva_boundary(x, BND_CONS, l), "boundary_conserved")
            check(L.cueva_boundary(x, BND_ELEC, l), "boundary_electric")
            check(L.cueva_flowinit(x, l), "flowinit")
            check(L.cueva_boundary(x, BND_FLUX, l), "boundary_flux_conserved")
            check(L.cueva_flux(x, l), "flux")
            check(L.cueva_fluxandsource(x, l), "fluxandsource")
            check(L.cueva_sumfin(x, l), "sumfin")
        check(L.cueva_varfin(x), "varfin")
        check(L.cueva_electricfield(x), "electricfield")
        cnt = (ctypes.c_int32 * 3)()
        fb = ctypes.c_int32()
        check(L.cueva_c2p(x, 2 if self.p.rec_prim else 1, cnt, ctypes.byref(fb)), "varprimitivecardano")
        check(L.cueva_boundary(x, BND_PRIM, 1), "boundary")
        return list(cnt)
