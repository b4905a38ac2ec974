"""cueva_reduce_globals_noalloc against cueva_reduce_globals.

The allocation-free variant takes device planes in the host layout and
writes the 37 diagnostics to device memory, with its block partial sums in
the library's fixed device array.  Two calls in a row, the second on a
smaller grid (the array keeps the first call's partials), must each equal
the context entry point bit for bit (same kernels, same summation order).
"""
import ctypes

import numpy as np
import pytest

import cueva

pytestmark = pytest.mark.gpu


class _Hip:
    """Device memory for the test through the HIP runtime the library itself
    links (loaded with it; a second HIP runtime in the process, such as
    torch's, would not see the device once the first one holds it)."""

    def __init__(self):
        cueva.load()
        self.rt = ctypes.CDLL("libamdhip64.so.7")

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        ptr = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(ptr), ctypes.c_size_t(a.nbytes)) == 0
        assert self.rt.hipMemcpy(ptr, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1) == 0
        return ptr

    def get(self, ptr, n):
        out = np.empty(n)
        assert self.rt.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ptr, ctypes.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self, *ptrs):
        for p in ptrs:
            self.rt.hipFree(p)


def _case(n_i, n_j, steps):
    p, prims = cueva.kh_problem(n_i, n_j, sigma=1.0e6, flux_solver=cueva.HLLC)
    with cueva.Solver(p) as s:
        s.upload(prims)
        s.step(steps)
        s.step_counters()
        want = np.array(list(s.globals().values()))
        pr, co = s.download()
    return p, pr, co, want


def test_noalloc_globals_match_context_entry_point():
    hip = _Hip()
    lib = cueva.load()
    for n_i, n_j, steps in ((128, 256, 3), (32, 64, 2)):  # the second call on a smaller grid
        print("noalloc globals", n_i, n_j, flush=True)
        p, pr, co, want = _case(n_i, n_j, steps)
        d_pr, d_co = hip.put(pr), hip.put(co)
        d_out = hip.put(np.full(37, np.nan))
        rc = lib.cueva_reduce_globals_noalloc(ctypes.byref(p), d_pr, d_co, d_out, None)
        assert rc == 0
        assert hip.rt.hipDeviceSynchronize() == 0
        got = hip.get(d_out, 37)
        hip.free(d_pr, d_co, d_out)
        assert np.array_equal(got.view("<u8"), want.view("<u8")), np.abs(got - want).max()


def test_noalloc_rejects_bad_arguments():
    p = cueva.default_params()
    lib = cueva.load()
    assert lib.cueva_reduce_globals_noalloc(ctypes.byref(p), None, None, None, None) == 1
    p.imax = 4  # below the smallest grid
    one = ctypes.c_void_p(16)
    assert lib.cueva_reduce_globals_noalloc(ctypes.byref(p), one, one, one, None) == 1
