"""The GPU's failure and fix-up branches against the reference's own runs.

* Primitive recovery alone (cueva_c2p) on the crafted state grid of
  tests/golden/c2p_states.*: every result bit, the three counters and the
  fix-up records (cell, order and both printed values) equal the
  reference's; eps1 = 0 in a context without an HLL+MPx face solve, and
  1e-6 after one, as in the reference process; each stopping grid returns
  the reference's status at the cell where it stops (the first of two in its
  loop order).
* One step of crafted data that stops the reference (tests/crafted.py,
  spin: superluminal face states): the per-subroutine sequence gives the
  reference's stage right-hand sides and face fluxes bit for bit, with 1612
  (HLL) / 806 (HLLC) superluminal face fallbacks among them, and stops with
  its status at its step, stage and cell; the fused step stops there too,
  runs no later step of the call, and leaves the state the failing launch
  worked on.
The oracle is pinned to the same fixtures in tests/test_c2p_states.py and
tests/test_oracle_golden.py.
"""
import json
import os

import numpy as np
import pytest

import crafted
import cueva
from oracle.binding import Oracle
from test_c2p_states import KIND, load, planes, records_match, result_of, same_bits, set_cell


def same_results(a, b):
    """Bit for bit, except that a NaN equals a NaN: the sign bit of a NaN
    differs between the host's and the GPU's arithmetic (x86 makes negative
    quiet NaNs), and the reference prints and reads NaN without it."""
    a, b = np.asarray(a, dtype="<f8"), np.asarray(b, dtype="<f8")
    return bool(np.all((a.view("<u8") == b.view("<u8")) | (np.isnan(a) & np.isnan(b))))

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def say(*a):
    print(*a, flush=True)


def solver_for(meta, flux_solver=cueva.HLLC):
    imax, jmax = meta["grid"]
    p = cueva.default_params(imax=imax, jmax=jmax, flux_solver=flux_solver, bc_x=cueva.BC_COPY,
                             bc_y=cueva.BC_PERIODIC, sigma=1.0e6, longx=2.0, longy=1.0, crmath=1)
    return cueva.Solver(p)


def test_gpu_recovery_matches_reference_bitwise():
    meta, arrays = load()
    say("c2p states", meta["grid"])
    prims, cons = planes(meta, arrays)
    with solver_for(meta) as s:
        s.upload(prims)
        s.upload_cons(cons)
        rc, cnt, _ = s.c2p()
        assert rc == 0
        assert cnt == meta["counters"]
        got = s.download(cons=False)
        for kind, name in enumerate(KIND):
            hij, vals = s.fixup_records(kind)
            assert np.all(hij[:, 0] == 1)
            assert records_match(hij[:, 1:], vals, arrays, name), name
    res = result_of(got, meta)
    assert same_results(res, arrays["result"]), int((res.view("u8") != arrays["result"].view("u8")).sum())


def test_gpu_recovery_clamp_after_hll_face_solve():
    """An HLL+MPx face solve in the context sets eps1 = 1e-6, as hll_flow_mpx
    sets the reference's module variable."""
    meta, arrays = load()
    prims, cons = planes(meta, arrays)
    with solver_for(meta, cueva.HLL) as s:
        s.upload(prims)
        assert s.lib.cueva_flux(s.ctx, 1) == 0  # on the uploaded state; only its side effect matters
        s.upload(prims)
        s.upload_cons(cons)
        rc, cnt, _ = s.c2p()
        assert rc == 0 and cnt == meta["counters_eps1"]
        got = s.download(cons=False)
    assert same_results(result_of(got, meta), arrays["result_eps1"])


@pytest.mark.parametrize("k", range(4))
def test_gpu_recovery_stops_where_the_reference_stops(k):
    meta, arrays = load()
    stop = meta["stops"][k]
    say("c2p stop", k, stop["message"])
    prims, cons = planes(meta, arrays)
    for c in stop["cells"]:
        set_cell(prims, cons, c["i"], c["j"], np.array(c["state"]))
    with solver_for(meta) as s:
        s.upload(prims)
        s.upload_cons(cons)
        rc, _, _ = s.c2p()
        code, where = s.last_error()
        # sticky until a new state is uploaded
        rc2, _, _ = s.c2p()
    first = stop["cells"][0]
    assert rc == code == rc2 == stop["code"]
    assert where == (1, 2, first["i"], first["j"])


def stop_case(name):
    meta = json.load(open(os.path.join(GOLD, name + ".json")))
    return meta, crafted.problem(meta)


def digest(a):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


@pytest.mark.parametrize("name", ["stop_hll_spin_32x64", "stop_hllc_spin_32x64"])
def test_gpu_routines_through_superluminal_faces(name):
    meta, (p, prims) = stop_case(name)
    say("stop case, routines", name)
    im, jm = p.imax, p.jmax
    L = cueva.load()
    with cueva.Solver(p) as s:
        x = s.ctx
        s.upload(prims)
        for call in (lambda: L.cueva_boundary(x, cueva.abi.BND_PRIM, 1), lambda: L.cueva_startconvar(x),
                     lambda: L.cueva_cleanvarint(x), lambda: L.cueva_swap1(x)):
            assert call() == 0
        for l in (1, 2):
            for m in range(1, l):
                assert L.cueva_sumint(x, l, m) == 0
            for call in (L.cueva_varint, L.cueva_electricvarint, L.cueva_flowinit, L.cueva_flux,
                         L.cueva_fluxandsource, L.cueva_sumfin):
                assert call(x, l) == 0
            ref = meta["stages"].get(str(l))
            if ref is None:
                break
            faces = [s.plane(3 + q // 14, q % 14)[5:5 + jm + 2, 5:5 + im + 2] for q in range(28)]
            rhs = [s.plane(4 + l, q)[6:6 + jm + 1, 6:6 + im + 1] for q in range(17)]
            bad_f = [q for q in range(28) if digest(faces[q]) != ref["faces_sha256"][q]]
            bad_r = [q for q in range(17) if digest(rhs[q]) != ref["rhs_sha256"][q]]
            assert not bad_f and not bad_r, (l, bad_f, bad_r)
        assert L.cueva_varfin(x) == 0 and L.cueva_electricfield(x) == 0
        rc, _, _ = s.c2p()
        code, where = s.last_error()
    assert rc == code == meta["stop"]["code"]
    if "i" in meta["stop"]:
        st = meta["stop"]
        assert where == (st["h"], st["l"], st["i"], st["j"])
    else:  # the reference names no cell for get_hllc_vel: the oracle's, on the step and stage it stopped in
        o = Oracle(p)
        o.upload(prims)
        assert o.step_status(1) == code
        assert where == o.last_stop()
        assert where[1] == max(int(k) for k in meta["stages"]) + 1


@pytest.mark.parametrize("name", ["stop_hll_spin_32x64", "stop_hllc_spin_32x64"])
def test_gpu_fused_step_stops_at_the_failing_step(name):
    """cueva_step(3) runs only the failing first step: status and (h, l, i, j)
    as the reference's stop (and the oracle's); the primitives are then those
    at the start of the failing step."""
    meta, (p, prims) = stop_case(name)
    say("stop case, fused", name)
    with cueva.Solver(p) as s:
        s.upload(prims)
        rc = s.step_status(3)
        code, where = s.last_error()
        assert s.step_status(1) == code  # sticky
        g_p, g_c = s.download()
    o = Oracle(p)
    o.upload(prims)
    assert o.step_status(3) == code == rc == meta["stop"]["code"]
    assert where == o.last_stop()
    sl = (slice(None), slice(6, 6 + p.jmax + 1), slice(6, 6 + p.imax + 1))
    assert same_results(g_p[sl], prims[sl])
