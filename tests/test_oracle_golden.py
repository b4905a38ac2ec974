"""The CPU restatement (oracle/) reproduces the reference bit for bit.

The fixtures in tests/golden/ come from the reference compiled from its own
sources (oracle/refbuild/build_ref.sh) and run single-threaded on seeded
Kelvin-Helmholtz data (tests/golden/make_golden.py).  Every field of every
dump must have the same SHA-256 as the reference's, and the fix-up counters
must match, and the 37 global diagnostics must equal what the reference's
own lector wrote after the last step.  This pins every oracle routine on
the path: MP5/MP7/MP9, HLL/HLLE/HLLC, the IMEX SSP2(2,2,2)/SSP3(4,3,3) stage
sums, the E solve, the boundary maps, the Cardano recovery and the global
reductions.
"""
import glob
import hashlib
import json
import os

import numpy as np
import pytest

import crafted
import cueva
from oracle.binding import Oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
FIELDS = list(cueva.PRIM_NAMES) + list(cueva.CONS_NAMES)


def load_case(name):
    meta = json.load(open(os.path.join(GOLD, name + ".json")))
    npz = os.path.join(GOLD, name + ".npz")
    arrays = dict(np.load(npz)) if os.path.exists(npz) else {}
    return meta, arrays


def problem(meta):
    return crafted.problem(meta)


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def interior(prims, cons, p):
    both = np.concatenate([prims, cons])
    return both[:, 6:6 + p.jmax + 1, 6:6 + p.imax + 1]


def same_values(a, b):
    """Equal as the reference's shortest round-trip text reads back (NaN as NaN)."""
    a, b = np.asarray(a, dtype=float).reshape(-1), np.asarray(b, dtype=float).reshape(-1)
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


def check_face_events(o, arrays, name):
    """The oracle's face fallbacks and density fixes are the faces the
    reference printed, in its order: kind (fallback from an x face / a y face
    only; the density fix of y_L, whose message says test_rho_L, or of
    y_R / x_L / x_R, whose messages all say test_rho_R), face and value."""
    kl, vals = o.face_event_list()
    kind = np.where(kl[:, 0] <= 2, kl[:, 0], 3)
    assert np.array_equal(kind, arrays["face_event_kind"]), name
    assert np.array_equal(kl[:, 2:4], arrays["face_event_ij"]), name
    assert same_values(vals, arrays["face_event_val"]), name


SMALL = sorted(os.path.basename(f)[:-5] for f in glob.glob(os.path.join(GOLD, "kh_*_32x64.json")))
LARGE = sorted(os.path.basename(f)[:-5] for f in glob.glob(os.path.join(GOLD, "*_128x256.json")))


def same_or_both_nan(x, y):
    """Bit for bit, a NaN equal to a NaN."""
    x, y = np.asarray(x, dtype="<f8"), np.asarray(y, dtype="<f8")
    return bool(np.all((x.view("<u8") == y.view("<u8")) | (np.isnan(x) & np.isnan(y))))


def run_case(name):
    meta, arrays = load_case(name)
    p, prims = problem(meta)
    assert digest(prims) == meta["init_sha256"], "initial-data generator changed"
    o = Oracle(p, libm="c")  # the C library, as the reference build
    o.upload(prims)
    done = 0
    fixups = {k: ([], []) for k in range(3)}  # per kind: (h, i, j) and values of every step
    for d in sorted(int(k) for k in meta["dumps"]):
        while done < d:
            o.step(1)
            done += 1
            for kind in range(3):
                ij, vals = o.fixups(kind)
                fixups[kind][0].extend((done, int(i), int(j)) for i, j in ij)
                fixups[kind][1].extend(vals.tolist())
        pr, co = o.download()
        inner = interior(pr, co, p)
        ref = meta["dumps"][str(d)]
        assert o.counters() == ref["counters"], (name, d)
        bad = [f for k, f in enumerate(FIELDS) if digest(inner[k]) != ref["sha256"][f]]
        if "h%d" % d in arrays:
            a = arrays["h%d" % d]
            # where the reference holds NaN (REC_PRIM = 1 on these grids), the
            # same points must be NaN: the sign and payload bits of a NaN follow
            # the operand order its compiler chose, which nothing here pins
            bad = [f for k, f in enumerate(FIELDS) if f in bad and not same_or_both_nan(inner[k], a[k])]
            diff = {f: float(np.abs(inner[k] - a[k]).max()) for k, f in enumerate(FIELDS) if f in bad}
        else:
            diff = {}
        assert not bad, (name, d, bad, diff)
    if "fixups" in meta and done == meta["nsteps"]:
        # the records of the reference's fix-up files, all steps (12_varprimitive.f95:544,601,663)
        for kind, fn in enumerate(("rho", "p", "v")):
            assert len(fixups[kind][0]) == meta["fixups"][fn]["n"], (name, fn)
            if meta["fixups"][fn]["n"]:
                assert np.array_equal(np.array(fixups[kind][0]), arrays["fixup_%s_hij" % fn]), (name, fn)
                assert same_values(np.array(fixups[kind][1]), arrays["fixup_%s_val" % fn]), (name, fn)
    if meta.get("face_events"):
        check_face_events(o, arrays, name)
    if "globals_final" in meta:
        # the reference's own lector output after the last step (22_writer.f95:818-991)
        assert done == meta["nsteps"]
        got = dict(zip(cueva.GLOBAL_NAMES, o.globals().tolist()))
        bad = {n: (got[n], v) for n, v in meta["globals_final"].items()
               if got[n] != v and not (np.isnan(got[n]) and np.isnan(v))}
        assert not bad, (name, bad)
    if meta["expected_stop_step"]:
        stop = meta.get("stop", {"code": 22})
        with pytest.raises(RuntimeError, match="status %d" % stop["code"]):
            o.step(meta["expected_stop_step"] - done)
        if "i" in stop:  # the reference's (h, i, j, l) of its stop message
            h, l, i, j = o.last_stop()
            assert (done + h, l, i, j) == (stop["h"], stop["l"], stop["i"], stop["j"]), (name, o.last_stop())


@pytest.mark.parametrize("name", SMALL)
def test_oracle_bit_exact_small(name):
    run_case(name)


@pytest.mark.slow
@pytest.mark.parametrize("name", LARGE)
def test_oracle_bit_exact_128x256_100_steps(name):
    run_case(name)


def test_fixture_set_complete():
    assert len(SMALL) >= 5 and len(LARGE) >= 2, (SMALL, LARGE)


def test_oracle_stops_where_the_reference_stops():
    """cfg1 (128x256, HLL, sigma = 1e3) cannot run its 200 steps: the
    reference's Cardano recovery finds det < 0 at step 135
    (12_varprimitive.f95:374-425); the oracle stops at the same step with
    status 22 and runs every step before it."""
    stop = json.load(open(os.path.join(GOLD, "kh_cfg1_hll_mp5_128x256_stop.json")))
    meta, _ = load_case(stop["case"])
    p, prims = problem(meta)
    o = Oracle(p)
    o.upload(prims)
    o.step(stop["stop_step"] - 1)
    with pytest.raises(RuntimeError, match="status 22"):
        o.step(1)


STOP_CASES = sorted(os.path.basename(f)[:-5] for f in glob.glob(os.path.join(GOLD, "stop_*.json")))


@pytest.mark.parametrize("name", STOP_CASES)
def test_oracle_stop_cases(name):
    """One step of crafted data on which the reference stops
    (tests/crafted.py: superluminal face states).  The oracle prints-equivalent
    face events are the reference's, in order; it stops with the reference's
    status (and, where the reference names it, at its cell and stage); the
    right-hand sides of every stage the reference completed, and the face
    fluxes of the last completed face pass, have the reference's SHA-256."""
    meta, arrays = load_case(name)
    p, prims = problem(meta)
    assert digest(prims) == meta["init_sha256"]
    o = Oracle(p, libm="c")
    o.upload(prims)
    st = o.step_status(1)
    assert st == meta["stop"]["code"], (st, meta["stop"])
    h, l, i, j = o.last_stop()
    if "i" in meta["stop"]:
        assert (h, l, i, j) == (meta["stop"]["h"], meta["stop"]["l"], meta["stop"]["i"], meta["stop"]["j"])
    check_face_events(o, arrays, name)
    stages = meta["stages"]
    assert stages
    im, jm = p.imax, p.jmax
    for ls, ref in stages.items():
        for q in range(17):
            out = o.rhs(int(ls) - 1, q)
            assert digest(out[6:6 + jm + 1, 6:6 + im + 1]) == ref["rhs_sha256"][q], (name, ls, q)
    last_faces = str(max(int(x) for x in stages))
    if meta["stop"]["code"] >= 20 and last_faces == str(2 if meta["order"] == 2 else 4):
        ref = stages[last_faces]
        for q in range(28):
            out = o.face(q // 14, q % 14)
            assert digest(out[5:5 + jm + 2, 5:5 + im + 2]) == ref["faces_sha256"][q], (name, q)
