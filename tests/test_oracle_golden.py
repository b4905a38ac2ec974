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
    ni, nj = meta["native_grid"]
    return cueva.kh_problem(ni, nj, sigma=meta["sigma"], flux_solver=meta["flux_solver"],
                            rec_mp=meta["rec_mp"], order=meta["order"], rotated=True, seed=meta["seed"],
                            **meta.get("init", {}))


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def interior(prims, cons, p):
    both = np.concatenate([prims, cons])
    return both[:, 6:6 + p.jmax + 1, 6:6 + p.imax + 1]


SMALL = sorted(os.path.basename(f)[:-5] for f in glob.glob(os.path.join(GOLD, "*_32x64.json")))
LARGE = sorted(os.path.basename(f)[:-5] for f in glob.glob(os.path.join(GOLD, "*_128x256.json")))


def run_case(name):
    meta, arrays = load_case(name)
    p, prims = problem(meta)
    assert digest(prims) == meta["init_sha256"], "initial-data generator changed"
    o = Oracle(p, libm="c")  # the C library, as the reference build
    o.upload(prims)
    done = 0
    for d in sorted(int(k) for k in meta["dumps"]):
        o.step(d - done)
        done = d
        pr, co = o.download()
        inner = interior(pr, co, p)
        ref = meta["dumps"][str(d)]
        assert o.counters() == ref["counters"], (name, d)
        bad = [f for k, f in enumerate(FIELDS) if digest(inner[k]) != ref["sha256"][f]]
        if "h%d" % d in arrays:
            a = arrays["h%d" % d]
            diff = {f: float(np.abs(inner[k] - a[k]).max()) for k, f in enumerate(FIELDS) if f in bad}
        else:
            diff = {}
        assert not bad, (name, d, bad, diff)
    if "globals_final" in meta:
        # the reference's own lector output after the last step (22_writer.f95:818-991)
        assert done == meta["nsteps"]
        got = dict(zip(cueva.GLOBAL_NAMES, o.globals().tolist()))
        bad = {n: (got[n], v) for n, v in meta["globals_final"].items() if got[n] != v}
        assert not bad, (name, bad)
    if meta["expected_stop_step"]:
        with pytest.raises(RuntimeError, match="status 22"):
            o.step(meta["expected_stop_step"] - done)


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
