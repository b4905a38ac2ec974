"""The Fortran host driver (mhd-kh-instability_amd/fortran/): the reference's
time loop calling reference-named shims over the iso_c_binding interface.

It must produce the same fields, bit for bit, as the fused cueva_step on the
same data, and its per-hglb global diagnostics must be cueva_reduce_globals'.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import cueva

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "mhd-kh-instability_amd", "lib", "cueva_gpu_driver")


def test_driver_built_and_checks_arguments():
    if not os.path.exists(DRIVER):
        pytest.skip("Fortran driver not built (needs flang)")
    res = subprocess.run([DRIVER], capture_output=True, text=True, timeout=60)
    assert res.returncode == 2 and "usage" in res.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("fs,rec,order,kw", [(cueva.HLLC, 0, 2, {}), (cueva.HLLE, 1, 2, {"bz0": 0.1}),
                                             (cueva.HLL, 2, 4, {}), (cueva.HLL_MCL, 0, 2, {"bz0": 0.1}),
                                             (cueva.LAX, 0, 2, {"bz0": 0.1}), (cueva.HLLC_MCL, 0, 2, {"bz0": 0.1}),
                                             (cueva.HLLE_MCL, 0, 2, {"bz0": 0.1}), (cueva.HLLC, 0, 2, {"primitive_rec": 1}),
                                             (cueva.HLLC_MCL, 0, 2, {"slop": 3}),
                                             (cueva.HLL_MCL, 0, 2, {"primitive_rec": 1, "limit": 1, "slop": 1}),
                                             (cueva.HLL, 0, 1, {"mirk": 1}), (cueva.HLLC, 0, 2, {"mirk": 2}),
                                             (cueva.LAX, 0, 2, {"rec_prim": 1, "mirk": 2, "bz0": 0.1}),
                                             (cueva.LAX, 0, 2, {"rec_prim": 1, "mirk": 2})],
                         ids=["hllc_mp5", "hlle_mp7", "hll_mp9_ssp3", "hll_mcl", "lax", "hllc_mcl", "hlle_mcl",
                              "hllc_mp5_primrec", "hllc_mcl_superbee", "hll_mcl_minmod_primrec_minmod", "hll_mirk1",
                              "hllc_mirk2", "lax_recprim_mirk2_bz", "lax_recprim_mirk2"])
def test_driver_matches_fused_step(tmp_path, fs, rec, order, kw):
    if not os.path.exists(DRIVER):
        pytest.skip("Fortran driver not built (needs flang)")
    ni, nj, steps, sigma = 32, 64, 3, 1.0e6
    bz0 = kw.get("bz0", 0.0)
    prefix = str(tmp_path / "run")
    print("driver", fs, rec, order, flush=True)
    args = [DRIVER, str(ni), str(nj), str(steps), str(fs), str(rec), str(order), repr(sigma), "1", "0.5", "0.01",
            repr(bz0), "20261202", "1", prefix]
    if any(kw.get(k) for k in ("primitive_rec", "slop", "limit", "mirk", "rec_prim")):
        # HLOC RESTART_IN HINI PRIMITIVE_REC SLOP LIMIT MIRK REC_PRIM
        args += ["0", "-", "1"] + [str(kw.get(k, d)) for k, d in (("primitive_rec", 0), ("slop", 2), ("limit", 2),
                                                                   ("mirk", 0), ("rec_prim", 0))]
    res = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    p, prims = cueva.kh_problem(ni, nj, sigma=sigma, flux_solver=fs, rec_mp=rec, order=order, rotated=True, **kw)
    raw = open(prefix + ".bin", "rb").read()
    cnt = np.frombuffer(raw[:12], dtype="<i4").tolist()
    arr = np.frombuffer(raw[12:], dtype="<f8").reshape(19, p.jmax + 13, p.imax + 13)
    glob = np.loadtxt(prefix + "_globals.dat", ndmin=2)
    with cueva.Solver(p) as s:
        s.upload(prims)
        want_glob = []
        for h in range(steps):
            s.step(1)
            c, _ = s.step_counters()
            want_glob.append(list(s.globals().values()))
        pr, co = s.download()
    assert cnt == c
    ours = np.concatenate([pr, co])
    sl = (slice(None), slice(6, 6 + p.jmax + 1), slice(6, 6 + p.imax + 1))
    assert np.array_equal(arr[sl], ours[sl])
    assert glob.shape == (steps, 39)
    assert np.array_equal(glob[:, 2:], np.array(want_glob))
    assert np.array_equal(glob[:, 0], np.arange(1, steps + 1))


@pytest.mark.gpu
def test_driver_stops_where_the_oracle_stops(tmp_path):
    """REC_PRIM = 1 with the IMEX E solve at sigma = 1e6 and a guide field: the
    recovery of the second stage of step 3 finds det < 0 (as the oracle does);
    the driver prints the reference's status and (h, l, i, j) and stops."""
    if not os.path.exists(DRIVER):
        pytest.skip("Fortran driver not built (needs flang)")
    from oracle.binding import Oracle
    p, prims = cueva.kh_problem(32, 64, sigma=1.0e6, flux_solver=cueva.LAX, rotated=True, rec_prim=1, bz0=0.1)
    o = Oracle(p)
    o.upload(prims)
    assert o.step_status(3) == 22
    args = [DRIVER, "32", "64", "3", str(cueva.LAX), "0", "2", repr(1.0e6), "1", "0.5", "0.01", "0.1", "20261202",
            "1", str(tmp_path / "run"), "0", "-", "1", "0", "2", "2", "0", "1"]
    res = subprocess.run(args, capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    assert res.returncode != 0
    assert int(res.stdout.split("returned status")[1].split()[0]) == 22, res.stdout
    assert tuple(int(x) for x in res.stdout.split("--->")[1].split()[:4]) == tuple(o.last_stop()), res.stdout


def _run_driver(cwd, ni, nj, fs, rec, order, sigma, bz0, steps, hloc, restart_in="-", hini=1, hglb=1, primrec=0,
                slop=2, limit=2, mirk=0, rec_prim=0):
    for d in ("data_restart", "data_global"):
        os.makedirs(os.path.join(cwd, d), exist_ok=True)
    args = [DRIVER, str(ni), str(nj), str(steps), str(fs), str(rec), str(order), repr(sigma), "1", "0.5", "0.01",
            repr(bz0), "20261202", str(hglb), "run", str(hloc), restart_in, str(hini), str(primrec), str(slop),
            str(limit), str(mirk), str(rec_prim)]
    res = subprocess.run(args, capture_output=True, text=True, timeout=120, cwd=cwd)
    assert res.returncode == 0, res.stdout + res.stderr
    return open(os.path.join(cwd, "run.bin"), "rb").read()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["kh_hllc_mp5_32x64_restart", "kh_hlle_mp7_bz_32x64_restart"])
def test_driver_restart_files_are_the_reference_s(tmp_path, name):
    """Restart dumps of the driver against the SHA-256 of the reference's own
    (its lector after the same steps), and a run resumed from one of them
    against the uninterrupted run, bit for bit."""
    import hashlib
    import json
    if not os.path.exists(DRIVER):
        pytest.skip("Fortran driver not built (needs flang)")
    fix = json.load(open(os.path.join(ROOT, "tests", "golden", name + ".json")))
    case = json.load(open(os.path.join(ROOT, "tests", "golden", fix["case"] + ".json")))
    n1, n2 = fix["steps"]
    assert n1 % 2 == 0 and n2 % 2 == 0
    ni, nj = case["native_grid"]
    conf = (ni, nj, case["flux_solver"], case["rec_mp"], case["order"], case["sigma"], case["init"].get("bz0", 0.0))
    print("driver restart", name, flush=True)
    whole = str(tmp_path / "whole")
    end_whole = _run_driver(whole, *conf, n1 + n2, hloc=2)
    dump = lambda d, k: open(os.path.join(d, cueva.restart_file_name(k)), "rb").read()  # noqa: E731
    # dumps every 2 steps: number n1/2 is step n1, (n1+n2)/2 the last step
    assert hashlib.sha256(dump(whole, n1 // 2)).hexdigest() == fix["restart_sha256"][str(n1)]
    assert hashlib.sha256(dump(whole, (n1 + n2) // 2)).hexdigest() == fix["restart_sha256"][str(n1 + n2)]
    # resumed from the step-n1 file, numbering steps on from n1 + 1
    resumed = str(tmp_path / "resumed")
    os.makedirs(resumed)
    src = os.path.join(resumed, "all_variables.dat")
    open(src, "wb").write(dump(whole, n1 // 2))
    end_resumed = _run_driver(resumed, *conf, n2, hloc=2, restart_in=src, hini=n1 + 1)
    assert end_resumed == end_whole
    assert dump(resumed, (n1 + n2) // 2) == dump(whole, (n1 + n2) // 2)
    assert not os.path.exists(os.path.join(resumed, cueva.restart_file_name(n1 // 2)))
    # the reference's per-quantity global files: "h t value" records of the
    # same values as the driver's own table
    table = np.loadtxt(os.path.join(whole, "run_globals.dat"), ndmin=2)
    vy = np.loadtxt(os.path.join(whole, "data_global", "2d_data_vy_integral_vol.dat"), ndmin=2)
    k = cueva.GLOBAL_NAMES.index("sum_vy")
    assert vy.shape == (n1 + n2, 3)
    assert np.array_equal(vy[:, 0], table[:, 0]) and np.array_equal(vy[:, 2], table[:, 2 + k])
    assert np.array_equal(vy[:, 1], table[:, 1])
    assert len(os.listdir(os.path.join(whole, "data_global"))) == len(cueva.GLOBAL_NAMES)


def write_restart_text(path, prims):
    """A restart file of host planes in the reference's record order (i
    outer, j inner, 14 values a record, 03_restart_primitive.f95:50-61),
    shortest round-trip decimals, which list-directed input reads back exactly."""
    nj, ni = prims.shape[1:]
    with open(path, "w") as fh:
        for i in range(ni):
            for j in range(nj):
                fh.write(" ".join(repr(float(v)) for v in prims[:, j, i]) + "\n")


@pytest.mark.gpu
@pytest.mark.parametrize("name,steps", [("kh_hll_lowp_32x64", None), ("kh_hllc_lowp_32x64", None),
                                        ("kh_lax_recprim_32x64", 1)])
def test_driver_fixup_files_are_the_reference_s(tmp_path, name, steps):
    """Crafted data whose recovery takes the pressure fix-up every step
    (tests/crafted.py): the driver's fix-up files are the reference's own
    byte for byte (12_varprimitive.f95:33-40,601) and its fields after the
    fixture's steps are the reference's.  With REC_PRIM = 1 every recovery
    of the first step writes the files' headers (one per stage, one after
    varfin)."""
    import hashlib
    import json
    import crafted
    if not os.path.exists(DRIVER):
        pytest.skip("Fortran driver not built (needs flang)")
    meta = json.load(open(os.path.join(ROOT, "tests", "golden", name + ".json")))
    p, prims = crafted.problem(meta)
    src = str(tmp_path / "initial.dat")
    write_restart_text(src, prims)
    assert np.array_equal(cueva.read_restart(src, p.imax, p.jmax), prims)
    ni, nj = meta["native_grid"]
    print("driver fix-up files", name, flush=True)
    init = meta.get("init", {})
    steps = steps or meta["nsteps"]
    raw = _run_driver(str(tmp_path), ni, nj, meta["flux_solver"], meta["rec_mp"], meta["order"], meta["sigma"], 0.0,
                      steps, 0, restart_in=src, hglb=0, rec_prim=init.get("rec_prim", 0), mirk=init.get("mirk", 0))
    for kind, fn in (("rho", "fix-up_densidad.dat"), ("p", "fix-up_presion.dat"),
                     ("v", "fix-up_velocity_drift.dat")):
        got = hashlib.sha256(open(str(tmp_path / fn), "rb").read()).hexdigest()
        assert got == meta["fixups"][kind]["sha256"], (kind, open(str(tmp_path / fn)).read()[:800])
    cnt = np.frombuffer(raw[:12], dtype="<i4").tolist()
    assert cnt == meta["dumps"][str(steps)]["counters"]
    arr = np.frombuffer(raw[12:], dtype="<f8").reshape(19, p.jmax + 13, p.imax + 13)
    inner = arr[:, 6:6 + p.jmax + 1, 6:6 + p.imax + 1]
    fields = list(cueva.PRIM_NAMES) + list(cueva.CONS_NAMES)
    ref = meta["dumps"][str(steps)]["sha256"]
    bad = [f for k, f in enumerate(fields) if hashlib.sha256(np.ascontiguousarray(inner[k]).tobytes()).hexdigest()
           != ref[f]]
    assert not bad, bad
