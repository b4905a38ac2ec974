"""Restart files in the reference's format (22_writer.f95:1116-1172 writes
them, 03_restart_primitive.f95:50-61 reads them), on the CPU.

The Fortran reader and writer of the GPU driver (fortran/cueva_io.f95),
built into a small copy program, must give back the reference's own restart
file byte for byte; the Python reader must give back the doubles of the
reference's binary dump of the same step.  The reference's file comes from
its own build (oracle/_ref, run with its lector) where that is present; the
SHA-256 of the files it writes is pinned in tests/golden/*_restart.json,
which the GPU test of the driver checks its own files against.
"""
import hashlib
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import cueva

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FC = os.environ.get("FC_CUEVA", "/opt/rocm/lib/llvm/bin/flang")
FDIR = os.path.join(ROOT, "mhd-kh-instability_amd", "fortran")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _copy_program(tmp_path):
    if not os.path.exists(FC):
        pytest.skip("no Fortran compiler")
    exe = str(tmp_path / "restart_copy")
    subprocess.run([FC, "-O2", "-J", str(tmp_path), os.path.join(FDIR, "cueva_io.f95"),
                    os.path.join(ROOT, "tests", "native", "restart_copy.f95"), "-o", exe], check=True,
                   capture_output=True)
    return exe


def _reference_restart(tmp_path, case, nsteps):
    """The reference's restart file after nsteps (its lector with h = 0) and
    its binary dump of the same step."""
    exe = os.path.join(ROOT, "oracle", "_ref", case, "cueva_ref")
    if not os.path.exists(exe):
        pytest.skip("reference build absent")
    meta = json.load(open(os.path.join(GOLDEN, case + ".json")))
    ni, nj = meta["native_grid"]
    p, prims = cueva.kh_problem(ni, nj, sigma=meta["sigma"], flux_solver=meta["flux_solver"], rec_mp=meta["rec_mp"],
                                order=meta["order"], rotated=True, seed=meta["seed"], **meta.get("init", {}))
    run = tmp_path / "ref"
    for d in ("data_restart", "data_print", "data_local"):
        (run / d).mkdir(parents=True)
    prims.astype("<f8").tofile(str(run / "in.bin"))
    res = subprocess.run([exe, "in.bin", str(nsteps), "out", str(nsteps), "-lector", "0"], cwd=str(run),
                         capture_output=True, text=True, timeout=120, env=dict(os.environ, OMP_NUM_THREADS="1"))
    assert res.returncode == 0, res.stdout + res.stderr
    text = (run / "data_restart" / "all_variables_0000.dat").read_bytes()
    dump = np.frombuffer((run / ("out_h%05d.bin" % nsteps)).read_bytes()[12:], dtype="<f8")
    return p, text, dump


@pytest.mark.parametrize("name", ["kh_hllc_mp5_32x64_restart", "kh_hlle_mp7_bz_32x64_restart"])
def test_reader_and_writer_give_back_the_reference_file(tmp_path, name):
    fix = json.load(open(os.path.join(GOLDEN, name + ".json")))
    n1 = fix["steps"][0]
    p, text, dump = _reference_restart(tmp_path, fix["case"], n1)
    assert hashlib.sha256(text).hexdigest() == fix["restart_sha256"][str(n1)]
    exe = _copy_program(tmp_path)
    src, dst = tmp_path / "in.dat", tmp_path / "out.dat"
    src.write_bytes(text)
    res = subprocess.run([exe, str(p.imax), str(p.jmax), str(src), str(dst)], capture_output=True, text=True,
                         timeout=60)
    assert res.returncode == 0, res.stdout + res.stderr
    assert dst.read_bytes() == text
    assert res.stdout.split() == ["data_restart/all_variables_0007.dat", "data_restart/all_variables_0123.dat",
                                  "data_restart/all_variables_1234.dat"]
    assert res.stdout.split() == [cueva.restart_file_name(n) for n in (7, 123, 1234)]
    # like lector, the writer does not overwrite an existing file (and, unlike
    # a failed OPEN with status='new' under flang's runtime, leaves it there)
    again = subprocess.run([exe, str(p.imax), str(p.jmax), str(src), str(dst)], capture_output=True, text=True,
                           timeout=60)
    assert again.returncode != 0 and dst.read_bytes() == text
    # the Python reader gives back the reference's doubles, ghosts included
    planes = cueva.read_restart_text(text.decode(), p.imax, p.jmax)
    n = planes.size
    assert np.array_equal(planes.ravel(), dump[:n])


def test_python_reader_follows_list_directed_input():
    imax = jmax = 0  # 13 x 13 points
    rng = np.random.default_rng(7)
    want = rng.standard_normal((14, 13, 13))
    lines = []
    for i in range(13):
        for j in range(13):
            v = want[:, j, i]
            # a record over two lines, comma separators, a D exponent and a
            # repeat count; trailing values on a record's last line are skipped
            first = ", ".join(repr(float(x)) for x in v[:5])
            rest = " ".join(("%.17E" % x).replace("E", "D") for x in v[5:12])
            lines += [first, rest + " " + repr(float(v[12])) + " " + repr(float(v[13])) + " 99.0"]
    got = cueva.read_restart_text("\n".join(lines), imax, jmax)
    assert np.array_equal(got, want)
    rep = "\n".join(["3*0.5 11*1.0"] * 169)
    got = cueva.read_restart_text(rep, 0, 0)
    assert np.all(got[:3] == 0.5) and np.all(got[3:] == 1.0)
    with pytest.raises(ValueError):
        cueva.read_restart_text("1.0 2.0", 0, 0)
