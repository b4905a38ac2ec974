"""cueva_params from the reference's own parameters module.

A Fortran host that keeps the reference's compile-time parameters module
(parameters.f95:10-334, or one of its parameters_*.f95 variants) fills the
library's parameter block with cueva_params_from_module
(mhd-kh-instability_amd/fortran/cueva_params_ref.f95).  This test builds it
against the reference's own files (in this container only: it reads
/root/reference) with a probe program, and checks every field against the
constants folded from the same file's text (with Fortran's literal kinds:
parameters_TM.f95's Lx carries a single precision exponent).  A variant that
selects a branch the library does not run is refused with every such
constant named: parameters_CE.f95 (flux_rec_mp5 = 1), parameters_REC_old.f95
(BOUND = 3; its MIRK = 2 with order 2 is run).
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/cueva_2D_base"
FC = "/opt/rocm/lib/llvm/bin/flang"
FDIR = os.path.join(ROOT, "mhd-kh-instability_amd", "fortran")
LIB = os.path.join(ROOT, "mhd-kh-instability_amd", "lib")


class _Int(int):
    """A Fortran default integer: integer arithmetic stays integer and
    division truncates toward zero."""

    def _w(op):
        def f(a, b):
            if not isinstance(b, int):
                return NotImplemented
            return _Int(op(int(a), int(b)))
        return f

    __add__ = __radd__ = _w(lambda a, b: a + b)
    __mul__ = __rmul__ = _w(lambda a, b: a * b)
    __sub__ = _w(lambda a, b: a - b)
    __rsub__ = _w(lambda a, b: b - a)
    __truediv__ = _w(lambda a, b: (abs(a) // abs(b)) * (1 if (a < 0) == (b < 0) else -1))
    __rtruediv__ = _w(lambda a, b: (abs(b) // abs(a)) * (1 if (a < 0) == (b < 0) else -1))
    __pow__ = _w(lambda a, b: a ** b)

    def __neg__(self):
        return _Int(-int(self))


def _literal(m):
    """A Fortran numeric literal as a value of its kind: d exponent -> double,
    other real literals -> default (single precision) real, else integer."""
    t = m.group(0)
    if re.search(r"[dD]", t):
        return "D('%s')" % re.sub(r"[dD]", "e", t)
    if re.search(r"[.eE]", t):
        return "S('%s')" % t
    return "I(%s)" % t


def constants(path):
    """name -> value of the `parameter` constants of a parameters module, each
    right-hand side evaluated as Fortran folds it: over the constants it
    names, with literal kinds kept (1.d0 double, 0.1 single, 2 integer),
    mixed operands promoted to the wider kind and the intrinsics these files
    use."""
    import math

    import numpy as np
    text = open(path).read()
    exprs = {}
    for m in re.finditer(r"^\s*(?:integer|doubleprecision)\s*,\s*parameter\s*::\s*(\w+)\s*=\s*([^!\n]+)", text,
                         flags=re.M | re.I):
        exprs[m.group(1).lower()] = m.group(2).strip()
    fns = {"sqrt": np.sqrt, "real": np.float32, "dble": np.float64, "min": min, "max": max, "abs": abs,
           "floor": lambda x: _Int(math.floor(x)), "ceiling": lambda x: _Int(math.ceil(x)),
           "D": np.float64, "S": np.float32, "I": _Int}
    lit = r"(?<![\w.])(?:\d+\.?\d*|\.\d+)(?:[eEdD][+-]?\d+)?"
    cache = {}

    def val(name):
        name = name.lower()
        if name not in cache:
            e = re.sub(lit, _literal, exprs[name])
            e = re.sub(r"(?<![\w.'])([A-Za-z_]\w*)\b(?!\s*[('])", lambda m: "V('%s')" % m.group(1), e)
            cache[name] = eval(e, {"__builtins__": {}}, dict(fns, V=val))
        return cache[name]

    out = {}
    for n in ("imax", "jmax", "order", "flux_solver", "rec_mp", "primitive_rec", "slop", "limit", "BOUND", "MIRK",
              "DIM", "REC_PRIM", "REC_ELEC", "VDRIFT", "flux_rec_mp5", "REC_4VECTOR", "conserved_to_primitive",
              "low_vel", "SLC", "source_rec", "source_rec_mpx", "var_source_rec"):
        out[n] = int(val(n))
    for n in ("Lx", "Ly", "CFL", "sigma_0", "gamma", "Delt1"):
        out[n] = float(val(n))
    return out


def expected_refusals(c):
    """The constants cueva_params_from_module must refuse: branches of the
    reference's step the library does not run."""
    bad = [n for n in ("DIM", "REC_ELEC", "VDRIFT", "flux_rec_mp5", "REC_4VECTOR", "low_vel",
                       "SLC") if c[n] != (2 if n == "DIM" else 0)]
    if c["REC_PRIM"] not in (0, 1) or (c["REC_PRIM"] == 1 and 2 <= c["flux_solver"] <= 4):
        bad.append("REC_PRIM")
    if c["MIRK"] not in (0, 1, 2) or (c["order"] == 2 and c["MIRK"] == 1):  # imex_scheme.f95:51-56
        bad.append("MIRK")
    if c["BOUND"] not in (1, 2, 6):
        bad.append("BOUND")
    if c["conserved_to_primitive"] != 1:
        bad.append("conserved_to_primitive")
    if c["source_rec"] or c["source_rec_mpx"] or c["var_source_rec"]:
        bad.append("source_rec")
    if not 1 <= c["order"] <= 4:
        bad.append("order")
    return sorted(bad)


def build_probe(tmp, variant):
    """Compile the reference's parameters module of `variant`, the library's
    interfaces, cueva_params_ref.f95 and the probe; link with libcueva.so."""
    src = [os.path.join(REF, variant), os.path.join(FDIR, "cueva_gpu.f95"),
           os.path.join(FDIR, "cueva_params_ref.f95"), os.path.join(ROOT, "tests", "native", "params_probe.f95")]
    objs = []
    for k, s in enumerate(src):
        o = os.path.join(tmp, "u%d.o" % k)
        subprocess.run([FC, "-O1", "-c", s, "-J", tmp, "-o", o], check=True, capture_output=True)
        objs.append(o)
    exe = os.path.join(tmp, "probe")
    subprocess.run([FC] + objs + ["-L", LIB, "-lcueva", "-Wl,-rpath," + LIB, "-o", exe], check=True,
                   capture_output=True)
    return exe


@pytest.mark.parametrize("variant", ["parameters.f95", "parameters_TM.f95", "parameters_CE.f95",
                                     "parameters_REC_old.f95"])
def test_params_from_reference_module(tmp_path, variant):
    if not os.path.isdir(REF) or not os.path.exists(FC):
        pytest.skip("needs the reference sources and flang (build container only)")
    if not os.path.exists(os.path.join(LIB, "libcueva.so")):
        pytest.skip("libcueva.so not built")
    exe = build_probe(str(tmp_path), variant)
    res = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert res.returncode == 0, res.stderr
    got = {}
    for line in res.stdout.splitlines():
        t = line.split()
        if len(t) == 2 and t[0].isidentifier():
            got[t[0]] = float(t[1]) if any(c in t[1] for c in ".Ee") else int(t[1])
    c = constants(os.path.join(REF, variant))
    refused = sorted(re.findall(r"not run by this library: (\w+) =", res.stdout))
    assert refused == expected_refusals(c), res.stdout
    assert got["status"] == (1 if refused else 0), res.stdout
    for n in ("imax", "jmax", "order", "flux_solver", "rec_mp", "primitive_rec", "slop", "limit"):
        assert got[n] == c[n], n
    assert got["mirk"] == c["MIRK"] and got["rec_prim"] == c["REC_PRIM"]
    if "BOUND" not in refused:
        assert (got["bc_x"], got["bc_y"]) == {1: (0, 0), 2: (1, 1), 6: (0, 1)}[c["BOUND"]]
    assert got["longx"] == c["Lx"] and got["longy"] == c["Ly"]
    assert got["cfl"] == c["CFL"] and got["sigma"] == c["sigma_0"] and got["gamma"] == c["gamma"]
    # Delt1 as the variant's own expression (parameters.f95:306-313: CFL min(Delx, Dely, Delz);
    # parameters_TM.f95:304: CFL min(Delx, Dely))
    assert got["dt"] == c["Delt1"]
    assert c["Delt1"] == c["CFL"] * min(c["Lx"] / float(c["imax"]), c["Ly"] / float(c["jmax"]))


def test_literal_kinds():
    """The constant folder keeps Fortran's kinds: a default real literal is
    single precision, integer division truncates."""
    import numpy as np
    lit = re.compile(r"\d+\.?\d*(?:[eEdD][+-]?\d+)?")
    assert _literal(lit.match("0.1")) == "S('0.1')" and _literal(lit.match("1.d-3")) == "D('1.e-3')"
    assert _Int(7) / _Int(2) == 3 and _Int(-7) / _Int(2) == -3 and isinstance(_Int(7) / 2.0, float)
    assert np.float64(40) * np.float32("0.01") != 0.4
