"""Benchmark of one CUEVA 2D resistive-RMHD step on one MI355X.

Workload: the 2048x4096 Kelvin-Helmholtz shear layer (periodic x, outflow
y), sigma = 1e6, MP5 + HLLC, IMEX SSP2(2,2,2), CFL 0.4.  A step is one full
time step of the whole grid; inputs are resident in HBM before timing
starts.  Two forms are timed: plain FP64 (bitwise with the reference build)
and the mixed form with the learned start of the primitive recovery; `value`
is the latter, the plain one and the accuracy after 100 steps (relative L1
of every field against plain FP64) are reported beside it.  One JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--config cfg2] [--form F] [--no-cpu]
  python bench.py --copy-calibration   # plain copy of known size, for rocprof counters
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "mhd-kh-instability_amd")
for _p in (PKG, ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6    # MI355X FP64 (vector and matrix)
METRIC = "zone-updates/s, 2048x4096 KH: plain FP64 vs mixed+learned-c2p at equal L1"

# compulsory bytes of one SSP2(2,2,2) step per zone-update (stage k reads U^n,
# P^n and earlier stage right-hand sides, writes its own; the last stage
# writes U^{n+1}, P^{n+1}): 91 doubles
STEP_BYTES_PER_ZONE = 728


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def copy_calibration(nbytes=1 << 30):
    import torch
    a = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "launches": 3}))


def cpu_baseline(workload_cfg, budget_s=20.0):
    """Time the reference binary (oracle/_ref) on the host cores, on a bounded
    128x256 sample of the same problem (same scheme and sigma), by
    differential wall clock; fall back to the oracle restatement."""
    import numpy as np
    import cueva
    from cueva import problems
    ni, nj, sigma, fs, rec, order, v0, bx, bz = problems.CONFIGS[workload_cfg]
    sample_ni, sample_nj = 128, 256
    zones = sample_ni * sample_nj
    name = {(7, 0, 2): "kh_hllc_mp5_128x256", (5, 0, 2): "kh_cfg1_hll_mp5_128x256"}.get((fs, rec, order))
    exe = os.path.join(ROOT, "oracle", "_ref", name or "-", "cueva_ref")
    p, prims = cueva.kh_problem(sample_ni, sample_nj, sigma=sigma, flux_solver=fs, rec_mp=rec, order=order,
                                rotated=True, v0=v0, bx0=bx, bz0=bz)
    ncpu = len(os.sched_getaffinity(0))
    threads = max(1, min(16, ncpu))
    if name and os.path.exists(exe):
        with tempfile.TemporaryDirectory() as tmp:
            inp = os.path.join(tmp, "in.bin")
            prims.astype("<f8").tofile(inp)

            def run(n, th):
                env = dict(os.environ, OMP_NUM_THREADS=str(th))
                t0 = time.perf_counter()
                subprocess.run([exe, inp, str(n), os.path.join(tmp, "o")], env=env, check=True,
                               capture_output=True)
                return time.perf_counter() - t0

            t1 = run(1, threads)
            t2 = run(2, threads)
            per = max(t2 - t1, 1e-3)
            n = int(max(2, min(40, budget_s / per)))
            tn = run(n + 2, threads)
            rate = zones * n / max(tn - t2, 1e-9)
            return {"value": rate, "unit": "zone-updates/s", "cores": threads, "kind": "reference",
                    "sample": "reference (flang -O2 -fopenmp, %s) on the %dx%d version of the same KH problem "
                              "(rotated frame), %d steps by differential wall clock T(%d)-T(2), "
                              "OMP_NUM_THREADS=%d" % (name, sample_ni, sample_nj, n, n + 2, threads)}
    from oracle.binding import Oracle
    o = Oracle(p)
    o.upload(prims)
    o.step(1)
    t0 = time.perf_counter()
    o.step(2)
    dt = time.perf_counter() - t0
    return {"value": zones * 2 / dt, "unit": "zone-updates/s", "cores": 1, "kind": "port",
            "sample": "oracle restatement, %dx%d, 2 steps, 1 thread" % (sample_ni, sample_nj)}


# the forms the metric compares; crmath 1 = the reference's correctly rounded
# acos/cos/cbrt (bitwise form), 0 = device math library
FORMS = {
    "plain_fp64": dict(precision=0, crmath=1, c2p_mode=0),
    "mixed": dict(precision=1, crmath=0, c2p_mode=0),
    "mixed_learned": dict(precision=1, crmath=0, c2p_mode=1),
}
# compulsory FP64 bytes per cell of the marching face kernel (every form: the
# stage planes are FP64, the FP32 work stays on chip).  Every launch reads the
# 16 stage + 5 primitive planes the stencils touch.  A stage launch also reads
# the other 9 U^n planes of the next stage's state and writes its 17-plane
# right-hand side and that 17-plane state; the last launch reads the same 9
# U^n planes and the earlier stages' right-hand sides (17 each) and writes
# U^{n+1} (14 planes).
MARCH_STAGE_BYTES = (16 + 5 + 9 + 17 + 17) * 8


def march_last_bytes(lmax):
    return (16 + 5 + 9 + 17 * (lmax - 1) + 14) * 8


ACCURACY_STEPS = 100


def load_packed():
    from cueva import learned_c2p
    if os.path.exists(learned_c2p.PARAMS_FILE):
        return learned_c2p.load_params()
    return None


def time_form(cfg, form, steps, warmup, packed):
    import cueva
    from cueva import problems
    p, prims = problems.config(cfg, **FORMS[form])
    s = cueva.Solver(p)
    if packed is not None:
        s.set_c2p_params(packed)
    s.upload(prims)
    del prims
    log("%s: warmup %d steps" % (form, warmup))
    s.step(warmup)
    s.step_counters()
    s.sync()
    s.set_profiling(True)
    log("%s: timed %d steps" % (form, steps))
    ms = s.time_steps(steps)
    s.sync()
    kt = s.kernel_times()
    s.set_profiling(False)
    cnt, fallback = s.step_counters()
    s.close()
    zones = p.imax * p.jmax
    return {"ms_per_step": ms / steps, "value": zones * steps / (ms / 1e3),
            "kernel_ms_per_step": {k: v[0] / steps for k, v in kt.items()},
            "kernel_launches": {k: v[1] for k, v in kt.items()},
            "fixups": cnt, "c2p_fallbacks": fallback, "p": p}


def accuracy(cfg, form, packed, steps=ACCURACY_STEPS):
    """Relative L1 of every field of `form` against plain FP64 after `steps` steps."""
    import numpy as np
    import cueva
    from cueva import problems
    out = {}
    for f in ("plain_fp64", form):
        p, prims = problems.config(cfg, **FORMS[f])
        with cueva.Solver(p) as s:
            if packed is not None:
                s.set_c2p_params(packed)
            s.upload(prims)
            s.step(steps)
            s.step_counters()
            pr, co = s.download()
        sl = (slice(None), slice(6, 6 + p.jmax + 1), slice(6, 6 + p.imax + 1))
        out[f] = np.concatenate([pr, co])[sl]
        del pr, co
    names = list(cueva.PRIM_NAMES) + list(cueva.CONS_NAMES)
    a, b = out[form], out["plain_fp64"]
    return {n: float(np.abs(a[k] - b[k]).sum() / max(np.abs(b[k]).sum(), 1e-300)) for k, n in enumerate(names)}


GROWTH_STEPS = 280  # cfg2 stops at step 290 in every form (the reference's HLLC contact solve)


def growth(cfg, form, packed, nsteps=GROWTH_STEPS, every=10):
    """KH vy-amplitude growth rate of `form` and of plain FP64 on the same samples."""
    import cueva
    from cueva import diagnostics, problems
    res = {}
    for f in ("plain_fp64", form):
        p, prims = problems.config(cfg, **FORMS[f])
        dt = cueva.derived(p)["dt"]
        with cueva.Solver(p) as s:
            if packed is not None:
                s.set_c2p_params(packed)
            s.upload(prims)
            del prims
            res[f] = diagnostics.kh_series(s, nsteps, every, dt, integral=True)
    t0, a0, _, stop0 = res["plain_fp64"]
    win = diagnostics.growth_window(a0)
    out = {"steps": nsteps, "every": every, "amplitude": "(max vy - min vy)/2 over the interior",
           "integral": "half the log slope of sum(vy^2 dx dy), the reference's measure"}
    if win is None:
        out["note"] = "amplitude did not reach 1.5x its initial value"
        return out
    out["window_t"] = [float(t0[win[0]]), float(t0[win[1] - 1])]
    for f, (t, a, vi, stop) in res.items():
        whole = len(a) >= win[1]
        out[f] = {"rate": diagnostics.growth_rate(t, a, win) if whole else None,
                  "rate_integral": diagnostics.growth_rate_of_integral(t, vi, win) if whole else None,
                  "A_first": float(a[0]), "A_last": float(a[-1]),
                  "stopped": None if stop is None else [stop[0], str(stop[1])]}
    if out[form]["rate"] is not None:
        out["rel_diff_vs_plain_fp64"] = out[form]["rate"] / out["plain_fp64"]["rate"] - 1.0
        out["rel_diff_vs_plain_fp64_integral"] = (out[form]["rate_integral"] /
                                                  out["plain_fp64"]["rate_integral"] - 1.0)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="cfg2")
    ap.add_argument("--form", default=None, choices=sorted(FORMS),
                    help="time one form only (default: plain FP64 and mixed + learned c2p)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the 100-step accuracy and growth runs")
    ap.add_argument("--copy-calibration", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the step on a single MI355X "
              "(multi-GPU slab decomposition is a design note only); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return

    import torch
    from cueva import problems

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    cfg = args.config
    ni, nj, sigma, fs, rec, order, v0, bx, bz = problems.CONFIGS[cfg]
    log("workload %s: %dx%d KH, sigma=%g, solver=%d, mp=%d, order=%d" % (cfg, ni, nj, sigma, fs, 5 + 2 * rec, order))
    packed = load_packed()
    if args.form:
        forms = [args.form]
    else:
        forms = ["plain_fp64", "mixed", "mixed_learned"] if packed is not None else ["plain_fp64", "mixed"]
    if "mixed_learned" in forms and packed is None:
        sys.exit("no trained parameters (cueva/c2p_mlp.json) for the learned form")
    res = {f: time_form(cfg, f, args.steps, args.warmup, packed) for f in forms}
    head = forms[-1]
    r = res[head]
    p = r.pop("p")
    for f in forms[:-1]:
        res[f].pop("p")
    kt = r["kernel_ms_per_step"]
    dom = max(kt, key=lambda k: kt[k])
    dom_avg_ms = kt[dom] / (r["kernel_launches"][dom] / args.steps)
    cells = (p.imax + 1) * (p.jmax + 1)
    lmax = 2 if order == 2 else 4
    # the face launches of one step: lmax - 1 stage launches and the last one
    alg_bytes = cells * (MARCH_STAGE_BYTES * (lmax - 1) + march_last_bytes(lmax)) / lmax if dom == "faces" else None
    achieved = alg_bytes / (dom_avg_ms / 1e3) / 1e9 if alg_bytes else None
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_%s_%s_%s.json" % (dom, cfg, head))
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get("bytes_per_launch")
    form_name = {"plain_fp64": "plain FP64", "mixed": "mixed FP32/FP64, Cardano c2p",
                 "mixed_learned": "mixed FP32/FP64 + learned c2p (BF16 MLP on MFMA)"}[head]
    out = {
        "metric": METRIC, "value": r["value"], "unit": "zone-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if head == "plain_fp64" else "f32+f64 (bf16 MLP)",
        "data": "synthetic (seeded KH initial data)",
        "config": {"workload": "%s: KH %dx%d, sigma=%g, MP%d + %s, IMEX %s, CFL 0.4" % (
            cfg, ni, nj, sigma, 5 + 2 * rec, {5: "HLL", 6: "HLLE", 7: "HLLC"}[fs],
            "SSP2(2,2,2)" if order == 2 else "SSP3(4,3,3)"), "form": form_name},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": traffic,
                     "alg_bytes_per_launch": alg_bytes, "avg_launch_ms": dom_avg_ms},
        "step_roofline": {"bytes_per_zone_update": STEP_BYTES_PER_ZONE,
                          "achieved_GBs": r["value"] * STEP_BYTES_PER_ZONE / 1e9,
                          "frac": r["value"] * STEP_BYTES_PER_ZONE / 1e9 / HBM_PEAK_GBS},
        "kernel_ms_per_step": kt, "fixups": r["fixups"], "c2p_fallbacks": r["c2p_fallbacks"],
        "forms": {f: {k: v for k, v in res[f].items() if k != "kernel_launches"} for f in forms},
    }
    if "plain_fp64" in res and head != "plain_fp64":
        out["speedup_vs_plain_fp64"] = r["value"] / res["plain_fp64"]["value"]
        if not args.no_accuracy:
            log("accuracy: %d steps of %s against plain FP64" % (ACCURACY_STEPS, head))
            out["rel_l1_vs_plain_fp64_after_%d_steps" % ACCURACY_STEPS] = accuracy(cfg, head, packed)
            log("growth: %d steps of plain FP64 and %s" % (GROWTH_STEPS, head))
            out["kh_vy_growth"] = growth(cfg, head, packed)
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(cfg)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
