"""Benchmark of one CUEVA 2D resistive-RMHD step on one MI355X.

Default workload: the 2048x4096 Kelvin-Helmholtz shear layer (periodic x,
outflow y), sigma = 1e6, MP5 + HLLC, IMEX SSP2(2,2,2), CFL 0.4, plain FP64.
A step is one full time step of the whole grid; inputs are resident in HBM
before timing starts.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--config cfg2] [--no-cpu]
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

# compulsory bytes of the face kernel per face pair: 16 stage planes + 5
# primitive planes (vx vy vz p rho) read once, 28 face-flux planes written
FACES_BYTES_PER_PAIR = (16 + 5 + 28) * 8
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="cfg2")
    ap.add_argument("--no-cpu", action="store_true")
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
    import cueva
    from cueva import problems

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    cfg = args.config
    ni, nj, sigma, fs, rec, order, v0, bx, bz = problems.CONFIGS[cfg]
    log("workload %s: %dx%d KH, sigma=%g, solver=%d, mp=%d, order=%d" % (cfg, ni, nj, sigma, fs, 5 + 2 * rec, order))
    p, prims = cueva.kh_problem(ni, nj, sigma=sigma, flux_solver=fs, rec_mp=rec, order=order, v0=v0, bx0=bx,
                                bz0=bz)
    s = cueva.Solver(p)
    s.upload(prims)
    del prims
    zones = ni * nj
    log("warmup %d steps" % args.warmup)
    s.step(args.warmup)
    s.step_counters()
    torch.cuda.synchronize()
    s.sync()
    s.set_profiling(True)
    log("timed %d steps" % args.steps)
    ms = s.time_steps(args.steps)
    torch.cuda.synchronize()
    s.sync()
    kt = s.kernel_times()
    s.set_profiling(False)
    cnt, _ = s.step_counters()
    t_step = ms / args.steps
    value = zones * args.steps / (ms / 1e3)
    dom = max(kt, key=lambda k: kt[k][0])
    dom_ms, dom_n = kt[dom]
    dom_avg_s = dom_ms / dom_n / 1e3
    pairs = (p.imax + 2) * (p.jmax + 2)
    if dom == "faces":
        alg_bytes = pairs * FACES_BYTES_PER_PAIR
    else:
        alg_bytes = None
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_%s_%s.json" % (dom, cfg))
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get("bytes_per_launch")
    achieved = alg_bytes / dom_avg_s / 1e9 if alg_bytes else None
    out = {
        "metric": METRIC, "value": value, "unit": "zone-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": t_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded KH initial data)",
        "config": {"workload": "%s: KH %dx%d, sigma=%g, MP%d + %s, IMEX %s, CFL 0.4, plain FP64" % (
            cfg, ni, nj, sigma, 5 + 2 * rec, {5: "HLL", 6: "HLLE", 7: "HLLC"}[fs],
            "SSP2(2,2,2)" if order == 2 else "SSP3(4,3,3)", ), "form": "plain FP64"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": traffic,
                     "alg_bytes_per_launch": alg_bytes, "avg_launch_ms": dom_ms / dom_n},
        "step_roofline": {"bytes_per_zone_update": STEP_BYTES_PER_ZONE,
                          "achieved_GBs": value * STEP_BYTES_PER_ZONE / 1e9,
                          "frac": value * STEP_BYTES_PER_ZONE / 1e9 / HBM_PEAK_GBS},
        "kernel_ms_per_step": {k: v[0] / args.steps for k, v in kt.items()},
        "fixups": cnt,
    }
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(cfg)
    s.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
