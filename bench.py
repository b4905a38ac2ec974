"""Benchmark of one CUEVA 2D resistive-RMHD step on one MI355X.

Workload: the 2048x4096 Kelvin-Helmholtz shear layer (periodic x, outflow
y), sigma = 1e6, MP5 + HLLC, IMEX SSP2(2,2,2), CFL 0.4.  A step is one full
time step of the whole grid; inputs are resident in HBM before timing
starts.  Three forms are timed in the same run: plain FP64 (bitwise with the
reference build; `value`), the mixed form with the Cardano recovery, and the
mixed form with the learned start of the recovery; the mixed forms' speedup
over plain FP64, their accuracy after 100 steps (relative L1 of every field
against plain FP64, with the equal-L1 bar of 1e-4 on rho checked) and the
KH growth rates are reported beside it.  One JSON line.

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
FP64_PEAK_TFLOPS = 78.6    # MI355X FP64 vector
# VALU issue roof: 256 CUs x 4 SIMDs, one wave64 instruction per SIMD every
# 4 cycles (16 lanes a cycle, FP64 and unpacked FP32 alike) at 2.4 GHz
VALU_PEAK_INSTR = 1024 * 2.4e9 / 4
EQUAL_L1_BAR = 1.0e-4      # relative L1 on rho after 100 steps, the mixed forms' accuracy bar
METRIC = "zone-updates/s, 2048x4096 KH: plain FP64 vs mixed+learned-c2p at equal L1"

# compulsory bytes of one step per zone-update: stage k of lmax reads U^n
# and P^n (19 planes) and the k-1 earlier stage right-hand sides (17 planes
# each: 14 flux differences, 3 stiff E sources) and writes its own; the last
# stage writes U^{n+1} and P^{n+1} (19) instead.  SSP2(2,2,2): 91 doubles =
# 728 B; SSP2(3,3,2)-LUM: 161 = 1288 B; SSP3(4,3,3): 248 = 1984 B.


def step_bytes_per_zone(lmax):
    reads = 19 * lmax + 17 * lmax * (lmax - 1) // 2
    writes = 17 * (lmax - 1) + 19
    return 8 * (reads + writes)


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


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(workload_cfg, budget_s=12.0):
    """Time the reference binary (oracle/_ref) on the host cores, on a bounded
    128x256 sample of the same problem (same scheme and sigma), by
    differential wall clock, with every core this job may use and with one
    thread; fall back to the oracle restatement (one thread)."""
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
    # every core this job may use: its affinity mask, within the host's
    # per-GPU share that OMP_NUM_THREADS states where it is set
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = max(1, min(ncpu, share) if share > 0 else ncpu)
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

            def rate_at(th):
                t1 = run(1, th)
                t2 = run(2, th)
                per = max(t2 - t1, 1e-3)
                n = int(max(2, min(60, budget_s / per)))
                tn = run(n + 2, th)
                return zones * n / max(tn - t2, 1e-9), n

            rate, n = rate_at(threads)
            rate1, n1 = rate_at(1)
            return {"value": rate, "unit": "zone-updates/s", "cores": threads, "kind": "reference",
                    "cpu_model": cpu_model(), "nproc": os.cpu_count(), "affinity_cores": ncpu,
                    "one_thread": {"value": rate1, "cores": 1, "steps": n1},
                    "sample": "reference (flang -O2 -fopenmp, %s) on the %dx%d version of the same KH problem "
                              "(rotated frame), %d steps by differential wall clock T(%d)-T(2) with "
                              "OMP_NUM_THREADS=%d (every core of this job), and %d steps with one thread"
                              % (name, sample_ni, sample_nj, n, n + 2, threads, n1)}
    from oracle.binding import Oracle
    o = Oracle(p)
    o.upload(prims)
    o.step(1)
    t0 = time.perf_counter()
    o.step(2)
    dt = time.perf_counter() - t0
    return {"value": zones * 2 / dt, "unit": "zone-updates/s", "cores": 1, "kind": "port",
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "sample": "oracle restatement, %dx%d, 2 steps, 1 thread" % (sample_ni, sample_nj)}


# the forms the metric compares; crmath 1 = the reference's correctly rounded
# acos/cos/cbrt (bitwise form), 0 = device math library; hllc_seed 1 = the
# reference build's first Newton point of the HLLC contact solve (bitwise
# form), 0 = the clamped root of the quadratic it refines (the mixed forms,
# which are not bitwise and converge from it in fewer iterations)
FORMS = {
    "plain_fp64": dict(precision=0, crmath=1, c2p_mode=0, hllc_seed=1),
    "mixed": dict(precision=1, crmath=0, c2p_mode=0, hllc_seed=0),
    "mixed_learned": dict(precision=1, crmath=0, c2p_mode=1, hllc_seed=0),
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


def time_form(cfg, form, steps, warmup, packed, host_io=False):
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
    zones = p.imax * p.jmax
    out = {"ms_per_step": ms / steps, "value": zones * steps / (ms / 1e3),
           "kernel_ms_per_step": {k: v[0] / steps for k, v in kt.items()},
           "kernel_launches": {k: v[1] for k, v in kt.items()},
           "fixups": cnt, "c2p_fallbacks": fallback, "p": p}
    if host_io:
        # the host side of the boundary: the state down (14 primitive + 5
        # conserved planes) and up again (14 primitive planes), synchronous
        # calls on pageable host arrays, as a host that crossed PCIe every
        # step would pay; the drop-in path keeps the state in HBM between
        # the reference's output calls
        t0 = time.perf_counter()
        prims_h, _ = s.download(cons=True)
        t1 = time.perf_counter()
        s.upload(prims_h)
        s.sync()
        t2 = time.perf_counter()
        down, up = (t1 - t0) * 1e3, (t2 - t1) * 1e3
        out["host_boundary"] = {
            "download_ms": down, "upload_ms": up,
            "bytes_down": int(prims_h.nbytes * 19 // 14), "bytes_up": int(prims_h.nbytes),
            "pcie_inclusive_value": zones / ((ms / steps + down + up) / 1e3),
            "note": "zone-updates/s of a host that downloads and re-uploads the state every step; "
                    "not the metric's value"}
        del prims_h
    s.close()
    return out


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
    res = {f: time_form(cfg, f, args.steps, args.warmup, packed, host_io=(f == forms[0])) for f in forms}
    head = forms[0]  # plain FP64 unless one form was asked for
    r = res[head]
    p = r["p"]
    for f in forms:
        res[f].pop("p")
    zones = p.imax * p.jmax
    lmax = order  # stages of the IMEX scheme
    step_bytes = step_bytes_per_zone(lmax)

    def rooflines(form):
        """HBM and VALU rooflines of the form's face launches (its dominant
        kernel class): compulsory bytes per launch = the step's compulsory
        bytes per zone-update (728 B for SSP2(2,2,2)) spread over the lmax
        face launches of a step; VALU
        instructions per launch from the committed rocprofv3 counter summary."""
        fr = res[form]
        kt = fr["kernel_ms_per_step"]
        avg_ms = kt["faces"] / (fr["kernel_launches"]["faces"] / args.steps)
        alg = step_bytes * zones / lmax
        achieved = alg / (avg_ms / 1e3) / 1e9
        prof = os.path.join(ROOT, "profiles", "faces_%s_%s.json" % (cfg, form))
        prof = json.load(open(prof)) if os.path.exists(prof) else {}
        hbm = {"bound": "hbm", "kernel": "faces", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
               "frac": achieved / HBM_PEAK_GBS, "traffic": prof.get("hbm_bytes_per_launch"),
               "alg_bytes_per_launch": alg, "avg_launch_ms": avg_ms}
        valu = None
        if prof.get("valu_instr_per_launch"):
            rate = prof["valu_instr_per_launch"] / (avg_ms / 1e3)
            valu = {"bound": "valu", "kernel": "faces", "achieved": rate, "peak": VALU_PEAK_INSTR,
                    "unit": "wave64 VALU instructions/s", "frac": rate / VALU_PEAK_INSTR,
                    "instr_per_launch": prof["valu_instr_per_launch"]}
            if prof.get("fp64_flop_per_launch"):
                tf = prof["fp64_flop_per_launch"] / (avg_ms / 1e3) / 1e12
                valu["fp64_tflops"] = tf
                valu["fp64_frac"] = tf / FP64_PEAK_TFLOPS
        return hbm, valu

    hbm, valu = rooflines(head)
    wl = "%s: KH %dx%d, sigma=%g, MP%d + %s, IMEX %s, CFL 0.4" % (
        cfg, ni, nj, sigma, 5 + 2 * rec, {5: "HLL", 6: "HLLE", 7: "HLLC"}[fs],
        {2: "SSP2(2,2,2)", 3: "SSP2(3,3,2)-LUM", 4: "SSP3(4,3,3)"}[order])
    form_name = {"plain_fp64": "plain FP64", "mixed": "mixed FP32/FP64, Cardano recovery",
                 "mixed_learned": "mixed FP32/FP64 + learned recovery start (BF16 MLP on MFMA)"}
    out = {
        "metric": METRIC, "value": r["value"], "unit": "zone-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if head == "plain_fp64" else "f32+f64",
        "data": "synthetic (seeded KH initial data)",
        "config": {"workload": wl, "form": form_name[head]},
        "roofline": hbm,
        "valu_roofline": valu,
        "step_roofline": {"bytes_per_zone_update": step_bytes,
                          "achieved_GBs": r["value"] * step_bytes / 1e9,
                          "frac": r["value"] * step_bytes / 1e9 / HBM_PEAK_GBS},
        "kernel_ms_per_step": r["kernel_ms_per_step"], "fixups": r["fixups"], "c2p_fallbacks": r["c2p_fallbacks"],
        "host_boundary": r.get("host_boundary"),
        "forms": {},
    }
    for f in forms:
        e = {k: v for k, v in res[f].items() if k not in ("kernel_launches", "host_boundary")}
        e["form"] = form_name[f]
        if f != head:
            e["speedup_vs_plain_fp64"] = res[f]["value"] / r["value"]
            # all forms keep every plane in HBM in FP64 (the FP32 arithmetic
            # stays on chip), so their compulsory bytes are the same
            e["roofline"], e["valu_roofline"] = rooflines(f)
        out["forms"][f] = e
    if "mixed" in res and "mixed_learned" in res:
        # what the learned start gains over the method without it (the
        # Cardano recovery of the same mixed form): its recovery pass time
        c_ms = res["mixed"]["kernel_ms_per_step"]["final"]
        l_ms = res["mixed_learned"]["kernel_ms_per_step"]["final"]
        out["forms"]["mixed_learned"]["vs_cardano_recovery"] = {
            "recovery_pass_ms": l_ms, "cardano_recovery_pass_ms": c_ms, "time_ratio": l_ms / c_ms,
            "fallbacks": res["mixed_learned"]["c2p_fallbacks"],
            "note": "the recovery pass is bound by its HBM traffic, which both share: the learned start "
                    "replaces the Cardano solve for the cells it is taken for and saves no time"}
    if not args.no_accuracy and len(forms) > 1:
        for f in forms[1:]:
            log("accuracy: %d steps of %s against plain FP64" % (ACCURACY_STEPS, f))
            acc = accuracy(cfg, f, packed)
            over = sorted(k for k, v in acc.items() if v > EQUAL_L1_BAR)
            out["forms"][f]["rel_l1_vs_plain_fp64_after_%d_steps" % ACCURACY_STEPS] = acc
            out["forms"][f]["equal_l1"] = {"bar": EQUAL_L1_BAR, "field": "rho", "rho": acc["rho"],
                                           "passed": acc["rho"] <= EQUAL_L1_BAR, "fields_over_bar": over}
        last = forms[-1]
        log("growth: %d steps of plain FP64 and %s" % (GROWTH_STEPS, last))
        out["forms"][last]["kh_vy_growth"] = growth(cfg, last, packed)
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(cfg)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
