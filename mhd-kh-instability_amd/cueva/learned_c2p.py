"""Learned start point of the primitive recovery: data, training, parameters.

The network (include/cueva.h, csrc/cueva_mlp.hpp) maps the features of a
conserved state to (log W, log(p/C2)).  Training runs on the device through
the library's own kernels (cueva.Mlp); this module builds the samples and
drives it:

* harvest(): (conserved state, W, p) of interior cells of the code's own
  Kelvin-Helmholtz runs (the plain FP64 solver): the benchmark
  configurations 1 and 2 and two 256x512 variants;
* synthetic(): states drawn over W in [1, 10] and p/rho in [1e-3, 1e2] with
  random field directions, made conserved by the same relations as
  06_conserved_var.f95:77-140;
* train(): normalisation statistics, Adam on the device, packed parameters.

    python -m cueva.learned_c2p --out params.json     # on a GPU host

The packed parameters the package ships are in c2p_mlp.json beside this file.
"""
import argparse
import json
import os
import time

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
PARAMS_FILE = os.path.join(HERE, "c2p_mlp.json")
GAMMA = 4.0 / 3.0


def cons_from_prims(rho, p, v, E, B, gamma=GAMMA):
    """D, tau, S of primitive states (06_conserved_var.f95:77-140); v, E, B shape (3, n)."""
    v2 = (v[0] ** 2 + v[1] ** 2) + v[2] ** 2
    W = 1.0 / np.sqrt(1.0 - v2)
    eps = p / ((gamma - 1.0) * rho)
    h = rho * (1.0 + eps) + p
    hw2 = h * W ** 2
    e2b2 = (E[0] ** 2 + E[1] ** 2 + E[2] ** 2) + (B[0] ** 2 + B[1] ** 2 + B[2] ** 2)
    D = rho * W
    tau = 0.5 * e2b2 + hw2 - p
    ExB = np.stack([E[1] * B[2] - E[2] * B[1], E[2] * B[0] - E[0] * B[2], E[0] * B[1] - E[1] * B[0]])
    S = ExB + hw2 * v
    return D, tau, S, W


def features(D, tau, S, E, B):
    """Network inputs (log(D/C2), |S - E x B|/C2, log C2), as the device forms them."""
    e2b2 = 0.5 * (((((E[0] ** 2 + E[1] ** 2) + E[2] ** 2) + B[0] ** 2) + B[1] ** 2) + B[2] ** 2)
    sfx = (S[0] - E[1] * B[2]) + E[2] * B[1]
    sfy = (S[1] - E[2] * B[0]) + E[0] * B[2]
    sfz = (S[2] - E[0] * B[1]) + E[1] * B[0]
    C1 = (sfx ** 2 + sfy ** 2) + sfz ** 2
    C2 = tau - e2b2
    return np.stack([np.log(D / C2), np.sqrt(C1) / C2, np.log(C2)], axis=1), C2


def targets(W, p, C2):
    return np.stack([np.log(W), np.log(p / C2)], axis=1)


def synthetic(n, seed=1, gamma=GAMMA):
    """n states covering W in [1, 10], p/rho in [1e-3, 1e2], |B| in [1e-3, 1]."""
    rng = np.random.default_rng(seed)
    rho = np.exp(rng.uniform(np.log(0.1), np.log(10.0), n))
    p = rho * np.exp(rng.uniform(np.log(1e-3), np.log(1e2), n))
    W = 1.0 + np.exp(rng.uniform(np.log(1e-5), np.log(9.0), n))
    vmag = np.sqrt(1.0 - 1.0 / W ** 2)
    d = rng.normal(size=(3, n))
    v = vmag * d / np.linalg.norm(d, axis=0)
    bmag = np.exp(rng.uniform(np.log(1e-3), np.log(1.0), n))
    d = rng.normal(size=(3, n))
    B = bmag * d / np.linalg.norm(d, axis=0)
    vxb = np.stack([v[1] * B[2] - v[2] * B[1], v[2] * B[0] - v[0] * B[2], v[0] * B[1] - v[1] * B[0]])
    E = -vxb + 0.1 * np.linalg.norm(vxb, axis=0) * rng.normal(size=(3, n))
    D, tau, S, W = cons_from_prims(rho, p, v, E, B, gamma)
    x, C2 = features(D, tau, S, E, B)
    return x, targets(W, p, C2)


def from_fields(prims, cons, p):
    """Samples of every interior cell of downloaded planes (cueva.Solver.download)."""
    sl = (slice(6, 6 + p.jmax + 1), slice(6, 6 + p.imax + 1))
    f = {n: prims[k][sl].ravel() for k, n in enumerate(abi.PRIM_NAMES)}
    D, tau = cons[0][sl].ravel(), cons[1][sl].ravel()
    S = np.stack([cons[2][sl].ravel(), cons[3][sl].ravel(), cons[4][sl].ravel()])
    E = np.stack([f["Ex"], f["Ey"], f["Ez"]])
    B = np.stack([f["Bx"], f["By"], f["Bz"]])
    v2 = (f["Vx"] ** 2 + f["Vy"] ** 2) + f["Vz"] ** 2
    W = 1.0 / np.sqrt(1.0 - v2)
    x, C2 = features(D, tau, S, E, B)
    return x, targets(W, f["p"], C2)


# harvest runs: (label, problem factory, steps between snapshots, snapshots);
# the last one (the largest grid) supplies whatever the others fall short of
def _runs():
    from . import problems
    return [
        ("cfg1 128x256 HLL sigma=1e3", lambda: problems.config("cfg1"), 15, 8),
        ("256x512 HLLC v0=0.5 Bz=0.1", lambda: problems.kh_problem(256, 512, sigma=1.0e6, flux_solver=abi.HLLC,
                                                                    v0=0.5, bx0=0.01, bz0=0.1), 5, 7),
        ("256x512 HLLC v0=0.9", lambda: problems.kh_problem(256, 512, sigma=1.0e6, flux_solver=abi.HLLC,
                                                             v0=0.9, bx0=0.001), 10, 8),
        ("cfg2 2048x4096 HLLC sigma=1e6", lambda: problems.config("cfg2"), 40, 7),
    ]


def harvest(n_target, seed=2, log=print):
    """Samples from the code's own KH runs (plain FP64): snapshots of the
    benchmark configurations 1 and 2 through the phase they can be run
    (the reference's own solvers stop them at steps 135 and 290), and of
    256x512 runs with a guide field and at v0 = 0.9, until they stop."""
    rng = np.random.default_rng(seed)
    xs, ys = [], []
    runs = _runs()
    total = 0
    for r, (label, make, every, snaps) in enumerate(runs):
        prm, prims = make()
        left = len(runs) - r
        per = max(1, (n_target - total) // (left * snaps))
        got = 0
        with abi.Solver(prm) as s:
            s.upload(prims)
            del prims
            for snap in range(snaps):
                try:
                    s.step(every)
                    s.step_counters()
                except abi.CuevaError as e:  # a run that stops still gave its earlier snapshots
                    log("harvest: %s stops (%s)" % (label, e))
                    break
                x, y = from_fields(*s.download(), prm)
                ok = np.all(np.isfinite(x), axis=1) & np.all(np.isfinite(y), axis=1)
                idx = rng.choice(np.flatnonzero(ok), size=min(per, int(ok.sum())), replace=False)
                xs.append(x[idx])
                ys.append(y[idx])
                got += idx.size
        total += got
        log("harvest: %s: %d samples" % (label, got))
    return np.concatenate(xs), np.concatenate(ys)


def pack(weights, in_mean, in_std, out_mean, out_std):
    return np.concatenate([weights, in_mean, in_std, out_mean, out_std]).astype(np.float32)


def init_weights(seed=0):
    """Glorot-uniform weights, zero biases, in the packed trainable layout."""
    rng = np.random.default_rng(seed)

    def glorot(fan_out, fan_in):
        a = np.sqrt(6.0 / (fan_in + fan_out))
        return rng.uniform(-a, a, (fan_out, fan_in)).ravel()

    return np.concatenate([glorot(32, 3), np.zeros(32), glorot(32, 32), np.zeros(32), glorot(2, 32),
                           np.zeros(2)]).astype(np.float32)


def train(x, y, steps=30000, batch=4096, lr=3e-3, seed=0, log=print):
    """Normalise, train on the device with Adam (lr decayed tenfold twice),
    return the packed parameters."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(x.shape[0])
    x, y = x[perm], y[perm]
    in_mean, in_std = x.mean(axis=0), x.std(axis=0) + 1e-12
    out_mean, out_std = y.mean(axis=0), y.std(axis=0) + 1e-12
    xn = ((x - in_mean) / in_std).astype(np.float32)
    yn = ((y - out_mean) / out_std).astype(np.float32)
    with abi.Mlp(pack(init_weights(seed), in_mean, in_std, out_mean, out_std)) as m:
        m.set_data(xn, yn)
        done = 0
        for phase, frac in enumerate((0.6, 0.3, 0.1)):
            n = int(steps * frac)
            chunk = 1000
            while n > 0:
                k = min(chunk, n)
                loss = m.train(batch, k, lr * (0.1 ** phase))
                n -= k
                done += k
                log("step %d loss %.3e" % (done, loss))
        return m.params()


def save(packed, path=PARAMS_FILE, meta=None):
    with open(path, "w") as fh:
        json.dump({"packed": [float(v) for v in packed], "meta": meta or {}}, fh)


def load_params(path=PARAMS_FILE):
    with open(path) as fh:
        return np.array(json.load(fh)["packed"], dtype=np.float32)


def newton_start_quality(packed, x, y):
    """Host check of a parameter set: relative error of the predicted W."""
    with abi.Mlp(packed) as m:
        xn = (x - packed[1250:1253]) / packed[1253:1256]
        yn = m.forward(xn.astype(np.float32))
    logw = yn[:, 0] * packed[1258] + packed[1256]
    return np.abs(np.expm1(logw - y[:, 0]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=PARAMS_FILE)
    ap.add_argument("--harvest", type=int, default=2_000_000)
    ap.add_argument("--synthetic", type=int, default=2_000_000)
    ap.add_argument("--steps", type=int, default=30000)
    args = ap.parse_args()
    t0 = time.time()
    xh, yh = harvest(args.harvest)
    print("harvested %d KH samples in %.1f s" % (xh.shape[0], time.time() - t0), flush=True)
    xs, ys = synthetic(args.synthetic)
    x, y = np.concatenate([xh, xs]), np.concatenate([yh, ys])
    packed = train(x, y, steps=args.steps, log=lambda s: print(s, flush=True))
    err_h = newton_start_quality(packed, xh[:200000], yh[:200000])
    err_s = newton_start_quality(packed, xs[:200000], ys[:200000])
    meta = {"samples_kh": int(xh.shape[0]), "samples_synthetic": int(xs.shape[0]), "steps": args.steps,
            "w_rel_err_kh_median": float(np.median(err_h)), "w_rel_err_kh_max": float(err_h.max()),
            "w_rel_err_syn_median": float(np.median(err_s)), "w_rel_err_syn_p99": float(np.quantile(err_s, 0.99)),
            "train_seconds": time.time() - t0}
    save(packed, args.out, meta)
    print(json.dumps(meta), flush=True)


if __name__ == "__main__":
    main()
