"""The CPU restatement of the learned-recovery network (oracle/mlp_oracle.py),
pinned independently of the device: BF16 rounding against known encodings,
backpropagation against central finite differences of the unrounded network.
"""
import numpy as np

from oracle import mlp_oracle as mo


def test_bf16_rounding_known_values():
    ulp = 2.0 ** -7  # bf16 spacing at 1
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -8 + 2 ** -20, -2.5, 0.1], dtype=np.float32)
    # ties go to the even mantissa: 1+2^-8 -> 1, 1+3*2^-8 -> 1+2^-6; above a tie rounds up;
    # 0.1 = 0x3DCCCCCD -> 0x3DCD (0.10009765625)
    want = np.array([1.0, 1.0, 1.0 + 2 * ulp, 1.0 + ulp, -2.5, 0.10009765625])
    assert np.array_equal(mo.bf16(x), want)


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = rng.normal(scale=0.4, size=mo.NPARAM)
    x = rng.normal(size=(64, 3))
    t = rng.normal(size=(64, 2))
    loss, g = mo.loss_and_grad(p, x, t, rounding=False)
    h = 1e-6
    idx = rng.choice(mo.NPARAM, size=60, replace=False)
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        lp, _ = mo.loss_and_grad(pp, x, t, rounding=False)
        lm, _ = mo.loss_and_grad(pm, x, t, rounding=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])
