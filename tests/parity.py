"""Shared helpers of the parity tests."""
import numpy as np

import cueva

FIELDS = list(cueva.PRIM_NAMES) + list(cueva.CONS_NAMES)


def interior(prims, cons, p):
    both = np.concatenate([prims, cons]) if cons is not None else prims
    return both[:, 6:6 + p.jmax + 1, 6:6 + p.imax + 1]


def rel_l1(a, b):
    """Per-field relative L1 distance sum|a-b| / sum|b| over the interior."""
    out = {}
    for k, f in enumerate(FIELDS[:a.shape[0]]):
        den = np.abs(b[k]).sum()
        num = np.abs(a[k] - b[k]).sum()
        out[f] = float(num / den) if den > 0 else (0.0 if num == 0 else np.inf)
    return out


def ndiff(a, b):
    return {f: int(np.count_nonzero(a[k] != b[k])) for k, f in enumerate(FIELDS[:a.shape[0]])}
