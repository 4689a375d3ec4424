"""Comparison helpers shared by the parity tests."""
from __future__ import annotations

import numpy as np


def rel_err(a, b) -> float:
    """max |a - b| / max |b| over the field (norm-relative error)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.max(np.abs(b)) if b.size else 0.0
    if scale == 0.0:
        return float(np.max(np.abs(a))) if a.size else 0.0
    return float(np.max(np.abs(a - b)) / scale)


def assert_close(name, a, b, tol):
    e = rel_err(a, b)
    assert np.all(np.isfinite(a)), f"{name}: non-finite values"
    assert e <= tol, f"{name}: relative error {e:.3e} > {tol:.1e}"
    return e


def sorted_pairs(off, nbr):
    """Per-target neighbour ids sorted ascending, as one int64 array of
    (target << 32 | neighbour)."""
    off = np.asarray(off, dtype=np.int64)
    nbr = np.asarray(nbr, dtype=np.int64)
    tgt = np.repeat(np.arange(len(off) - 1, dtype=np.int64), np.diff(off))
    return np.sort((tgt << 32) | nbr)
