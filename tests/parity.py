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


def _fma(a: float, b: float, c: float) -> float:
    """a * b + c with one rounding (exact rational arithmetic, then the
    correctly rounded conversion to double)."""
    from fractions import Fraction
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def squared_distance(dx, fused=None):
    """The reference's squared distance of one pair (asymmetricSearch.f90:
    209-237): each square and each partial sum rounded on its own, in axis
    order. fused = "x": dx0^2 fused into the sum with dx1^2, then dx2^2 fused
    into that; fused = "y": dx1^2 fused into dx0^2 instead (the two forms a
    compiler that contracts a * a + b builds) -- the mistakes the tests must
    be able to see."""
    if fused is None:
        dr = dx[0] * dx[0] + dx[1] * dx[1]
        for d in dx[2:]:
            dr = dr + d * d
        return dr
    dr = _fma(dx[0], dx[0], dx[1] * dx[1]) if fused == "x" else _fma(dx[1], dx[1], dx[0] * dx[0])
    for d in dx[2:]:
        dr = _fma(d, d, dr)
    return dr


def rounding_boundary_set(dim: int, npairs: int = 48, seed: int = 7):
    """Isolated particle pairs whose acceptance sqrt(dr) < 2 h flips when the
    squared distance is formed with a fused multiply-add: for each pair the
    cutoff 2 h is the larger of the two rounded square roots, so the
    reference's rounding and a fused one decide it differently. Pairs sit on
    a lattice 8 cutoffs apart (no pair sees another). Returns (x[n, dim],
    h[n], accept[npairs]) with accept the reference's decision."""
    import math
    rng = np.random.default_rng(seed)
    xs, hs, acc = [], [], []
    k = 0
    while len(acc) < npairs:
        k += 1
        side = int(np.ceil(npairs ** (1.0 / dim))) + 1
        cell = np.array(np.unravel_index(len(acc), (side,) * dim), dtype=np.float64) * 8.0
        qi = cell + rng.uniform(-0.5, 0.5, dim)
        u = rng.normal(size=dim)
        qj = qi + u / np.linalg.norm(u) * rng.uniform(0.9, 1.1)
        dx = [float(a - b) for a, b in zip(qi, qj)]
        ref = squared_distance(dx)
        for f in ("x", "y"):
            fu = squared_distance(dx, f)
            s_ref, s_fu = math.sqrt(ref), math.sqrt(fu)
            if s_ref != s_fu:
                cut = max(s_ref, s_fu)
                xs += [qi, qj]
                hs += [cut / 2, cut / 2]
                acc.append(s_ref < cut)
                break
        if k > 200000:
            raise RuntimeError("no discriminating pairs found")
    return np.array(xs), np.array(hs), np.array(acc)


# real(1.2_4, 8) - 1 (asymmetricSearch.f90:21): the grid's margin factor
GRID_MARGIN = 0.2000000476837158203125


def grid_rounding_cloud(dim: int, n: int = 64, seed: int = 3):
    """A seeded point cloud whose background-grid bounds (asymmetricSearch.f90:
    49-131, maxc = max + m (max - min), minc = min - m (max - min)) come out
    differently when the margin product is fused into the sum: returns x[n,
    dim] with at least one such axis for maxc and one for minc."""
    rng = np.random.default_rng(seed)
    for _ in range(100000):
        x = rng.uniform(-1.0, 1.0, (n, dim)) * rng.uniform(0.5, 3.0, dim) + rng.uniform(-5, 5, dim)
        mn, mx = x.min(axis=0), x.max(axis=0)
        up = lo = False
        for d in range(dim):
            delta = float(mx[d] - mn[d])
            m = GRID_MARGIN * delta
            up |= float(mx[d]) + m != _fma(GRID_MARGIN, delta, float(mx[d]))
            lo |= float(mn[d]) - m != _fma(-GRID_MARGIN, delta, float(mn[d]))
        if up and lo:
            return x
    raise RuntimeError("no cloud found")
