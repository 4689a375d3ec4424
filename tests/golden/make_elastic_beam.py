"""Writes tests/golden/elastic_beam_20k.{json,npz}: the oracle's FP64 run of
the 2D elastic cantilever (mysph_amd.cases.elastic_beam: 0.2 x 0.02 m,
dx = 0.5 mm, 16,120 particles with three clamped columns, E = 2 GPa,
nu = 0.3975, artificial stress eps = 0.3, first bending mode at 0.01 m/s)
over 20,000 leap-frog steps.

The reference cannot run this case itself: its solid constants are fixed by
particle type (eos.f90:219-248, viscosity.f90:14-33), and no type has
E = 2 GPa. The oracle is the C restatement pinned bit-for-bit to the
reference on the reference's own beams (tests/test_oracle_golden.py); this
fixture carries that pin to the longer, stiffer case.

Recorded:
  - the tip deflection (mean y displacement of the free-end column) every
    500 steps;
  - x, v, rho, p and stress of every 80th particle after 20,000 steps;
  - the oracle's own sensitivity: the same run with the initial velocities
    perturbed by 1e-15 (relative, seeded), as the relative change of every
    recorded quantity. A GPU result is held to max(floor, 10 x sensitivity).

Run in the build container (about 15 minutes on 8 cores):
    python tests/golden/make_elastic_beam.py
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from mysph_amd import cases  # noqa: E402

STEPS = 20000
EVERY = 500
SAMPLE = 80
FIELDS = ("x", "v", "rho", "p", "stress")


def tip_mask(case):
    x = case["x"][:, 0]
    return x > x.max() - 0.5 * (x.max() - np.sort(np.unique(x))[-2])


def run(case):
    o = oracle.OracleSim(case)
    tip = tip_mask(case)
    y0 = case["x"][tip, 1].copy()
    series = []
    for _ in range(STEPS // EVERY):
        o.run(EVERY)
        series.append(float(np.mean(o["x"][tip, 1] - y0)))
    nt = case["n_real"]
    out = {k: np.array(o[k][:nt:SAMPLE]) for k in FIELDS}
    return np.array(series), out


def rel(a, b):
    s = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / s) if s > 0 else float(np.max(np.abs(a)))


def main():
    oracle.set_threads(os.cpu_count() or 1)
    case = cases.elastic_beam()
    t = time.time()
    series, fields = run(case)
    print(f"run: {time.time() - t:.0f} s, tip deflection {series[-1]:.6e} m", flush=True)
    pert = dict(case)
    rng = np.random.default_rng(7)
    pert["v"] = case["v"] * (1 + 1e-15 * rng.standard_normal(case["v"].shape))
    s2, f2 = run(pert)
    sens = {"tip": rel(s2, series)}
    sens.update({k: rel(f2[k], fields[k]) for k in FIELDS})
    print("sensitivity", sens, flush=True)
    np.savez_compressed(os.path.join(HERE, "elastic_beam_20k.npz"), tip=series,
                        **{k: fields[k] for k in FIELDS})
    meta = {
        "case": "mysph_amd.cases.elastic_beam() defaults",
        "n_particles": int(case["n_real"]),
        "steps": STEPS,
        "tip_every": EVERY,
        "sample_every": SAMPLE,
        "dt": float(case["config"].dt),
        "sensitivity_1e-15": sens,
        "producer": "oracle/sph_oracle.c (FP64), tests/golden/make_elastic_beam.py",
    }
    with open(os.path.join(HERE, "elastic_beam_20k.json"), "w") as f:
        json.dump(meta, f, indent=1)


if __name__ == "__main__":
    main()
