"""Writes tests/golden/dam_break_3d_1000.{npz,json}: the oracle's state after
1000 leap-frog steps of the benchmark's full workload, the 3D dam break at
dx = 0.0075 (1,024,000 fluid + 366,864 boundary particles).

The oracle (oracle/sph_oracle.c, pinned to the compiled reference by
tests/test_oracle_golden.py) needs 25-50 minutes of host time for this, far
more than a test may take, so the result is committed as a fixture: every
SAMPLE-th fluid particle's x, v, rho, p and e, and for the whole fluid set
per-field sums, largest magnitudes and SHA-256 hashes. The GPU test
(tests/test_gpu_full_size.py) holds the GPU run to the 1000-step bars on the
sample and on the whole-set sums. Run in the build container:
    python tests/golden/make_full_size.py [threads]
"""
from __future__ import annotations

import hashlib
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from mysph_amd import cases  # noqa: E402

NAME = "dam_break_3d_1000"
DX = 0.0075
STEPS = 1000
SAMPLE = 256
FIELDS = ("x", "v", "rho", "p", "e")


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def summary(a):
    """Whole-set figures of one field: exact (math.fsum) sum per component
    and largest magnitude."""
    a = np.asarray(a, dtype=np.float64).reshape(len(a), -1)
    return {"sum": [math.fsum(a[:, c]) for c in range(a.shape[1])],
            "absmax": float(np.max(np.abs(a)))}


def main(threads: int):
    case = cases.dam_break_3d(dx=DX)
    nt = case["n_real"]
    oracle.set_threads(threads)
    orc = oracle.OracleSim(case)
    t0 = time.time()
    for k in range(STEPS // 100):
        orc.run(100)
        print(f"oracle: {(k + 1) * 100} steps, {time.time() - t0:.0f} s", flush=True)
    pick = np.arange(0, nt, SAMPLE)
    arrays = {"index": pick.astype(np.int32)}
    meta = {"dx": DX, "steps": STEPS, "sample": SAMPLE, "n_real": nt, "n_dummy": case["n_dummy"],
            "threads": threads, "oracle_seconds": time.time() - t0}
    for k in FIELDS:
        a = orc[k][:nt]
        arrays[k] = a[pick].copy()
        meta["sha_" + k] = sha(a)
        meta["summary_" + k] = summary(a)
    np.savez_compressed(os.path.join(HERE, NAME + ".npz"), **arrays)
    with open(os.path.join(HERE, NAME + ".json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print("wrote", NAME, flush=True)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else (os.cpu_count() or 1))
