"""Helpers for the tests of the Fortran host program (mysph_amd/bin/mysph_host)."""
from __future__ import annotations

import os
import struct
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "mysph_amd", "bin", "mysph_host")
LIB = os.path.join(ROOT, "mysph_amd", "libsphgpu.so")


def ensure_built():
    """Build the host if it is missing (it needs libsphgpu.so, which the
    package build makes)."""
    from mysph_amd import build as B
    if not os.path.exists(LIB):
        B.build()
    if not os.path.exists(HOST):
        B.build_fortran()
    return HOST


def write_toml(d, text: str):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "sph.toml"), "w") as f:
        f.write(text)


def run_host(d, *args, check=True, timeout=600):
    ensure_built()
    r = subprocess.run([HOST, str(d), *map(str, args)], capture_output=True, text=True,
                       timeout=timeout, cwd=str(d))
    if check and r.returncode != 0:
        raise AssertionError(f"mysph_host failed ({r.returncode}):\n{r.stdout[-3000:]}\n{r.stderr[-3000:]}")
    return r


def write_particle_file(path, case):
    """A case dict in the particle-file format the host reads (the records
    cases_m.f90's write_dump writes)."""
    n = case["n_real"] + case["n_dummy"]

    def rec(f, name, arr, kind):
        a = np.ascontiguousarray(arr, dtype="<f8" if kind == 8 else "<i4").ravel()
        f.write(name.ljust(8).encode())
        f.write(struct.pack("<iq", kind, a.size))
        f.write(a.tobytes())

    with open(path, "wb") as f:
        rec(f, "ntotal", [case["n_real"]], 4)
        rec(f, "ndummy", [case["n_dummy"]], 4)
        rec(f, "dim", [case["dim"]], 4)
        for k in ("x", "v", "mass", "rho", "p", "e", "c", "h"):
            rec(f, k, np.asarray(case[k], dtype=np.float64)[:n], 8)
        for k in ("type", "state", "boundary"):
            rec(f, k, np.asarray(case[k], dtype=np.int32)[:n], 4)


def read_dat(path):
    """Columns of a Type_<t>_<i>.dat file: (names, float array)."""
    with open(path) as f:
        names = f.readline().split()
        rows = [list(map(float, line.split())) for line in f if line.strip()]
    return names, np.array(rows)
