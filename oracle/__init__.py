"""CPU oracle for the SPH hot path. TEST INFRASTRUCTURE ONLY.

``liboracle.so`` is built from ``sph_oracle.c`` (a restatement of the
reference routines, each citing the reference file:line it follows). Only
tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg may
import this package; the product (mysph_amd) never does.

``ref_driver`` (oracle/_ref, built by build_ref.sh from the reference's own
Fortran sources) is the checker of the checker: tests/golden/make_golden.py
runs it to write the golden fixtures that pin this restatement.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
SRC = os.path.join(HERE, "sph_oracle.c")
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")

# -O3 on an x86-64-v3 target with contraction off: no FMA is formed and
# nothing is reassociated, so results equal the reference's default build.
CFLAGS = ["-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
          "-fPIC", "-shared", "-Wall", "-Wno-unknown-pragmas"]


def build(verbose: bool = False) -> str:
    cmd = ["gcc", *CFLAGS, "-o", LIB_PATH + ".tmp", SRC, "-lm"]
    subprocess.run(cmd, check=True, capture_output=not verbose)
    os.replace(LIB_PATH + ".tmp", LIB_PATH)
    return LIB_PATH


def build_ref(verbose: bool = False) -> str | None:
    """Build oracle/_ref/ref_driver from the reference sources when they are
    present (this container); returns the driver path or None."""
    script = os.path.join(HERE, "build_ref.sh")
    if not os.path.isdir("/root/reference/src"):
        return None
    subprocess.run(["bash", script], check=True, capture_output=not verbose)
    return REF_DRIVER if os.path.exists(REF_DRIVER) else None


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        L.oracle_bggs.restype = ctypes.c_int
        L.oracle_bggs.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  dp, dp, ctypes.POINTER(ctypes.c_int32), ctypes.c_int64,
                                  ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int32),
                                  dp, dp, ctypes.POINTER(ctypes.c_int64), dp]
        L.oracle_kernel.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, dp,
                                    ctypes.c_double, dp, dp]
        L.oracle_derivatives.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.oracle_run.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int]
        L.oracle_external_force.argtypes = [ctypes.c_void_p, ctypes.c_void_p, dp]
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def set_threads(n: int) -> int:
    return lib().oracle_set_threads(int(n))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def kernel(dim: int, skf: int, r: float, dx, h: float):
    dxa = np.ascontiguousarray(dx, dtype=np.float64)
    w = np.zeros(1)
    dw = np.zeros(dim)
    lib().oracle_kernel(dim, skf, float(r), _dp(dxa), float(h), _dp(w), _dp(dw))
    return w[0], dw


def bggs(dim: int, skf: int, ntotal: int, x, h, state=None):
    """Neighbour lists as the reference BGGS builds them (pair order kept).
    Returns (offsets, nbr, w, dwdx, grid9)."""
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    h = np.ascontiguousarray(h, dtype=np.float64)
    n = h.shape[0]
    st = None if state is None else np.ascontiguousarray(state, dtype=np.int32)
    off = np.zeros(ntotal + 1, dtype=np.int64)
    needed = np.zeros(1, dtype=np.int64)
    grid = np.zeros(9)
    L = lib()
    stp = None if st is None else st.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    rc = L.oracle_bggs(dim, skf, ntotal, n, _dp(x), _dp(h), stp, 0,
                       off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), None, None, None,
                       needed.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), _dp(grid))
    cap = int(needed[0])
    nbr = np.zeros(max(cap, 1), dtype=np.int32)
    w = np.zeros(max(cap, 1))
    dw = np.zeros(max(cap, 1) * dim)
    rc = L.oracle_bggs(dim, skf, ntotal, n, _dp(x), _dp(h), stp, cap,
                       off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                       nbr.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _dp(w), _dp(dw),
                       needed.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), _dp(grid))
    if rc != 0:
        raise RuntimeError(f"oracle_bggs failed: {rc}")
    return off, nbr[:cap], w[:cap], dw[:cap * dim].reshape(cap, dim), grid


class OracleArrays(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("ntotal", ctypes.c_int32), ("n", ctypes.c_int32),
                ("pad", ctypes.c_int32)] + [
        (name, ctypes.c_void_p) for name in (
            "x", "v", "disp", "mass", "rho", "p", "e", "c", "h", "eta", "divx", "divv",
            "stress", "type", "state", "boundary", "shear", "shear_prev", "dsdt", "v_prev",
            "rho_prev", "e_prev", "dvdt", "drhodt", "dedt")]


class OracleSim:
    """Particle state on the host, advanced by the oracle's leap_frog."""

    VEC = ("x", "v", "disp", "v_prev", "dvdt")
    TEN = ("stress", "shear", "shear_prev", "dsdt")
    SCA = ("mass", "rho", "p", "e", "c", "h", "eta", "divx", "divv", "rho_prev", "e_prev",
           "drhodt", "dedt")

    def __init__(self, case: dict, cfg=None):
        self.cfg = (cfg if cfg is not None else case["config"]).copy()
        self.dim = int(case["dim"])
        self.ntotal = int(case["n_real"])
        self.n = int(case["n_real"]) + int(case["n_dummy"])
        d, n = self.dim, self.n
        self.a = {}
        for k in self.VEC:
            self.a[k] = np.zeros((n, d))
        for k in self.TEN:
            self.a[k] = np.zeros((n, d, d))
        for k in self.SCA:
            self.a[k] = np.zeros(n)
        for k in ("x", "v", "mass", "rho", "p", "e", "c", "h"):
            self.a[k][:] = np.asarray(case[k], dtype=np.float64).reshape(self.a[k].shape)
        if "shear" in case:
            self.a["shear"][:] = case["shear"]
        for k in ("type", "state", "boundary"):
            self.a[k] = np.ascontiguousarray(case[k], dtype=np.int32).copy()
        self.step = 1
        self._arr = OracleArrays()
        self._arr.dim, self._arr.ntotal, self._arr.n = d, self.ntotal, n
        for k, v in self.a.items():
            setattr(self._arr, k, v.ctypes.data)

    def derivatives(self):
        rc = lib().oracle_derivatives(ctypes.byref(self.cfg), ctypes.byref(self._arr))
        if rc != 0:
            raise RuntimeError(f"oracle_derivatives failed: {rc}")

    def external_force(self) -> np.ndarray:
        """The external force alone at the current state (BGGS, then
        external_force), ntotal x dim."""
        out = np.zeros((self.ntotal, self.dim))
        rc = lib().oracle_external_force(ctypes.byref(self.cfg), ctypes.byref(self._arr), _dp(out))
        if rc != 0:
            raise RuntimeError(f"oracle_external_force failed: {rc}")
        return out

    def run(self, nsteps: int):
        rc = lib().oracle_run(ctypes.byref(self.cfg), ctypes.byref(self._arr), self.step,
                              int(nsteps))
        if rc != 0:
            raise RuntimeError(f"oracle_run failed: {rc}")
        self.step += nsteps

    def __getitem__(self, k):
        return self.a[k]


def read_dump(path: str) -> dict:
    """Reader for ref_driver output (name, kind, count, data records)."""
    out = {}
    b = open(path, "rb").read()
    off = 0
    while off < len(b):
        name = b[off:off + 8].decode().strip()
        off += 8
        kind = int(np.frombuffer(b, "<i4", 1, off)[0])
        off += 4
        cnt = int(np.frombuffer(b, "<i8", 1, off)[0])
        off += 8
        dt = "<f8" if kind == 8 else "<i4"
        out[name] = np.frombuffer(b, dt, cnt, off).copy()
        off += cnt * (8 if kind == 8 else 4)
    return out
