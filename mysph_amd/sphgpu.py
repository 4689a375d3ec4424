"""Host-side binding of libsphgpu (include/sphgpu.h) and a Simulation object
that mirrors the reference's step seam.

The reference drives its hot path through ``single_step`` (called once per
step by ``leap_frog``, src/TimeIntegration/timeIntegration.f90:172-176). Here:

  ``Simulation.derivatives()``  one ``single_step`` evaluation at the resident
                                state (singleStep.f90:58-144)
  ``Simulation.advance(n)``     n leap-frog steps on the device
                                (timeIntegration.f90:137-252)
  ``Simulation.get(field)``     read a field back in the caller's particle order
  ``Simulation.pairs()``        the BGGS neighbour lists (asymmetricSearch.f90:160-257)

Errors come back as ``SphGpuError`` with the library's message; the reference
aborts with ``error stop`` in the same situations.

There is no CPU fallback: without the built library (``python -m
mysph_amd.build``) every call raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import config as C

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsphgpu.so")

STATUS = {
    0: "ok", -1: "E_OUT_OF_RANGE", -2: "E_NONFINITE", -3: "E_PAIR_CAPACITY",
    -4: "E_BAD_CONFIG", -5: "E_HIP", -6: "E_ARG", -7: "E_CAPACITY",
}

EXPORTED = (
    "sphgpu_abi_version", "sphgpu_create", "sphgpu_upload", "sphgpu_derivatives",
    "sphgpu_advance", "sphgpu_download", "sphgpu_upload_field", "sphgpu_export_pairs",
    "sphgpu_pairs_noalloc", "sphgpu_grid_info",
    "sphgpu_timing", "sphgpu_step", "sphgpu_last_error", "sphgpu_destroy",
    "sphgpu_neighbour_stats", "sphgpu_checkpoint_size", "sphgpu_checkpoint_save",
    "sphgpu_checkpoint_restore", "sphgpu_context_info",
    "sphgpu_batch_bytes", "sphgpu_batch_create", "sphgpu_batch_add", "sphgpu_batch_advance",
    "sphgpu_batch_status", "sphgpu_batch_run_error", "sphgpu_batch_download", "sphgpu_batch_remove",
    "sphgpu_batch_info", "sphgpu_batch_timing", "sphgpu_batch_last_error", "sphgpu_batch_destroy",
)

# sphgpu_batch_status states (negative values are the failure's sphgpu_status)
RUNNING, FINISHED = 0, 1


class SphGpuError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libsphgpu.so (no fallback if it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libsphgpu.so not built ({path}); run python -m mysph_amd.build")
    L = ctypes.CDLL(path)
    vp, dp, i64 = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int64
    ip = ctypes.POINTER(ctypes.c_int32)
    L.sphgpu_abi_version.restype = ctypes.c_int
    L.sphgpu_create.argtypes = [ctypes.POINTER(C.Config), i64, ctypes.POINTER(vp)]
    L.sphgpu_upload.argtypes = [vp, i64, i64, dp, dp, dp, dp, dp, dp, dp, dp, dp, ip, ip, ip]
    L.sphgpu_derivatives.argtypes = [vp]
    L.sphgpu_advance.argtypes = [vp, ctypes.c_int32]
    L.sphgpu_download.argtypes = [vp, ctypes.c_int32, dp]
    L.sphgpu_upload_field.argtypes = [vp, ctypes.c_int32, dp]
    L.sphgpu_pairs_noalloc.argtypes = [ctypes.c_int32, ctypes.c_int32, i64, i64, vp, vp, vp, vp, vp,
                                       i64, ctypes.POINTER(i64), vp]
    L.sphgpu_export_pairs.argtypes = [vp, ctypes.POINTER(i64), ip, i64, ctypes.POINTER(i64)]
    L.sphgpu_grid_info.argtypes = [vp, dp]
    L.sphgpu_neighbour_stats.argtypes = [vp, ctypes.POINTER(i64)]
    L.sphgpu_timing.argtypes = [vp, dp, dp, ctypes.POINTER(i64)]
    L.sphgpu_step.argtypes = [vp]
    L.sphgpu_step.restype = i64
    L.sphgpu_last_error.argtypes = [vp]
    L.sphgpu_last_error.restype = ctypes.c_char_p
    L.sphgpu_destroy.argtypes = [vp]
    L.sphgpu_destroy.restype = None
    L.sphgpu_checkpoint_size.argtypes = [vp, ctypes.POINTER(i64)]
    L.sphgpu_checkpoint_save.argtypes = [vp, vp, i64]
    L.sphgpu_checkpoint_restore.argtypes = [vp, i64, ctypes.POINTER(vp)]
    L.sphgpu_context_info.argtypes = [vp, ctypes.POINTER(C.Config), ctypes.POINTER(i64)]
    p32 = ctypes.POINTER(ctypes.c_int32)
    p64 = ctypes.POINTER(i64)
    L.sphgpu_batch_bytes.argtypes = [ctypes.c_int32, ctypes.c_int32, i64, i64, ctypes.c_int32,
                                     ctypes.c_int32, p64]
    L.sphgpu_batch_create.argtypes = [ctypes.c_int32, ctypes.c_int32, i64, i64, ctypes.c_int32,
                                      ctypes.c_int32, ctypes.POINTER(vp)]
    L.sphgpu_batch_add.argtypes = [vp, ctypes.POINTER(C.Config), i64, i64, dp, dp, dp, dp, dp, dp,
                                   dp, dp, ip, ip, ip, i64, ctypes.c_int32, ctypes.c_double, p32]
    L.sphgpu_batch_advance.argtypes = [vp, ctypes.c_int32]
    L.sphgpu_batch_status.argtypes = [vp, ctypes.c_int32, p64, p32]
    L.sphgpu_batch_run_error.argtypes = [vp, ctypes.c_int32]
    L.sphgpu_batch_run_error.restype = ctypes.c_char_p
    L.sphgpu_batch_download.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, dp]
    L.sphgpu_batch_remove.argtypes = [vp, ctypes.c_int32]
    L.sphgpu_batch_info.argtypes = [vp, p64]
    L.sphgpu_batch_timing.argtypes = [vp, dp, p64, p64]
    L.sphgpu_batch_last_error.argtypes = [vp]
    L.sphgpu_batch_last_error.restype = ctypes.c_char_p
    L.sphgpu_batch_destroy.argtypes = [vp]
    L.sphgpu_batch_destroy.restype = None
    if L.sphgpu_abi_version() != 2:
        raise RuntimeError("libsphgpu ABI version mismatch")
    _lib = L
    return L


def _dp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def _f64(a, shape=None):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if shape is None else a.reshape(shape)


class Simulation:
    """One run resident on the current GPU (one context, one stream)."""

    def __init__(self, case: dict, cfg: C.Config | None = None, capacity: int | None = None):
        L = load_library()
        self._L = L
        self.cfg = (cfg if cfg is not None else case["config"]).copy()
        self.dim = int(case["dim"])
        if int(self.cfg.dim) != self.dim:
            raise ValueError("config dim does not match the case")
        self.n_real = int(case["n_real"])
        self.n_dummy = int(case["n_dummy"])
        self.n = self.n_real + self.n_dummy
        cap = int(capacity or self.n)
        self._ctx = ctypes.c_void_p()
        rc = L.sphgpu_create(ctypes.byref(self.cfg), cap, ctypes.byref(self._ctx))
        if rc != 0:
            raise SphGpuError(rc, "sphgpu_create failed (bad config or no device)")
        self.upload(case)

    def _check(self, rc: int):
        if rc != 0:
            msg = self._L.sphgpu_last_error(self._ctx)
            raise SphGpuError(rc, msg.decode() if msg else "")

    def upload(self, case: dict):
        d = self.dim
        self._keep = dict(
            x=_f64(case["x"]), v=_f64(case.get("v")), mass=_f64(case["mass"]),
            rho=_f64(case["rho"]), p=_f64(case.get("p")), e=_f64(case.get("e")),
            c=_f64(case.get("c")), h=_f64(case["h"]), shear=_f64(case.get("shear")),
            type=np.ascontiguousarray(case["type"], dtype=np.int32),
            state=np.ascontiguousarray(case.get("state", np.zeros(self.n)), dtype=np.int32),
            boundary=np.ascontiguousarray(case.get("boundary", np.zeros(self.n)), dtype=np.int32),
        )
        k = self._keep
        if k["x"].size != self.n * d:
            raise ValueError("x has the wrong size")
        self._check(self._L.sphgpu_upload(
            self._ctx, self.n_real, self.n_dummy, _dp(k["x"]), _dp(k["v"]), _dp(k["mass"]),
            _dp(k["rho"]), _dp(k["p"]), _dp(k["e"]), _dp(k["c"]), _dp(k["h"]), _dp(k["shear"]),
            _ip(k["type"]), _ip(k["state"]), _ip(k["boundary"])))

    def derivatives(self):
        self._check(self._L.sphgpu_derivatives(self._ctx))

    def advance(self, nsteps: int):
        self._check(self._L.sphgpu_advance(self._ctx, int(nsteps)))

    @property
    def step(self) -> int:
        return int(self._L.sphgpu_step(self._ctx))

    def get(self, name: str) -> np.ndarray:
        f = C.FIELD_NAMES[name]
        comp = C.field_components(f, self.dim)
        out = np.zeros(self.n * comp)
        self._check(self._L.sphgpu_download(self._ctx, f, _dp(out)))
        if comp == 1:
            return out
        if comp == self.dim:
            return out.reshape(self.n, self.dim)
        return out.reshape(self.n, self.dim, self.dim)

    def put(self, name: str, values):
        """Overwrite one field of the resident state (in the caller's
        particle order); the divergences only (sphgpu_upload_field)."""
        f = C.FIELD_NAMES[name]
        a = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        if a.size != self.n * C.field_components(f, self.dim):
            raise ValueError(f"{name} has the wrong size")
        self._check(self._L.sphgpu_upload_field(self._ctx, f, _dp(a)))

    def pairs(self):
        """(offsets[n_real+1], nbr) of the current state, ids ascending per target."""
        off = np.zeros(self.n_real + 1, dtype=np.int64)
        need = ctypes.c_int64(0)
        p64 = ctypes.POINTER(ctypes.c_int64)
        rc = self._L.sphgpu_export_pairs(self._ctx, off.ctypes.data_as(p64), None, 0,
                                         ctypes.byref(need))
        if rc not in (0, -3):
            self._check(rc)
        nbr = np.zeros(max(need.value, 1), dtype=np.int32)
        self._check(self._L.sphgpu_export_pairs(self._ctx, off.ctypes.data_as(p64), _ip(nbr),
                                                need.value, ctypes.byref(need)))
        return off, nbr[:need.value]

    def neighbour_stats(self) -> dict:
        """Statistics of the last neighbour search (nnps.f90:377-416)."""
        out = np.zeros(4, dtype=np.int64)
        self._check(self._L.sphgpu_neighbour_stats(
            self._ctx, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return dict(pairs=int(out[0]), max=int(out[1]), min=int(out[2]), zero=int(out[3]))

    def grid_info(self) -> np.ndarray:
        out = np.zeros(9)
        self._check(self._L.sphgpu_grid_info(self._ctx, _dp(out)))
        return out

    def timing(self):
        a, f, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        self._check(self._L.sphgpu_timing(self._ctx, ctypes.byref(a), ctypes.byref(f),
                                          ctypes.byref(n)))
        return a.value, f.value, n.value

    def checkpoint(self) -> np.ndarray:
        """Exact checkpoint of the run (bytes as a uint8 array)."""
        nb = ctypes.c_int64(0)
        self._check(self._L.sphgpu_checkpoint_size(self._ctx, ctypes.byref(nb)))
        buf = np.empty(nb.value, dtype=np.uint8)
        self._check(self._L.sphgpu_checkpoint_save(self._ctx, buf.ctypes.data, nb.value))
        return buf

    @classmethod
    def restore(cls, blob) -> "Simulation":
        """A run restored from ``checkpoint()``: continues bit for bit."""
        L = load_library()
        blob = np.ascontiguousarray(blob, dtype=np.uint8)
        self = cls.__new__(cls)
        self._L = L
        self._ctx = ctypes.c_void_p()
        rc = L.sphgpu_checkpoint_restore(blob.ctypes.data, blob.size, ctypes.byref(self._ctx))
        if rc != 0:
            msg = L.sphgpu_last_error(None)
            raise SphGpuError(rc, msg.decode() if msg else "")
        self.cfg = C.Config()
        counts = np.zeros(3, dtype=np.int64)
        self._check(L.sphgpu_context_info(self._ctx, ctypes.byref(self.cfg),
                                          counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        self.dim = int(self.cfg.dim)
        self.n_real, self.n_dummy = int(counts[0]), int(counts[1])
        self.n = self.n_real + self.n_dummy
        self._keep = {}
        return self

    def close(self):
        if getattr(self, "_ctx", None) and self._ctx.value:
            self._L.sphgpu_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def case_arrays(case: dict) -> dict:
    """Contiguous host arrays of a case, in the layout the C ABI takes."""
    n = int(case["n_real"]) + int(case["n_dummy"])
    return dict(
        x=_f64(case["x"]), v=_f64(case.get("v")), mass=_f64(case["mass"]),
        rho=_f64(case["rho"]), p=_f64(case.get("p")), e=_f64(case.get("e")),
        c=_f64(case.get("c")), h=_f64(case["h"]),
        type=np.ascontiguousarray(case["type"], dtype=np.int32),
        state=np.ascontiguousarray(case.get("state", np.zeros(n)), dtype=np.int32),
        boundary=np.ascontiguousarray(case.get("boundary", np.zeros(n)), dtype=np.int32),
    )


def batch_bytes(dim: int, precision: int, slots: int, cells: int, max_runs: int,
                verlet_cap: int = 0) -> int:
    L = load_library()
    out = ctypes.c_int64(0)
    rc = L.sphgpu_batch_bytes(dim, precision, slots, cells, max_runs, verlet_cap, ctypes.byref(out))
    if rc != 0:
        raise SphGpuError(rc, "bad batch dimensions")
    return out.value


class Batch:
    """Many runs resident in one arena on the current GPU, advanced together
    (sphgpu_batch_*). Runs are admitted with add() (False when the arena has
    no room), advanced with advance(), read with get() and released with
    remove()."""

    def __init__(self, dim: int, precision: int, slots: int, cells: int, max_runs: int = 256,
                 verlet_cap: int = 0):
        L = load_library()
        self._L = L
        self.dim = int(dim)
        self.precision = int(precision)
        self._b = ctypes.c_void_p()
        rc = L.sphgpu_batch_create(self.dim, self.precision, int(slots), int(cells), int(max_runs),
                                   int(verlet_cap), ctypes.byref(self._b))
        if rc != 0:
            msg = L.sphgpu_batch_last_error(None)
            raise SphGpuError(rc, msg.decode() if msg else "sphgpu_batch_create failed")
        self._n = {}

    def _check(self, rc: int):
        if rc != 0:
            msg = self._L.sphgpu_batch_last_error(self._b)
            raise SphGpuError(rc, msg.decode() if msg else "")

    def add(self, case: dict, nsteps: int, cfg: C.Config | None = None, win: int = 0,
            skin_frac: float = 0.0):
        """Admit a run; returns its handle, or None if the batch has no room."""
        cfg = (cfg if cfg is not None else case["config"]).copy()
        k = case_arrays(case)
        n_real, n_dummy = int(case["n_real"]), int(case["n_dummy"])
        run = ctypes.c_int32(-1)
        rc = self._L.sphgpu_batch_add(
            self._b, ctypes.byref(cfg), n_real, n_dummy, _dp(k["x"]), _dp(k["v"]), _dp(k["mass"]),
            _dp(k["rho"]), _dp(k["p"]), _dp(k["e"]), _dp(k["c"]), _dp(k["h"]), _ip(k["type"]),
            _ip(k["state"]), _ip(k["boundary"]), int(nsteps), int(win), float(skin_frac),
            ctypes.byref(run))
        if rc == -7:   # SPHGPU_E_CAPACITY
            return None
        self._check(rc)
        self._n[run.value] = n_real + n_dummy
        return run.value

    def advance(self, nsteps: int):
        self._check(self._L.sphgpu_batch_advance(self._b, int(nsteps)))

    def status(self, run: int):
        """(next step, state): state 0 running, 1 finished, < 0 failed."""
        step, state = ctypes.c_int64(), ctypes.c_int32()
        self._check(self._L.sphgpu_batch_status(self._b, int(run), ctypes.byref(step),
                                                ctypes.byref(state)))
        return step.value, state.value

    def run_error(self, run: int) -> str:
        m = self._L.sphgpu_batch_run_error(self._b, int(run))
        return m.decode() if m else ""

    def get(self, run: int, name: str) -> np.ndarray:
        f = C.FIELD_NAMES[name]
        comp = C.field_components(f, self.dim)
        n = self._n[run]
        out = np.zeros(n * comp)
        self._check(self._L.sphgpu_batch_download(self._b, int(run), f, _dp(out)))
        if comp == 1:
            return out
        if comp == self.dim:
            return out.reshape(n, self.dim)
        return out.reshape(n, self.dim, self.dim)

    def remove(self, run: int):
        self._check(self._L.sphgpu_batch_remove(self._b, int(run)))
        self._n.pop(run, None)

    def info(self) -> dict:
        out = np.zeros(8, dtype=np.int64)
        self._check(self._L.sphgpu_batch_info(self._b, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        keys = ("runs", "runnable", "slots_used", "slots", "cells_used", "cells", "max_runs",
                "verlet_cap")
        return {k: int(v) for k, v in zip(keys, out)}

    def timing(self):
        """(ms, steps, particle_steps) of the last advance."""
        ms, st, ps = ctypes.c_double(), ctypes.c_int64(), ctypes.c_int64()
        self._check(self._L.sphgpu_batch_timing(self._b, ctypes.byref(ms), ctypes.byref(st),
                                                ctypes.byref(ps)))
        return ms.value, st.value, ps.value

    def close(self):
        if getattr(self, "_b", None) and self._b.value:
            self._L.sphgpu_batch_destroy(self._b)
            self._b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
