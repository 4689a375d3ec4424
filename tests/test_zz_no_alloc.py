"""sphgpu_pairs_noalloc (the BGGS pair export with no allocation, its
scratch in a file-scope device array) against sphgpu_export_pairs, the entry
point it is a variant of: the same CSR offsets and the same ascending ids, bit
for bit, on calls in a row on one device with inputs of decreasing size, so
that each call starts from what the last one left in the shared scratch.

Device buffers come from the HIP runtime the library itself links (the
process's libamdhip64), through ctypes.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np
import pytest

from mysph_amd import cases
from mysph_amd import config as C
from mysph_amd.sphgpu import Simulation, load_library

pytestmark = pytest.mark.gpu

MAX_N = 1 << 20          # the entry point's stated maximum (real + boundary)


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libsphgpu links."""

    H2D, D2H = 1, 2

    def __init__(self):
        load_library()
        self.L = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), max(int(nbytes), 8)) == 0
        self.bufs.append(p)
        return p

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, self.H2D) == 0
        return p

    def get(self, p, count, dtype):
        out = np.zeros(count, dtype=dtype)
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, self.D2H) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


@pytest.fixture(scope="module")
def hip():
    h = Hip()
    yield h
    h.free()


def _lib():
    L = load_library()
    vp, i64 = ctypes.c_void_p, ctypes.c_int64
    L.sphgpu_pairs_noalloc.argtypes = [ctypes.c_int32, ctypes.c_int32, i64, i64, vp, vp, vp, vp, vp,
                                       i64, ctypes.POINTER(i64), vp]
    L.sphgpu_last_error.argtypes = [vp]
    L.sphgpu_last_error.restype = ctypes.c_char_p
    return L


def noalloc_pairs(hip, dim, skf, n_real, x, h, state, cap=None):
    """(rc, needed, offsets, ids) of one call; cap None: two calls, the first
    to size the id buffer."""
    L = _lib()
    n = len(h)
    xd, hd = hip.put(np.asarray(x, np.float64).ravel()), hip.put(np.asarray(h, np.float64))
    sd = hip.put(np.asarray(state, np.int32)) if state is not None else None
    off = hip.alloc(8 * (n_real + 1))
    need = ctypes.c_int64(-1)
    if cap is None:
        rc = L.sphgpu_pairs_noalloc(dim, skf, n_real, n, xd, hd, sd, off, None, 0, ctypes.byref(need), None)
        assert rc in (0, -3), L.sphgpu_last_error(None)
        cap = need.value
    nbr = hip.alloc(4 * max(cap, 1))
    rc = L.sphgpu_pairs_noalloc(dim, skf, n_real, n, xd, hd, sd, off, nbr, cap, ctypes.byref(need), None)
    o = hip.get(off, n_real + 1, np.int64)
    ids = hip.get(nbr, need.value, np.int32) if rc == 0 else None
    return rc, need.value, o, ids


def export_pairs(case):
    with Simulation(case) as sim:
        return sim.pairs()


def _custom_case(s):
    n = len(s["h"])
    case = dict(dim=s["dim"], n_real=s["ntotal"], n_dummy=n - s["ntotal"], x=s["x"], h=s["h"],
                state=s["state"], mass=np.ones(n), rho=np.full(n, 1000.0),
                type=np.where(np.arange(n) < s["ntotal"], 2, -2).astype(np.int32))
    cfg = C.base_config(s["dim"])
    cfg.skf = s["skf"]
    cfg.add_material(2, C.EOS_TAIT, 1000.0, C.tait_B(10.0), 7.0)
    case["config"] = cfg
    return case


def _inputs():
    """Decreasing sizes: 3D dam break (32k), the reference dam break (21k),
    then two seeded sets (400 with State == -1 particles and per-particle h,
    250 with the quintic kernel's 3h reach)."""
    from tests.golden.make_golden import custom_bggs_inputs
    sets = [cases.dam_break_3d(dx=0.03), cases.reference_dam_break()]
    custom = custom_bggs_inputs()
    sets += [_custom_case(custom[1]), _custom_case(custom[3])]
    return sets


def test_noalloc_equals_export_pairs_on_calls_in_a_row(hip):
    sizes = []
    for case in _inputs():
        n = case["n_real"] + case["n_dummy"]
        sizes.append(n)
        print(f"no-alloc pairs: dim {case['dim']}, {case['n_real']} real + {case['n_dummy']} boundary",
              flush=True)
        want_off, want_nbr = export_pairs(case)
        rc, need, off, ids = noalloc_pairs(hip, case["dim"], case["config"].skf, case["n_real"],
                                           case["x"], case["h"], case.get("state"))
        assert rc == 0, _lib().sphgpu_last_error(None)
        assert need == want_off[-1] == len(want_nbr)
        assert np.array_equal(off, want_off)
        assert np.array_equal(ids, want_nbr)
    assert sizes == sorted(sizes, reverse=True) and sizes[-1] < sizes[0]


def test_noalloc_reports_capacity_and_bad_arguments(hip):
    L = _lib()
    case = cases.reference_dam_break()
    nt, n = case["n_real"], case["n_real"] + case["n_dummy"]
    rc, need, off, _ = noalloc_pairs(hip, 2, 1, nt, case["x"], case["h"], case["state"], cap=1000)
    assert rc == -3 and need == 122267 and off[-1] == need      # E_PAIR_CAPACITY, sized
    null = ctypes.c_void_p(0)
    cnt = ctypes.c_int64(0)
    assert L.sphgpu_pairs_noalloc(2, 1, 10, MAX_N + 1, null, null, null, null, null, 0,
                                  ctypes.byref(cnt), None) == -7          # E_CAPACITY
    assert b"holds" in L.sphgpu_last_error(None)
    assert L.sphgpu_pairs_noalloc(4, 1, 10, 20, null, null, null, null, null, 0,
                                  ctypes.byref(cnt), None) == -6          # E_ARG


@pytest.mark.parametrize("dim", [2, 3])
def test_noalloc_exact_test_at_rounding_boundary(hip, dim):
    """Pairs whose acceptance a fused multiply-add in the squared distance
    would flip (tests.parity.rounding_boundary_set): the no-alloc export keeps
    exactly the oracle's pairs."""
    import oracle
    from tests.parity import rounding_boundary_set, sorted_pairs
    x, h, accept = rounding_boundary_set(dim)
    n = len(h)
    o_off, o_nbr, _, _, _ = oracle.bggs(dim, 1, n, x, h)
    rc, need, off, ids = noalloc_pairs(hip, dim, 1, n, x, h, None)
    assert rc == 0, _lib().sphgpu_last_error(None)
    assert need == 2 * int(accept.sum())
    assert np.array_equal(sorted_pairs(off, ids), sorted_pairs(o_off, o_nbr))
