"""Host-side checks that need no GPU: case generators, the config struct
layout shared with the C ABI, and that libsphgpu.so loads and exports every
entry point declared in include/sphgpu.h (no compute calls)."""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np
import pytest

import oracle
from mysph_amd import cases
from mysph_amd import config as C
from mysph_amd import sphgpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    text = open(os.path.join(ROOT, "include", "sphgpu.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|void|const char\*)\s+(sphgpu_\w+)\(",
                                 text, re.M)))


def test_library_exports_every_declared_entry_point():
    lib = sphgpu.load_library()
    names = declared_entry_points()
    assert len(names) >= 12
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(sphgpu.EXPORTED)
    assert lib.sphgpu_abi_version() == 2


def test_config_struct_layout_matches_c():
    assert ctypes.sizeof(C.Config) == oracle.lib().oracle_config_size()


def test_benchmark_dam_break_counts():
    c0 = cases.dam_break_2d(dx=0.02)
    assert c0["n_real"] == 5000 and c0["n_dummy"] == 1518
    assert cases.dam_break_3d_fluid_count(0.0075) == 1024000
    c = cases.dam_break_3d(dx=0.03)
    assert c["n_real"] == 20 * 20 * 40
    # jitter stays within +-0.01 dx of the lattice
    lat = (np.floor(c["x"][: c["n_real"]] / 0.03) + 0.5) * 0.03
    assert np.max(np.abs(c["x"][: c["n_real"]] - lat)) <= 0.01 * 0.03 + 1e-15
    # fixed boundary sits outside the tank interior
    b = c["x"][c["n_real"]:]
    inside = (b[:, 0] > 0) & (b[:, 0] < 1.6) & (b[:, 1] > 0) & (b[:, 1] < 0.6) & (b[:, 2] > 0)
    assert not inside.any()


def test_jitter_is_counter_based_and_uniform():
    a = cases.jitter_unit(np.arange(100000), 3)
    b = cases.jitter_unit(np.arange(50000, 100000), 3)
    assert np.array_equal(a[50000:], b)
    assert a.min() >= -1 and a.max() < 1 and abs(a.mean()) < 0.01


def test_dt_from_courant_form():
    c = cases.dam_break_3d(dx=0.0075)
    c0 = 10 * np.sqrt(2 * 9.81 * 1.2)
    assert abs(c["config"].dt - 0.2 * 1.2 * 0.0075 / (c0 * 1.12)) < 1e-18
    assert abs(c["config"].dt - 3.312e-5) < 1e-8


def test_checkpoint_restore_refuses_foreign_blobs():
    """Header checks run before any device call: garbage and truncated
    blobs are refused with a message."""
    from mysph_amd.sphgpu import Simulation, SphGpuError
    with pytest.raises(SphGpuError, match="not a checkpoint"):
        Simulation.restore(np.zeros(4096, dtype=np.uint8))
    with pytest.raises(SphGpuError, match="short checkpoint"):
        Simulation.restore(np.zeros(8, dtype=np.uint8))
