"""Batch API on the CPU: arena sizing and a clean error without a device."""
from __future__ import annotations

import pytest

from mysph_amd import config as C
from mysph_amd.sphgpu import Batch, SphGpuError, batch_bytes


def test_arena_bytes_scale_with_slots_and_cells():
    a = batch_bytes(2, C.MIXED, 1 << 20, 1 << 20, 64)
    b = batch_bytes(2, C.MIXED, 2 << 20, 1 << 20, 64)
    c = batch_bytes(2, C.MIXED, 1 << 20, 2 << 20, 64)
    assert b - a > (1 << 20) * 300          # hundreds of bytes per slot (48 candidates alone 192)
    assert 0 < c - a < (1 << 20) * 16       # a few bytes per cell
    # FP64 keeps (rho, p/rho^2, c, m/rho) and the derivatives in doubles, mixed
    # keeps FP32 copies: different per-slot records
    assert batch_bytes(2, C.FP64, 1 << 20, 1 << 20, 64) != a
    with pytest.raises(SphGpuError):
        batch_bytes(4, C.MIXED, 10, 10, 1)


def test_create_without_device_reports_reason():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(SphGpuError) as e:
        Batch(2, C.MIXED, 1024, 1024, 4)
    assert "sphgpu_batch_create" in str(e.value)
