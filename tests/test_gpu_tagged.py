"""GPU checks of the tagged entry points (ocu_grid_reduce_tagged,
ocu_grid_maxabs3_tagged and the ocu_norm*_finalize_tagged): the results go into pinned host memory and a
completion tag after them, so a host that polls the tag slot - no stream
synchronisation - finds the same values as the untagged entry points with
device results, bit for bit, for every reduction op and both precisions, on
several calls in a row with fresh tags."""
import ctypes as C
import time

import numpy as np
import pytest

from opencurrent_amd import _lib
from opencurrent_amd.device import DeviceGrid

OPS = {0: "maxabs", 1: "max", 2: "min", 3: "sum", 4: "sqrsum", 5: "checknan"}


class Pinned:
    """eight pinned host doubles and their device-side address"""

    def __init__(self, hip):
        self.hip = hip
        self.h = C.c_void_p()
        _lib.check(hip.ocu_host_malloc(C.byref(self.h), 64), "ocu_host_malloc")
        d = C.c_void_p()
        _lib.check(hip.ocu_host_device_pointer(C.byref(d), self.h), "ocu_host_device_pointer")
        self.dev = C.cast(d, _lib.DP)
        self.vals = (C.c_double * 8).from_address(self.h.value)
        self.tags = (C.c_ulonglong * 8).from_address(self.h.value)

    def wait(self, slot, tag, limit_s=10.0):
        t0 = time.monotonic()
        while self.tags[slot] != tag:
            assert time.monotonic() - t0 < limit_s, f"tag {tag} did not arrive in slot {slot}"
        return self.vals[slot - 1]

    def free(self):
        self.hip.ocu_host_free(self.h)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_tagged_reduce_matches_device_reduce(gpu, dtype):
    import torch
    hip = _lib.hip()
    pin = Pinned(hip)
    try:
        ws = torch.zeros(hip.ocu_reduce_workspace_bytes() // 8, dtype=torch.float64, device="cuda")
        out = torch.full((1,), np.nan, dtype=torch.float64, device="cuda")
        gen = torch.Generator(device="cuda").manual_seed(7101)
        tag = 1000
        for n in [(64, 48, 130), (9, 17, 33)]:
            g = DeviceGrid(*n, dtype=dtype)
            g.tensor.uniform_(-1, 1, generator=gen)
            for op in sorted(OPS):
                print(f"tagged reduce op={OPS[op]} dtype={np.dtype(dtype).name} grid={n}", flush=True)
                _lib.check(hip.ocu_grid_reduce(g.ptr(), op, ws.data_ptr(), C.cast(out.data_ptr(), _lib.DP), None),
                           "ocu_grid_reduce")
                torch.cuda.synchronize()
                want = out.item()
                tag += 1
                pin.vals[0] = np.nan
                _lib.check(hip.ocu_grid_reduce_tagged(g.ptr(), op, ws.data_ptr(), pin.dev, tag, None),
                           "ocu_grid_reduce_tagged")
                got = pin.wait(1, tag)
                assert np.array_equal(np.float64(got), np.float64(want)), (OPS[op], n, got, want)
        torch.cuda.synchronize()
    finally:
        pin.free()


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_maxabs3_matches_three_device_reduces(gpu, dtype):
    """ocu_grid_maxabs3_tagged: the three maxima of one pass equal three
    ocu_grid_reduce(op 0) calls bit for bit, on MAC-shaped grids of
    different extents (u: nx+1, v: ny+1, w: nz+1), where the largest
    interior magnitude sits in a different grid each call and a larger one
    in a ghost is left out (the first size has more rows than the walk has
    waves), and on smaller calls after it"""
    import torch
    hip = _lib.hip()
    pin = Pinned(hip)
    try:
        ws = torch.zeros(3 * hip.ocu_reduce_workspace_bytes() // 8, dtype=torch.float64, device="cuda")
        out = torch.full((1,), np.nan, dtype=torch.float64, device="cuda")
        gen = torch.Generator(device="cuda").manual_seed(7301)
        tag = 9000
        for n in [(100, 96, 40), (64, 48, 130), (9, 17, 33)]:
            nx, ny, nz = n
            gs = [DeviceGrid(nx + 1, ny, nz, dtype=dtype), DeviceGrid(nx, ny + 1, nz, dtype=dtype),
                  DeviceGrid(nx, ny, nz + 1, dtype=dtype)]
            for big in range(3):
                for q, g in enumerate(gs):
                    g.tensor.uniform_(-1, 1, generator=gen)
                    g.full()[0, 0, 0] = 7.0  # a ghost corner: outside the interior
                    if q == big:
                        g.interior()[nx // 2, ny // 3, nz - 1] = -3.5
                print(f"maxabs3 dtype={np.dtype(dtype).name} grid={n} largest in {big}", flush=True)
                want = []
                for g in gs:
                    _lib.check(hip.ocu_grid_reduce(g.ptr(), 0, ws.data_ptr(), C.cast(out.data_ptr(), _lib.DP), None),
                               "ocu_grid_reduce")
                    torch.cuda.synchronize()
                    want.append(out.item())
                tag += 1
                for q in range(4):
                    pin.vals[q] = np.nan
                _lib.check(hip.ocu_grid_maxabs3_tagged(gs[0].ptr(), gs[1].ptr(), gs[2].ptr(), ws.data_ptr(),
                                                       pin.dev, tag, None), "ocu_grid_maxabs3_tagged")
                pin.wait(3, tag)
                got = [pin.vals[q] for q in range(3)]
                assert np.array_equal(np.array(got), np.array(want)), (n, big, got, want)
                assert max(want) == 3.5 and want[big] == 3.5
        torch.cuda.synchronize()
    finally:
        pin.free()


@pytest.mark.gpu
def test_tagged_entry_points_reject_a_zero_tag(gpu):
    import torch
    hip = _lib.hip()
    g = DeviceGrid(8, 8, 8, dtype=np.float64)
    ws = torch.zeros(hip.ocu_reduce_workspace_bytes() // 8, dtype=torch.float64, device="cuda")
    out = torch.zeros((4,), dtype=torch.float64, device="cuda")
    dp = C.cast(out.data_ptr(), _lib.DP)
    assert hip.ocu_grid_reduce_tagged(g.ptr(), 0, ws.data_ptr(), dp, 0, None) != 0
    assert hip.ocu_grid_maxabs3_tagged(g.ptr(), g.ptr(), g.ptr(), ws.data_ptr(), dp, 0, None) != 0
    wsp = C.cast(ws.data_ptr(), _lib.DP)
    assert hip.ocu_norm_finalize_tagged(wsp, dp, 0, None) != 0
    assert hip.ocu_norm3_finalize_tagged(wsp, dp, 0, None) != 0
    assert hip.ocu_norm4_finalize_tagged(wsp, dp, 0, None) != 0


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_tagged_norm_finalizes_match_untagged(gpu, dtype):
    """ocu_norm_finalize_tagged / ocu_norm3_finalize_tagged /
    ocu_norm4_finalize_tagged: the results polled from pinned host memory
    equal the untagged finalizes' (device memory), bit for bit, and the tag
    lands in the slot after them"""
    import torch
    hip = _lib.hip()
    pin = Pinned(hip)
    try:
        parts = hip.ocu_norm_parts()
        ws = torch.zeros(4 * parts + 64, dtype=torch.float64, device="cuda")
        wsp = C.cast(ws.data_ptr(), _lib.DP)
        out = torch.full((8,), np.nan, dtype=torch.float64, device="cuda")
        outp = C.cast(out.data_ptr(), _lib.DP)
        gen = torch.Generator(device="cuda").manual_seed(7201)
        n = (40, 36, 70)
        u = DeviceGrid(*n, dtype=dtype)
        b = DeviceGrid(*n, dtype=dtype)
        r = DeviceGrid(*n, dtype=dtype)
        for g in (u, b):
            g.tensor.uniform_(-1, 1, generator=gen)
        h = 1.0 / 64
        tag = 5000

        def untagged(fn, k):
            _lib.check(fn(wsp, outp, None), "finalize")
            torch.cuda.synchronize()
            return out[:k].cpu().numpy().copy()

        def tagged(fn, k):
            nonlocal tag
            tag += 1
            for q in range(8):
                pin.vals[q] = np.nan
            _lib.check(fn(wsp, pin.dev, tag, None), "finalize_tagged")
            pin.wait(k, tag)
            return np.array([pin.vals[q] for q in range(k)])

        print(f"tagged norm finalize dtype={np.dtype(dtype).name}", flush=True)
        _lib.check(hip.ocu_mg_residual_norm(u.ptr(), b.ptr(), r.ptr(), h, 1.0, 1.0, 1.0, wsp, None), "norm")
        want = untagged(hip.ocu_norm_finalize, 2)
        got = tagged(hip.ocu_norm_finalize_tagged, 2)
        assert np.array_equal(got, want) and np.isfinite(want).all(), (got, want)
        if dtype == np.float64:
            print("tagged norm3/norm4 finalize", flush=True)
            _lib.check(hip.ocu_grid_stats3(u.ptr(), wsp, None), "stats3")
            want = untagged(hip.ocu_norm3_finalize, 3)
            got = tagged(hip.ocu_norm3_finalize_tagged, 3)
            assert np.array_equal(got, want) and np.isfinite(want).all(), (got, want)
            r32 = DeviceGrid(*n, dtype=np.float32)
            _lib.check(hip.ocu_ir_residual(u.ptr(), b.ptr(), 0.25, r32.ptr(), 0.5, h, 1.0, 1.0, 1.0, wsp, None),
                       "ir_residual")
            want = untagged(hip.ocu_norm4_finalize, 4)
            got = tagged(hip.ocu_norm4_finalize_tagged, 4)
            assert np.array_equal(got, want) and np.isfinite(want).all(), (got, want)
        torch.cuda.synchronize()
    finally:
        pin.free()
