"""GPU check of ocu_grid_reduce_noalloc, the reduction entry point that
allocates nothing (its partials live in a library-owned device array): for
every reduction op and both precisions, two calls in a row - a large grid,
then a smaller one - each equal bit for bit to ocu_grid_reduce with a
caller-provided workspace on the same grid."""
import ctypes as C

import numpy as np
import pytest

from opencurrent_amd import _lib
from opencurrent_amd.device import DeviceGrid

OPS = {0: "maxabs", 1: "max", 2: "min", 3: "sum", 4: "sqrsum", 5: "checknan"}


def _reduce_with_workspace(hip, torch, g, op):
    ws = torch.zeros(hip.ocu_reduce_workspace_bytes() // 8, dtype=torch.float64, device="cuda")
    out = torch.full((1,), np.nan, dtype=torch.float64, device="cuda")
    _lib.check(hip.ocu_grid_reduce(g.ptr(), op, ws.data_ptr(), C.cast(out.data_ptr(), _lib.DP), None),
               "ocu_grid_reduce")
    torch.cuda.synchronize()
    return out.item()


def _reduce_noalloc(hip, torch, g, op):
    out = torch.full((1,), np.nan, dtype=torch.float64, device="cuda")
    _lib.check(hip.ocu_grid_reduce_noalloc(g.ptr(), op, C.cast(out.data_ptr(), _lib.DP), None),
               "ocu_grid_reduce_noalloc")
    torch.cuda.synchronize()
    return out.item()


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("op", sorted(OPS))
def test_noalloc_reduce_matches_workspace_reduce(gpu, dtype, op):
    import torch
    hip = _lib.hip()
    sizes = [(96, 80, 130), (17, 9, 33)]   # second call smaller than the first
    gen = torch.Generator(device="cuda").manual_seed(7001 + op)
    for n in sizes:
        print(f"noalloc reduce op={OPS[op]} dtype={np.dtype(dtype).name} grid={n}", flush=True)
        g = DeviceGrid(*n, dtype=dtype)
        g.tensor.uniform_(-1, 1, generator=gen)
        if op == 5:
            # checknan reports the first non-finite value it meets
            g.interior()[n[0] // 2, n[1] // 3, n[2] - 1] = float("inf")
        want = _reduce_with_workspace(hip, torch, g, op)
        got = _reduce_noalloc(hip, torch, g, op)
        assert np.array_equal(np.float64(got), np.float64(want)), (OPS[op], n, got, want)
        if op in (0, 1, 2):
            host = g.to_host().interior().astype(np.float64)
            ref = {0: np.abs(host).max(), 1: host.max(), 2: host.min()}[op]
            assert got == ref


@pytest.mark.gpu
def test_noalloc_reduce_rejects_bad_arguments_without_launching(gpu):
    hip = _lib.hip()
    out = C.cast(16, _lib.DP)
    assert hip.ocu_grid_reduce_noalloc(None, 0, out, None) == -1
    g = DeviceGrid(8, 8, 8, dtype=np.float64)
    assert hip.ocu_grid_reduce_noalloc(g.ptr(), 9, out, None) == -1
    assert "bad op" in _lib.last_error()
    assert hip.ocu_grid_reduce_noalloc(g.ptr(), 0, None, None) == -1
