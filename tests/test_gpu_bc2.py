"""Two-layer ghost fills and point sampling on the GPU, bit-exact against
the oracle (grid3dboundary.cu:122-232, 350-549, 904-1016;
grid3dsample.cu:24-263): whole buffers for the fills, every sampled value
(including positions outside the box, wrapped or clamped) for sampling."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from fields import uniform
from opencurrent_amd import _lib
from opencurrent_amd import _types as T
from opencurrent_amd._types import BC, BCSet, HostGrid
from opencurrent_amd.device import DeviceGrid

pytestmark = pytest.mark.gpu

CELL_BCS = {
    "neumann_vals": lambda: BCSet(BC(T.BC_NEUMANN, .5, 0), BC(T.BC_NEUMANN, -.25, 0), BC(T.BC_NEUMANN, 1.5, 0),
                                  BC(T.BC_NEUMANN, 0, 0), BC(T.BC_NEUMANN, .125, 0), BC(T.BC_NEUMANN, -2, 0)),
    "dirichlet_none": lambda: BCSet(BC(T.BC_DIRICHELET, 1, 0), BC(T.BC_DIRICHELET, -.3, 0), BC(T.BC_NONE, 0, 0),
                                    BC(T.BC_DIRICHELET, .7, 0), BC(T.BC_NONE, 0, 0), BC(T.BC_NEUMANN, .3, 0)),
    "periodic": lambda: BCSet.uniform(T.BC_PERIODIC),
    "periodic_x_dirichlet": lambda: BCSet(BC(T.BC_PERIODIC, 0, 0), BC(T.BC_PERIODIC, 0, 0),
                                          BC(T.BC_DIRICHELET, .2, 0), BC(T.BC_NEUMANN, .1, 0),
                                          BC(T.BC_PERIODIC, 0, 0), BC(T.BC_PERIODIC, 0, 0)),
}


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("kind", sorted(CELL_BCS))
@pytest.mark.parametrize("n", [(5, 7, 9), (16, 3, 33), (2, 2, 2)])
def test_level2_nocorners_matches_oracle(gpu, n, kind, dtype):
    print(f"level2 nocorners n={n} {kind} {np.dtype(dtype).name}", flush=True)
    bc = CELL_BCS[kind]()
    g = HostGrid(*n, 2, 2, 2, 1, 0, 3, dtype=dtype)
    g.buf[:] = 2 * uniform(9990, g.buf.shape) - 1
    h = (.3, .7, 1.9)
    dg = DeviceGrid.from_host(g)
    assert O.lib().orc_bc_fill_level2_nocorners(g.ptr(), C.byref(bc), *h) == 1
    _lib.check(_lib.hip().ocu_bc_fill_level2_nocorners(dg.ptr(), C.byref(bc), *h, None), "level2")
    assert np.array_equal(dg.to_host().buf, g.buf)


MAC_BCS = {
    "closed": lambda: BCSet(BC(T.BC_FORCED_INFLOW_VARIABLE_SLIP, .5, 1), BC(T.BC_FORCED_INFLOW_VARIABLE_SLIP, -.25, 0),
                            BC(T.BC_FORCED_INFLOW_VARIABLE_SLIP, 0, .5), BC(T.BC_FORCED_INFLOW_VARIABLE_SLIP, 0, 1),
                            BC(T.BC_FORCED_INFLOW_VARIABLE_SLIP, .1, 0), BC(T.BC_FORCED_INFLOW_VARIABLE_SLIP, 0, 1)),
    "periodic": lambda: BCSet.uniform(T.BC_PERIODIC),
    "periodic_xy": lambda: BCSet(BC(T.BC_PERIODIC, 0, 0), BC(T.BC_PERIODIC, 0, 0), BC(T.BC_PERIODIC, 0, 0),
                                 BC(T.BC_PERIODIC, 0, 0), BC(T.BC_FORCED_INFLOW_VARIABLE_SLIP, 0, 1),
                                 BC(T.BC_NONE, 0, 0)),
}


def mac2(n, dtype):
    nx, ny, nz = n
    return (HostGrid(nx + 1, ny, nz, 2, 2, 2, 0, 1, 1, dtype=dtype),
            HostGrid(nx, ny + 1, nz, 2, 2, 2, 1, 0, 1, dtype=dtype),
            HostGrid(nx, ny, nz + 1, 2, 2, 2, 1, 1, 0, dtype=dtype))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("kind", sorted(MAC_BCS))
@pytest.mark.parametrize("n", [(6, 5, 7), (17, 4, 40)])
def test_mac_level2_matches_oracle(gpu, n, kind, dtype):
    print(f"mac level2 n={n} {kind} {np.dtype(dtype).name}", flush=True)
    bc = MAC_BCS[kind]()
    grids = mac2(n, dtype)
    for q, g in enumerate(grids):
        g.buf[:] = 2 * uniform(9991 + q, g.buf.shape) - 1
    dev = [DeviceGrid.from_host(g) for g in grids]
    assert O.lib().orc_bc_fill_mac_level2(*(g.ptr() for g in grids), C.byref(bc)) == 1
    _lib.check(_lib.hip().ocu_bc_fill_mac_level2(*(d.ptr() for d in dev), C.byref(bc), None), "mac level2")
    for d, g in zip(dev, grids):
        assert np.array_equal(d.to_host().buf, g.buf)


SAMPLE_BCS = {
    "neumann": lambda: BCSet(BC(T.BC_NEUMANN, .5, 0), BC(T.BC_NEUMANN, -.25, 0), BC(T.BC_NEUMANN, 1.5, 0),
                             BC(T.BC_NEUMANN, 0, 0), BC(T.BC_NEUMANN, .125, 0), BC(T.BC_NEUMANN, -2, 0)),
    "periodic": lambda: BCSet.uniform(T.BC_PERIODIC),
    "mixed": lambda: BCSet(BC(T.BC_PERIODIC, 0, 0), BC(T.BC_PERIODIC, 0, 0), BC(T.BC_DIRICHELET, 0, 0),
                           BC(T.BC_SCALAR_SLIP, 0, 0), BC(T.BC_NEUMANN, .75, 0), BC(T.BC_DIRICHELET, 1, 0)),
}


def device_floats(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("kind", sorted(SAMPLE_BCS))
def test_sample_points_matches_oracle(gpu, kind, dtype):
    import torch
    print(f"sample points {kind} {np.dtype(dtype).name}", flush=True)
    n = (12, 7, 20)
    h = (.125, .25, .0625)
    bc = SAMPLE_BCS[kind]()
    g = HostGrid(*n, dtype=dtype)
    g.interior()[:] = 2 * uniform(9995, n) - 1
    O.lib().orc_bc_fill(0, g.ptr(), C.byref(bc), *h)
    r = np.random.default_rng(9996)
    m = 5000
    L = [nn * hh for nn, hh in zip(n, h)]
    # positions inside, outside (clamped / wrapped) and on the cell lattice
    pos = [r.uniform(-.6 * l, 1.6 * l, m).astype(np.float32) for l in L]
    for p, hh in zip(pos, h):
        p[:200] = (np.round(p[:200] / hh) * hh).astype(np.float32)
    for period in ((12.0, 7.0, 20.0), (11.5, 6.0, 19.0)):
        want = np.zeros(m, dtype=dtype)
        ok = O.lib().orc_sample_points(want.ctypes.data, *(p.ctypes.data for p in pos), m, g.ptr(), C.byref(bc),
                                       *period, *h, .5, .5, .5)
        assert ok == 1
        dg = DeviceGrid.from_host(g)
        dp = [device_floats(p) for p in pos]
        out = torch.zeros(m, dtype=torch.float64 if dtype == np.float64 else torch.float32, device="cuda")
        _lib.check(_lib.hip().ocu_sample_points(out.data_ptr(), *(d.data_ptr() for d in dp), m, dg.ptr(), C.byref(bc),
                                                *period, *h, .5, .5, .5, None), "sample")
        got = out.cpu().numpy()
        assert np.array_equal(got, want), (period, np.abs(got - want).max())


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("kind", ["periodic", "closed"])
def test_sample_points_mac_matches_oracle(gpu, kind, dtype):
    import torch
    from fields import mac_grids
    n = (10, 6, 8)
    h = (.1, .2, .125)
    bc = BCSet.uniform(T.BC_PERIODIC) if kind == "periodic" else BCSet.uniform(T.BC_FORCED_INFLOW_VARIABLE_SLIP, 0, 1)
    grids = mac_grids(*n, dtype=dtype)[:3]
    for q, g in enumerate(grids):
        g.interior()[:] = 2 * uniform(9997 + q, g.interior().shape) - 1
    O.lib().orc_bc_fill_mac(*(g.ptr() for g in grids), C.byref(bc))
    r = np.random.default_rng(9998)
    m = 3000
    pos = [r.uniform(-.3 * nn * hh, 1.3 * nn * hh, m).astype(np.float32) for nn, hh in zip(n, h)]
    want = [np.zeros(m, dtype=dtype) for _ in range(3)]
    assert O.lib().orc_sample_points_mac(*(w.ctypes.data for w in want), *(p.ctypes.data for p in pos), m,
                                         *(g.ptr() for g in grids), C.byref(bc), *h) == 1
    dev = [DeviceGrid.from_host(g) for g in grids]
    dp = [device_floats(p) for p in pos]
    td = torch.float64 if dtype == np.float64 else torch.float32
    outs = [torch.zeros(m, dtype=td, device="cuda") for _ in range(3)]
    _lib.check(_lib.hip().ocu_sample_points_mac(*(o.data_ptr() for o in outs), *(d.data_ptr() for d in dp), m,
                                                *(d.ptr() for d in dev), C.byref(bc), *h, None), "sample mac")
    for o, w in zip(outs, want):
        assert np.array_equal(o.cpu().numpy(), w)
