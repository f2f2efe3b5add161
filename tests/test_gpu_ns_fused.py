"""The fused NS step (ocu_ns_step_fused) against the reference's
operator-by-operator step (eqn_incompressns3d.cu:331-393) on the GPU and
against the oracle: whole buffers (ghosts and padding included) of u, v, w,
T and p after each step must be bit-identical.

Covered beyond the Rayleigh-Benard parity test of test_gpu_solvers.py:
vertical direction along x, y and z with either sign, Neumann / Dirichlet /
no-fill temperature faces, closed and periodic flow boxes, inviscid and
non-diffusive runs, Adams-Bashforth 2 over unequal steps, odd sizes, and a
301^3 box whose allocation waste is under 3%, where the reference combines
over the whole allocation instead of the interior."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from fields import mac_grids, uniform
from opencurrent_amd import _types as T
from opencurrent_amd._types import BC, BCSet, HostGrid, dtype_code

pytestmark = pytest.mark.gpu


def make_params(n, dtype, vd, tbc, fbc, visc, kappa, scheme, ts, h=(1.0, 1.0, 1.0)):
    from opencurrent_amd._lib import NSParams
    p = NSParams()
    p.nx, p.ny, p.nz = n
    p.hx, p.hy, p.hz = (h[0] / n[0], h[1] / n[1], h[2] / n[2])
    p.max_divergence = 1e-8 if dtype == np.float64 else 1e-3
    p.temp_bc, p.flow_bc = tbc, fbc
    p.vertical_direction = vd
    p.gravity, p.bouyancy = -1.0, 50.0
    p.viscosity, p.thermal_diffusion = visc, kappa
    p.advection_scheme, p.time_step, p.cfl_factor = scheme, ts, .5
    return p


def initial(n, dtype, seed):
    u, v, w, t = mac_grids(*n, dtype=dtype)
    for q, g in enumerate((u, v, w)):
        g.interior()[:] = .3 * (2 * uniform(seed + q, g.interior().shape) - 1)
    t.interior()[:] = 2 * uniform(seed + 7, t.interior().shape) - 1
    return u, v, w, t


def closed():
    return BCSet.uniform(T.BC_FORCED_INFLOW_VARIABLE_SLIP)


def periodic_xy():
    b = BCSet.uniform(T.BC_PERIODIC)
    b.zneg = BC(T.BC_FORCED_INFLOW_VARIABLE_SLIP, 0.0, 1.0)
    b.zpos = BC(T.BC_FORCED_INFLOW_VARIABLE_SLIP, 0.0, 1.0)
    return b


def tbc_mixed():
    b = BCSet.uniform(T.BC_NEUMANN)
    b.xneg = BC(T.BC_DIRICHELET, .5, 0.0)
    b.ypos = BC(T.BC_NEUMANN, .25, 0.0)
    b.zneg = BC(T.BC_NONE, 0.0, 0.0)
    return b


def tbc_periodic_xy():
    b = BCSet.uniform(T.BC_PERIODIC)
    b.zneg = BC(T.BC_DIRICHELET, 1.0, 0.0)
    b.zpos = BC(T.BC_DIRICHELET, 0.0, 0.0)
    return b


CASES = {
    # name: (n, vertical, temp bc, flow bc, nu, kappa, scheme, time step)
    "closed_vx_ab2": ((12, 10, 14), T.DIR_XPOS, tbc_mixed, closed, .05, .1, T.IT_FIRST_ORDER_UPWIND,
                      T.TS_ADAMS_BASHFORD2),
    "closed_vyneg_centred": ((11, 16, 9), T.DIR_YNEG, tbc_mixed, closed, .05, 0.0, T.IT_SECOND_ORDER_CENTERED,
                             T.TS_FORWARD_EULER),
    "periodic_vz_inviscid_up2": ((16, 12, 10), T.DIR_ZPOS, tbc_periodic_xy, periodic_xy, 0.0, .1,
                                 T.IT_SECOND_ORDER_UPWIND, T.TS_ADAMS_BASHFORD2),
    "periodic_vzneg_up2": ((10, 14, 12), T.DIR_ZNEG, tbc_periodic_xy, periodic_xy, .05, .1,
                           T.IT_SECOND_ORDER_UPWIND, T.TS_FORWARD_EULER),
}


def run_gpu(p, init, dts):
    from opencurrent_amd.solvers import IncompressibleNS3D
    ns = IncompressibleNS3D(p, *init)
    out, oks = [], []
    for dt in dts:
        oks.append(bool(ns.advance_one_step(dt)))
        out.append([ns.field(q).buf.copy() for q in range(5)])
    return out, oks


# unequal steps exercise the Adams-Bashforth coefficient
DTS = [2e-3, 1.5e-3, 2.5e-3, 1e-3]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("case", sorted(CASES))
def test_fused_step_equals_reference_sequence_and_oracle(gpu, case, dtype):
    n, vd, tbc, fbc, nu, kappa, scheme, ts = CASES[case]
    print(f"fused NS step case={case} dtype={np.dtype(dtype).name} n={n}", flush=True)
    p = make_params(n, dtype, vd, tbc(), fbc(), nu, kappa, scheme, ts)
    init = initial(n, dtype, 8100)
    runs = {}
    for unfused in (0, 1):
        p.unfused_step = unfused
        runs[unfused] = run_gpu(p, init, DTS)
    ho = O.lib().orc_ns_create(dtype_code(dtype), C.byref(p), *(g.ptr() for g in init))
    assert ho
    try:
        for s, dt in enumerate(DTS):
            ok_o = O.lib().orc_ns_advance_one_step(ho, dt) == 1
            assert runs[0][1][s] == runs[1][1][s] == ok_o
            if dtype == np.float64:
                assert ok_o
            for q in range(5):
                a, b = runs[0][0][s][q], runs[1][0][s][q]
                assert np.array_equal(a, b), ("fused vs reference order", case, s, q, np.abs(a - b).max())
                want_h = _host_like(init, q, dtype)
                assert O.lib().orc_ns_get(ho, q, want_h.ptr()) == 1
                assert np.array_equal(a, want_h.buf), ("fused vs oracle", case, s, q,
                                                       np.abs(a - want_h.buf).max())
    finally:
        O.lib().orc_ns_destroy(ho)


def _host_like(init, q, dtype):
    """an empty HostGrid laid out like field q (0 u, 1 v, 2 w, 3 T, 4 p: p
    has the temperature's layout)"""
    return HostGrid.congruent(init[min(q, 3)])


def test_fused_step_whole_allocation_combination(gpu):
    """301^3 FP32 in a closed box: under 3% of the allocation lies outside
    each grid's interior, so the reference's combination runs over the whole
    allocation (ghosts get x*1 + 0*dt); fused and reference-order steps stay
    bit-identical over every element."""
    n = (301, 301, 301)
    dtype = np.float32
    p = make_params(n, dtype, T.DIR_YPOS, tbc_mixed(), closed(), .01, .01, T.IT_FIRST_ORDER_UPWIND,
                    T.TS_ADAMS_BASHFORD2)
    p.max_divergence = 1e-2
    init = initial(n, dtype, 8300)
    d = init[3].desc
    interior = d.nx * d.ny * d.nz
    assert (d.allocated - interior) / d.allocated <= .03
    # negative zeros in the ghosts: the whole-allocation combination turns them into +0
    init[3].buf[: 16] = -0.0
    dts = [1e-3, 2e-3]
    runs = {}
    for unfused in (0, 1):
        p.unfused_step = unfused
        runs[unfused] = run_gpu(p, init, dts)[0]
    for s in range(len(dts)):
        for q in range(5):
            a, b = runs[0][s][q], runs[1][s][q]
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (s, q)
