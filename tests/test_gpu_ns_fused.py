"""The fused NS step (ocu_ns_step_fused) against the reference's
operator-by-operator step (eqn_incompressns3d.cu:331-393) on the GPU and
against the oracle: whole buffers (ghosts and padding included) of u, v, w,
T and p after each step must be bit-identical.

Covered beyond the Rayleigh-Benard parity test of test_gpu_solvers.py:
vertical direction along x, y and z with either sign, Neumann / Dirichlet /
no-fill temperature faces, closed and periodic flow boxes, inviscid and
non-diffusive runs, Adams-Bashforth 2 over unequal steps, unequal sizes, and a
301^3 box whose allocation waste is under 3%, where the reference combines
over the whole allocation instead of the interior."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from fields import mac_grids, uniform
from opencurrent_amd import _lib
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
    "closed_vyneg_centred": ((14, 16, 10), T.DIR_YNEG, tbc_mixed, closed, .05, 0.0, T.IT_SECOND_ORDER_CENTERED,
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
            # odd sizes coarsen little, so the projection may stop short of
            # its tolerance: all three runs must agree on that too
            ok_o = O.lib().orc_ns_advance_one_step(ho, dt) == 1
            assert runs[0][1][s] == runs[1][1][s] == ok_o
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


def _reference_sequence(g, d, last, p, tbc, ab, keep):
    """the reference's explicit step (eqn_incompressns3d.cu:331-393) through
    the per-operator C entry points, in place on g = (u, v, w, T)"""
    hip = _lib.hip()
    u, v, w, t = g
    du, dv, dw, dt = d
    nx, ny, nz = p.nx, p.ny, p.nz
    h = (p.hx, p.hy, p.hz)
    chk = _lib.check
    chk(hip.ocu_self_advect_ex(u.ptr(), v.ptr(), w.ptr(), du.ptr(), dv.ptr(), dw.ptr(), nx, ny, nz, *h,
                               p.advection_scheme, p.flow_periodic, None), "self advect")
    if p.viscosity > 0:
        for f, df, n in ((u, du, (nx + 1, ny, nz)), (v, dv, (nx, ny + 1, nz)), (w, dw, (nx, ny, nz + 1))):
            chk(hip.ocu_laplacian_acc(f.ptr(), df.ptr(), *n, *h, p.viscosity, None), "laplacian")
    vert = {T.DIR_XPOS: du, T.DIR_XNEG: du, T.DIR_YPOS: dv, T.DIR_YNEG: dv}.get(p.vertical_direction, dw)
    chk(hip.ocu_thermal_force(vert.ptr(), t.ptr(), nx, ny, nz, p.vertical_direction, p.thermal_coef, None), "thermal")
    chk(hip.ocu_passive_advect_ex(u.ptr(), v.ptr(), w.ptr(), t.ptr(), dt.ptr(), nx, ny, nz, *h, p.advection_scheme,
                                  p.temp_periodic, None), "passive advect")
    if p.thermal_diffusion > 0:
        chk(hip.ocu_laplacian_acc(t.ptr(), dt.ptr(), nx, ny, nz, *h, p.thermal_diffusion, None), "laplacian T")
    for f, df, lf in ((t, dt, last[3]), (u, du, last[0]), (v, dv, last[1]), (w, dw, last[2])):
        chk(hip.ocu_grid_lincomb(f.ptr(), 1.0, f.ptr(), p.a_dt, df.ptr(), None), "lincomb")
        if ab:
            chk(hip.ocu_grid_lincomb(f.ptr(), 1.0, f.ptr(), p.a_ab, lf.ptr(), None), "lincomb ab")
        if f is t:
            chk(hip.ocu_bc_fill_nocorners(t.ptr(), C.byref(tbc), *h, None), "T fill")


@pytest.mark.parametrize("ab", [0, 1])
def test_fused_step_whole_allocation_combination(gpu, ab):
    """400 x 400 x 253 FP32 grids sharing the MAC layout: under 3% of the
    allocation lies outside each grid's interior, so the reference's
    combination runs over the whole allocation (ghosts and padding get
    x*1 + 0*a, turning -0.0 into +0.0).  One step of ocu_ns_step_fused
    equals the per-operator sequence over every element, for forward Euler
    and for an Adams-Bashforth step with given last tendencies."""
    import torch
    from opencurrent_amd._types import NSStepParams
    from opencurrent_amd.device import DeviceGrid
    nx, ny, nz = 400, 400, 253
    dtype = np.float32
    hosts = mac_grids(nx, ny, nz, dtype=dtype)
    d = hosts[3].desc
    assert (d.allocated - d.nx * d.ny * d.nz) / d.allocated <= .03
    for q, g in enumerate(hosts):
        g.buf[:] = (.3 if q < 3 else 1.0) * (2 * uniform(8400 + q, g.buf.shape) - 1)
        g.buf[:40] = -0.0
    tbc = tbc_mixed()
    p = NSStepParams()
    p.nx, p.ny, p.nz = nx, ny, nz
    p.advection_scheme, p.vertical_direction = T.IT_FIRST_ORDER_UPWIND, T.DIR_YNEG
    p.flow_periodic = p.temp_periodic = 0
    p.keep_tendencies, p.use_last = ab, ab
    p.hx, p.hy, p.hz = 1 / nx, 1 / ny, 1 / nz
    p.viscosity, p.thermal_diffusion = .01, .02
    p.thermal_coef = float(np.float32(1.0 * -1.0 * 50.0))
    p.a_dt = float(np.float32(1e-4))
    p.a_ab = float(np.float32(-3e-5)) if ab else 0.0
    ref = [DeviceGrid.from_host(g) for g in hosts]
    old = [DeviceGrid.from_host(g) for g in hosts]
    new = [DeviceGrid.congruent(g) for g in old]
    d_ref = [DeviceGrid.congruent(g) for g in old]
    d_new = [DeviceGrid.congruent(g) for g in old]
    # last tendencies as a step leaves them: values on the interior, zero
    # on ghosts and padding (tendency grids are cleared once and only their
    # interiors are written)
    last_h = mac_grids(nx, ny, nz, dtype=dtype)
    for q, g in enumerate(last_h):
        g.buf[:] = 0
        g.interior()[:] = 2 * uniform(8500 + q, g.interior().shape) - 1
    last = [DeviceGrid.from_host(g) for g in last_h]
    _reference_sequence(ref, d_ref, last, p, tbc, ab, ab)
    _lib.check(_lib.hip().ocu_ns_step_fused(*(g.ptr() for g in old), *(g.ptr() for g in new),
                                            *(g.ptr() for g in d_new), *(g.ptr() for g in last), C.byref(p),
                                            C.byref(tbc), None), "fused step")
    torch.cuda.synchronize()
    for q in range(4):
        a, b = new[q].tensor.view(torch.int32), ref[q].tensor.view(torch.int32)
        assert torch.equal(a, b), (q, int((a != b).sum()))
        if ab:
            assert torch.equal(d_new[q].tensor.view(torch.int32), d_ref[q].tensor.view(torch.int32)), q
