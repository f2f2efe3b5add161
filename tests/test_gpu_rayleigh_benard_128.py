"""The Rayleigh-Benard workload at its full size: 128^3 Boussinesq convection
(periodic x/y, no-slip z, T Dirichlet 1 / 0, linear profile plus 1e-3
seeded noise, Ra = 1e5, Pr = 0.71, second-order upwind, forward Euler,
CFL 0.5, projection to 1e-8).

* FP64: the first steps are bit-identical to the oracle's restatement of
  Eqn_IncompressibleNS3D (eqn_incompressns3d.cu:301-399) with the same
  get_max_stable_timestep, every field compared over its whole allocation.
* 100 steps in FP64 and in FP64 with the FP32-refined projection: after
  every step the divergence recomputed here in numpy from the downloaded MAC
  velocity is within the 1e-8 target, every field is finite, and at the end
  the FP32-refined run tracks the FP64 one.
"""
import ctypes as C

import numpy as np
import pytest

from opencurrent_amd import _types as T
from opencurrent_amd._types import BC, BCSet, HostGrid, dtype_code
from fields import mac_grids, uniform
import oracle_lib as O

pytestmark = pytest.mark.gpu

N = 128


def rb_params(mixed):
    from opencurrent_amd._lib import NSParams
    p = NSParams()
    p.nx = p.ny = p.nz = N
    p.hx = p.hy = p.hz = 1.0 / N
    p.max_divergence = 1e-8
    tb = BCSet.uniform(T.BC_PERIODIC)
    tb.zneg = BC(T.BC_DIRICHELET, 1.0, 0.0)
    tb.zpos = BC(T.BC_DIRICHELET, 0.0, 0.0)
    p.temp_bc = tb
    fb = BCSet.uniform(T.BC_PERIODIC)
    fb.zneg = BC(T.BC_FORCED_INFLOW_VARIABLE_SLIP, 0.0, 1.0)
    fb.zpos = BC(T.BC_FORCED_INFLOW_VARIABLE_SLIP, 0.0, 1.0)
    p.flow_bc = fb
    p.vertical_direction = T.DIR_ZPOS
    p.gravity, p.bouyancy, p.viscosity, p.thermal_diffusion = -1.0, 1e5 * .71, .71, 1.0
    p.advection_scheme, p.time_step, p.cfl_factor = T.IT_SECOND_ORDER_UPWIND, T.TS_FORWARD_EULER, .5
    p.mixed_projection = mixed
    return p


def rb_initial():
    u, v, w, t = mac_grids(N, N, N)
    z = (np.arange(N) + .5) / N
    t.interior()[:] = (1 - z)[None, None, :] + 1e-3 * (2 * uniform(2001, t.interior().shape) - 1)
    return u, v, w, t


def _log(s):
    print(s, flush=True)


def test_rb128_fp64_steps_match_oracle(gpu):
    from opencurrent_amd.solvers import IncompressibleNS3D
    p = rb_params(0)
    u, v, w, t = rb_initial()
    ho = O.lib().orc_ns_create(dtype_code(np.float64), C.byref(p), u.ptr(), v.ptr(), w.ptr(), t.ptr())
    assert ho
    ns = IncompressibleNS3D(p, u, v, w, t)
    try:
        for step in range(3):
            dt_o = O.lib().orc_ns_max_stable_timestep(ho)
            dt_g = ns.get_max_stable_timestep()
            assert dt_g == dt_o, step
            assert O.lib().orc_ns_advance_one_step(ho, dt_o) == 1
            assert ns.advance_one_step(dt_g)
            _log(f"step {step}: dt {dt_g:.6e}")
        for which in range(5):
            got = ns.field(which)
            want = HostGrid.congruent(got)
            assert O.lib().orc_ns_get(ho, which, want.ptr()) == 1
            assert np.array_equal(got.buf, want.buf), f"field {which}: {np.abs(got.buf - want.buf).max()}"
        assert np.abs(ns.field(2).interior()).max() > 0  # buoyancy has started the flow
    finally:
        O.lib().orc_ns_destroy(ho)


def _max_div(ns):
    U, V, W = (ns.field(q).interior() for q in range(3))
    h = 1.0 / N
    div = (U[1:] - U[:-1]) / h + (V[:, 1:] - V[:, :-1]) / h + (W[:, :, 1:] - W[:, :, :-1]) / h
    return float(np.abs(div).max()), (U, V, W)


def test_rb128_100_steps_fp64_and_fp32_refined(gpu):
    from opencurrent_amd.solvers import IncompressibleNS3D
    runs = {}
    for name, mixed in (("fp64", 0), ("fp32_refined", 1)):
        ns = IncompressibleNS3D(rb_params(mixed), *rb_initial())
        worst = 0.0
        for step in range(100):
            assert ns.advance_one_step(ns.get_max_stable_timestep()), (name, step)
            d, (U, V, W) = _max_div(ns)
            worst = max(worst, d)
            assert d <= 1e-8 * (1 + 1e-6), (name, step, d)
            assert not np.any(W[:, :, 0]) and not np.any(W[:, :, -1])  # no-slip: w = 0 on the walls
        fields = [ns.field(q).interior().copy() for q in range(4)]
        for q, f in enumerate(fields):
            assert np.all(np.isfinite(f)), (name, q)
        _log(f"{name}: 100 steps, worst max|div| {worst:.3e}, max|w| {np.abs(fields[2]).max():.3e}")
        runs[name] = fields
        del ns
    for q in range(4):
        a, b = runs["fp64"][q], runs["fp32_refined"][q]
        diff = float(np.abs(a - b).max())
        _log(f"field {q}: max|fp64 - fp32_refined| {diff:.3e} (max|fp64| {np.abs(a).max():.3e})")
        assert diff <= 1e-6 * max(1.0, float(np.abs(a).max())), q
