"""Eqn_ScalarAdvection3D<T> on the GPU (through the ocu_scalar_* C ABI)
against the oracle's restatement of eqn_scalaradvection3d.cpp:88-105
(passive advection, phi += dt*dphi, no-corner ghost fill), step by step and
bit for bit, for the reference's two schemes and second-order upwind, FP64
and FP32, on periodic and walled domains.  The time step is the equation's
own get_max_stable_timestep (eqn_scalaradvection3d.cpp:70-86), checked
against the same rule on the oracle's max|u|, |v|, |w|."""
import ctypes as C

import numpy as np
import pytest

from opencurrent_amd import _types as T
from opencurrent_amd._types import BC, BCSet, HostGrid
from fields import mac_grids, uniform
import oracle_lib as O

pytestmark = pytest.mark.gpu


def _bcs(kind):
    if kind == "periodic":
        return BCSet.uniform(T.BC_PERIODIC)
    b = BCSet.uniform(T.BC_NEUMANN, 0.0)
    b.xneg = BC(T.BC_DIRICHELET, 1.0, 0.0)
    b.xpos = BC(T.BC_DIRICHELET, 0.0, 0.0)
    b.zneg = BC(T.BC_NEUMANN, 0.5, 0.0)
    b.ypos = BC(T.BC_PERIODIC, 0, 0)
    b.yneg = BC(T.BC_PERIODIC, 0, 0)
    return b


def _params(n, h, bc, scheme):
    from opencurrent_amd._lib import ScalarParams
    p = ScalarParams()
    p.nx, p.ny, p.nz = n
    p.hx, p.hy, p.hz = h
    p.bc = bc
    p.advection_scheme = scheme
    return p


def _max_stable(u, v, w, h):
    """eqn_scalaradvection3d.cpp:70-86 on the oracle's max|.| reductions"""
    mu, mv, mw = (O.lib().orc_reduce(g.ptr(), 0) for g in (u, v, w))
    ts = []
    for hh, m in zip(h, (mu, mv, mw)):
        t = hh / m if m != 0 else float("inf")
        ts.append(t if np.isfinite(t) else 1e10)
    return min(ts)


@pytest.mark.parametrize("bckind", ["periodic", "walls"])
@pytest.mark.parametrize("scheme", [T.IT_FIRST_ORDER_UPWIND, T.IT_SECOND_ORDER_CENTERED, T.IT_SECOND_ORDER_UPWIND])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_scalar_advection_matches_oracle(gpu, dtype, scheme, bckind):
    from opencurrent_amd.solvers import ScalarAdvection3D
    n = (20, 24, 18)
    h = (1 / 20, 1 / 24, 1 / 18)
    bc = _bcs(bckind)
    u, v, w, phi = mac_grids(*n, dtype=dtype)
    for q, g in enumerate((u, v, w)):
        g.interior()[:] = (2 * uniform(1001 + q, g.interior().shape) - 1).astype(dtype)
    phi.interior()[:] = uniform(2001, phi.interior().shape).astype(dtype)
    eq = ScalarAdvection3D(_params(n, h, bc, scheme), u, v, w, phi)

    # the oracle's copy: set_parameters fills phi's ghosts (no corners)
    ph = phi.copy()
    dph = HostGrid.congruent(phi)
    assert O.lib().orc_bc_fill(1, ph.ptr(), C.byref(bc), *h) == 1
    assert np.array_equal(eq.field(3).buf, ph.buf)

    dt = eq.get_max_stable_timestep()
    assert dt == _max_stable(u, v, w, h)
    dt *= .5
    for step in range(6):
        assert eq.advance_one_step(dt)
        assert O.lib().orc_scalar_advect_step(u.ptr(), v.ptr(), w.ptr(), ph.ptr(), dph.ptr(), C.byref(bc), *n, *h,
                                              scheme, dt) == 1
        got = eq.field(3)
        assert np.array_equal(got.buf, ph.buf), f"step {step}: {np.abs(got.buf - ph.buf).max()}"
    d = eq.field(4)
    assert np.array_equal(d.interior(), dph.interior())
    assert np.all(np.isfinite(got.interior()))


def test_scalar_advection_advance_substeps(gpu):
    """Equation::advance (equation.cpp:31-56) walks dt down in
    ceil(left / max_stable) equal sub-steps, recomputed after each one: 2.5
    stable steps run as three"""
    import math
    from opencurrent_amd.solvers import ScalarAdvection3D
    n, h = (16, 16, 16), (1 / 16,) * 3
    bc = _bcs("periodic")
    u, v, w, phi = mac_grids(*n)
    u.interior()[:] = 1.0
    phi.interior()[:] = uniform(7, phi.interior().shape)
    a = ScalarAdvection3D(_params(n, h, bc, T.IT_FIRST_ORDER_UPWIND), u, v, w, phi)
    b = ScalarAdvection3D(_params(n, h, bc, T.IT_FIRST_ORDER_UPWIND), u, v, w, phi)
    dt = a.get_max_stable_timestep()
    assert dt == h[0]  # v = w = 0 give hy/0 = inf, replaced by 1e10
    assert a.advance(2.5 * dt)
    left, steps = 2.5 * dt, 0
    while left > 0:
        sub = left / math.ceil(left / dt)
        assert b.advance_one_step(sub)
        left -= sub
        steps += 1
    assert steps == 3
    assert np.array_equal(a.field(3).buf, b.field(3).buf)


def test_init_grid_layouts_are_checked_in_the_library(gpu):
    """ocu_ns_create / ocu_scalar_create copy whole allocations from the host
    descriptors, so a descriptor of another size or type is rejected there
    (NULL handle), not read past its end"""
    from opencurrent_amd import _lib
    from opencurrent_amd._lib import NSParams
    n = (8, 8, 8)
    u, v, w, t = mac_grids(*n)
    small = HostGrid(8, 8, 8)  # u needs (9, 8, 8) with padding (0, 1, 1)
    p = NSParams()
    p.nx, p.ny, p.nz = n
    p.hx = p.hy = p.hz = 1 / 8
    p.max_divergence = 1e-6
    p.flow_bc = BCSet.uniform(T.BC_PERIODIC)
    p.temp_bc = BCSet.uniform(T.BC_PERIODIC)
    p.vertical_direction = T.DIR_ZPOS
    p.advection_scheme, p.time_step, p.cfl_factor = T.IT_FIRST_ORDER_UPWIND, T.TS_FORWARD_EULER, .5
    host = _lib.host()
    assert not host.ocu_ns_create(T.OCU_F64, C.byref(p), small.ptr(), v.ptr(), w.ptr(), t.ptr())
    t32 = HostGrid.congruent(t, dtype=np.float32)
    assert not host.ocu_ns_create(T.OCU_F64, C.byref(p), u.ptr(), v.ptr(), w.ptr(), t32.ptr())
    h = host.ocu_ns_create(T.OCU_F64, C.byref(p), u.ptr(), v.ptr(), w.ptr(), t.ptr())
    assert h
    host.ocu_ns_destroy(h)
    sp = _params(n, (1 / 8,) * 3, BCSet.uniform(T.BC_PERIODIC), T.IT_FIRST_ORDER_UPWIND)
    assert not host.ocu_scalar_create(T.OCU_F64, C.byref(sp), u.ptr(), v.ptr(), small.ptr(), t.ptr())
