"""The reference's lock-exchange tests (tests/lockextest.cpp:28-206) on the
GPU Eqn_IncompressibleNS3D: a closed free-slip 128^3 box, heavy fluid
(temperature -1) in x < L/2 and light (+1) beyond, gravity along y, stepped
to t = 10 with dt = min(stable step, 0.1).

LockExDoubleTest (FP64, upwind, forward Euler, nu = kappa = 0.1): the
temperature stays uniform in z and antisymmetric about the box centre in
x, y to 1e-4.  LockExTest (FP32, centred, Adams-Bashforth 2, inviscid):
no NaN or infinity appears."""
import numpy as np
import pytest

from opencurrent_amd import _types as T
from opencurrent_amd._types import BCSet
from fields import mac_grids

pytestmark = pytest.mark.gpu


def params(n, dtype):
    from opencurrent_amd._lib import NSParams
    p = NSParams()
    p.nx = p.ny = p.nz = n
    p.hx, p.hy, p.hz = .4, .7, 1.0
    p.flow_bc = BCSet.uniform(T.BC_FORCED_INFLOW_VARIABLE_SLIP)
    p.temp_bc = BCSet.uniform(T.BC_NEUMANN)
    p.vertical_direction = T.DIR_YPOS  # Eqn_IncompressibleNS3DParams default (eqn_incompressns3d.h:62-75)
    p.gravity, p.bouyancy = -1.0, 1.0
    if dtype == np.float64:  # lockextest.cpp:108-124
        p.max_divergence = 1e-6
        p.viscosity = p.thermal_diffusion = .1
        p.advection_scheme, p.time_step, p.cfl_factor = T.IT_FIRST_ORDER_UPWIND, T.TS_FORWARD_EULER, 1.0
    else:  # lockextest.cpp:25-41
        p.max_divergence = 1e-4
        p.viscosity = p.thermal_diffusion = 0.0
        p.advection_scheme, p.time_step, p.cfl_factor = T.IT_SECOND_ORDER_CENTERED, T.TS_ADAMS_BASHFORD2, .7
    return p


def run(n, dtype):
    from opencurrent_amd.solvers import IncompressibleNS3D
    u, v, w, t = mac_grids(n, n, n, dtype)
    t.interior()[:] = np.where(np.arange(n) < n // 2, -1.0, 1.0)[:, None, None]
    ns = IncompressibleNS3D(params(n, dtype), u, v, w, t)
    # lockextest.cpp:80-83: for (float t = 0; t < 10; t += this_dt)
    time, steps = np.float32(0), 0
    while time < 10:
        dt = min(ns.get_max_stable_timestep(), .1)
        assert ns.advance_one_step(dt), f"step {steps}"
        time = np.float32(float(time) + dt)
        steps += 1
    return ns, steps


def test_lockex_double_symmetry(gpu):
    n = 128
    ns, steps = run(n, np.float64)
    fields = [ns.field(q).interior() for q in range(4)]
    for f in fields:
        assert np.all(np.isfinite(f))
    temp = fields[3]
    # uniform in z, antisymmetric about the centre in x and y
    assert np.abs(temp - temp[:, :, :1]).max() <= 1e-4
    assert np.abs(temp + temp[::-1, ::-1, :]).max() <= 1e-4
    assert steps >= 100  # dt <= 0.1 up to t = 10
    assert np.abs(temp).max() > .5  # the fronts have not diffused away


def test_lockex_float_stays_finite(gpu):
    ns, steps = run(128, np.float32)
    for q in range(4):
        assert np.all(np.isfinite(ns.field(q).interior())), q


def test_lockex_full_size_projection(gpu):
    """The lock-exchange workload at full size (512 x 512 x 256, hx = hy = 1,
    hz = 0.5, no-slip walls, FP64 equations with the FP32-refined
    projection to 1e-8): after three steps the divergence recomputed here in
    numpy from the downloaded MAC velocity stays within the projection's
    target, the wall-normal faces are zero and the fields are finite."""
    from opencurrent_amd.solvers import IncompressibleNS3D
    from opencurrent_amd._lib import NSParams
    nx, ny, nz = 512, 512, 256
    hx, hy, hz = 1.0, 1.0, .5
    p = NSParams()
    p.nx, p.ny, p.nz = nx, ny, nz
    p.hx, p.hy, p.hz = hx, hy, hz
    p.max_divergence = 1e-8
    walls = BCSet.uniform(T.BC_FORCED_INFLOW_VARIABLE_SLIP)
    for f in ("xpos", "xneg", "ypos", "yneg", "zpos", "zneg"):
        setattr(walls, f, T.BC(T.BC_FORCED_INFLOW_VARIABLE_SLIP, 0.0, 1.0))
    p.flow_bc = walls
    p.temp_bc = BCSet.uniform(T.BC_NEUMANN)
    p.vertical_direction = T.DIR_YPOS
    p.gravity, p.bouyancy, p.viscosity, p.thermal_diffusion = -1.0, 1.0, .1, .1
    p.advection_scheme, p.time_step, p.cfl_factor = T.IT_FIRST_ORDER_UPWIND, T.TS_FORWARD_EULER, 1.0
    p.mixed_projection = 1
    u, v, w, t = mac_grids(nx, ny, nz)
    t.interior()[:] = np.where(np.arange(nx) < nx // 2, -1.0, 1.0)[:, None, None]
    ns = IncompressibleNS3D(p, u, v, w, t)
    del u, v, w, t
    for step in range(3):
        assert ns.advance_one_step(min(ns.get_max_stable_timestep(), .1)), step
    U, V, W = (ns.field(q).interior() for q in range(3))
    for f in (U, V, W):
        assert np.all(np.isfinite(f))
    assert np.abs(V).max() > 0  # the fronts have started to move
    for f in (U[0], U[-1], V[:, 0], V[:, -1], W[:, :, 0], W[:, :, -1]):
        assert not np.any(f)
    div = (U[1:] - U[:-1]) / hx + (V[:, 1:] - V[:, :-1]) / hy + (W[:, :, 1:] - W[:, :, :-1]) / hz
    assert np.abs(div).max() <= 1e-8 * (1 + 1e-6), np.abs(div).max()
