"""CPU checks that pin the oracle's adjoints (which have no reference
counterpart) by the dot-product identity <J x, y> = <x, J^T y> in FP64.

The passive-advection operator dphi = A(u, phi) (with the no-corner ghost
fill of phi in front) is linear in phi for a fixed velocity, and linear in u
for the centred scheme; for the upwind scheme it is linear in u wherever no
face velocity changes sign, so the check perturbs a velocity bounded away
from zero."""
import ctypes as C

import numpy as np
import pytest

from opencurrent_amd import _types as T
from opencurrent_amd._types import BC, BCSet, HostGrid, MGConfig
from fields import mac_grids, uniform
import oracle_lib as O

N = (7, 6, 9)
H = (.5, .25, 1 / 3)
SCHEMES = {"upwind": T.IT_FIRST_ORDER_UPWIND, "centred": T.IT_SECOND_ORDER_CENTERED}


def bcs(kind):
    if kind == "periodic":
        return BCSet.uniform(T.BC_PERIODIC)
    if kind == "neumann":
        return BCSet.uniform(T.BC_NEUMANN)
    if kind == "dirichlet":
        return BCSet.uniform(T.BC_DIRICHELET)
    s = BCSet.uniform(T.BC_PERIODIC)
    s.zpos = BC(T.BC_DIRICHELET, 0.0, 0.0)
    s.zneg = BC(T.BC_NEUMANN, 0.0, 0.0)
    return s


def velocity(seed, away_from_zero=False, n=N):
    u, v, w, t = mac_grids(*n)
    for q, g in enumerate((u, v, w)):
        x = 2 * uniform(seed + q, g.buf.shape) - 1
        if away_from_zero:
            x = np.sign(x) * (.1 + .9 * np.abs(x))
        g.buf[:] = x
    return u, v, w, t


def advect(u, v, w, phi_interior, bc, scheme):
    """dphi interior of A(u, fill(phi))"""
    _, _, _, t = mac_grids(*N)
    t.interior()[:] = phi_interior
    O.lib().orc_bc_fill(1, t.ptr(), C.byref(bc), *H)
    d = HostGrid.congruent(t)
    assert O.lib().orc_passive_advect(u.ptr(), v.ptr(), w.ptr(), t.ptr(), d.ptr(), *N, *H, scheme) == 1
    return d.interior().copy(), t


def adjoint(u, v, w, phi_grid, mu, bc, scheme):
    """(fill^T A_phi^T mu, A_u^T mu) with the oracle's adjoints"""
    lam = HostGrid.congruent(phi_grid)
    lam.interior()[:] = mu
    out = HostGrid.congruent(phi_grid)
    au, av, aw, _ = mac_grids(*N)
    assert O.lib().orc_passive_advect_adjoint(u.ptr(), v.ptr(), w.ptr(), phi_grid.ptr(), lam.ptr(), out.ptr(),
                                              au.ptr(), av.ptr(), aw.ptr(), 1.0, *N, *H, scheme) == 1
    out.buf[:] -= lam.buf  # remove the identity part
    assert O.lib().orc_bc_adjoint_nocorners(out.ptr(), C.byref(bc)) == 1
    return out, (au, av, aw)


def faces(g, axis, n=N):
    """the velocity entries an advection reads: faces [0, n] of interior rows"""
    nx, ny, nz = n
    f = g.full()[1:, 1:, 1:]  # index 0 = cell 0
    return f[:nx + 1, :ny, :nz] if axis == 0 else f[:nx, :ny + 1, :nz] if axis == 1 else f[:nx, :ny, :nz + 1]


@pytest.mark.parametrize("scheme", list(SCHEMES))
@pytest.mark.parametrize("kind", ["periodic", "neumann", "dirichlet", "mixed"])
def test_advection_adjoint_in_phi(scheme, kind):
    bc, sch = bcs(kind), SCHEMES[scheme]
    u, v, w, _ = velocity(11)
    x = 2 * uniform(21, N) - 1
    y = 2 * uniform(22, N) - 1
    jx, grid = advect(u, v, w, x, bc, sch)
    jty, _ = adjoint(u, v, w, grid, y, bc, sch)
    lhs, rhs = float((jx * y).sum()), float((x * jty.interior()).sum())
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)
    # every ghost adjoint was folded into the interior
    full = jty.full()[:N[0] + 2, :N[1] + 2, :N[2] + 2]  # the ghost box (index 0 = ghost -1)
    assert np.all(full[0, 1:-1, 1:-1] == 0) and np.all(full[-1, 1:-1, 1:-1] == 0)


@pytest.mark.parametrize("scheme", list(SCHEMES))
def test_advection_adjoint_in_velocity(scheme):
    bc, sch = bcs("periodic"), SCHEMES[scheme]
    u, v, w, _ = velocity(31, away_from_zero=True)
    phi = 2 * uniform(41, N) - 1
    y = 2 * uniform(42, N) - 1
    du, dv, dw, _ = velocity(51)
    base, grid = advect(u, v, w, phi, bc, sch)
    t = 1e-3  # |t du| < .1 <= |u|: no face changes its upwind side
    up = [HostGrid.congruent(g) for g in (u, v, w)]
    for g, b, d in zip(up, (u, v, w), (du, dv, dw)):
        g.buf[:] = b.buf + t * d.buf
    pert, _ = advect(*up, phi, bc, sch)
    lhs = float(((pert - base) / t * y).sum())
    _, (au, av, aw) = adjoint(u, v, w, grid, y, bc, sch)
    rhs = sum(float((faces(a, q) * faces(d, q)).sum()) for q, (a, d) in enumerate(zip((au, av, aw), (du, dv, dw))))
    assert abs(lhs - rhs) <= 1e-10 * max(abs(lhs), 1.0)


def test_mac_normal_adjoint_folds_periodic_duplicates():
    bc = bcs("periodic")
    u, v, w, _ = mac_grids(*N)
    for q, g in enumerate((u, v, w)):
        g.buf[:] = 2 * uniform(61 + q, g.buf.shape) - 1
    want = [faces(g, q).copy() for q, g in enumerate((u, v, w))]
    assert O.lib().orc_mac_adjoint_normal(u.ptr(), v.ptr(), w.ptr(), C.byref(bc)) == 1
    for q, (g, wnt) in enumerate(zip((u, v, w), want)):
        got = faces(g, q)
        sl_last = [slice(None)] * 3
        sl_last[q] = -1
        sl_first = [slice(None)] * 3
        sl_first[q] = 0
        assert np.all(got[tuple(sl_last)] == 0)
        assert np.allclose(got[tuple(sl_first)], wnt[tuple(sl_first)] + wnt[tuple(sl_last)], rtol=0, atol=0)


def cells_and_face_ghosts(g):
    """interior cells plus the three low face-ghost layers, as a mask over buf"""
    m = np.zeros(g.full().shape, bool)
    nx, ny, nz = N
    m[1:nx + 1, 1:ny + 1, 1:nz + 1] = True
    m[0, 1:ny + 1, 1:nz + 1] = m[1:nx + 1, 0, 1:nz + 1] = m[1:nx + 1, 1:ny + 1, 0] = True
    return m


def test_divergence_adjoint_dot_product():
    """<D u, mu> = <u, D^T mu> over the faces [0,n] of interior rows; mu's
    ghost values are not read"""
    u, v, w, d = velocity(71)
    assert O.lib().orc_divergence(u.ptr(), v.ptr(), w.ptr(), d.ptr(), *N, *H) == 1
    mu = 2 * uniform(81, N) - 1
    lam = HostGrid.congruent(d)
    lam.buf[:] = 5.0  # ghosts and padding must not matter
    lam.interior()[:] = mu
    au, av, aw, _ = mac_grids(*N)
    assert O.lib().orc_divergence_adjoint(lam.ptr(), au.ptr(), av.ptr(), aw.ptr(), *N, *H) == 1
    lhs = float((d.interior() * mu).sum())
    rhs = sum(float((faces(g, q) * faces(a, q)).sum()) for q, (g, a) in enumerate(zip((u, v, w), (au, av, aw))))
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)
    for q, a in enumerate((au, av, aw)):  # nothing written off those faces
        faces(a, q)[:] = 0
        assert np.all(a.buf == 0)


@pytest.mark.parametrize("coef", [-1.0, .75])
def test_gradient_adjoint_dot_product(coef):
    """<G p, lam> = <p, G^T lam>: G p is the increment gradient_sub adds to
    the faces [0,n-1] of interior rows; p is read on the interior and the
    low face ghosts, where G^T lam is written"""
    u, v, w, p = mac_grids(*N)
    p.buf[:] = 2 * uniform(91, p.buf.shape) - 1
    assert O.lib().orc_gradient_sub(u.ptr(), v.ptr(), w.ptr(), p.ptr(), coef, *H) == 1
    lu, lv, lw, ap = velocity(95)
    ap.buf[:] = 0
    assert O.lib().orc_gradient_adjoint(lu.ptr(), lv.ptr(), lw.ptr(), ap.ptr(), coef, *H) == 1
    lhs = sum(float((g.buf * l.buf).sum()) for g, l in zip((u, v, w), (lu, lv, lw)))
    rhs = float((p.buf * ap.buf).sum())
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)
    m = cells_and_face_ghosts(ap)
    assert np.all(ap.full()[~m] == 0)


def flow_bcs(kind):
    if kind == "periodic":
        return BCSet.uniform(T.BC_PERIODIC)
    if kind == "closed":  # free-slip box
        return BCSet.uniform(T.BC_FORCED_INFLOW_VARIABLE_SLIP)
    s = BCSet.uniform(T.BC_PERIODIC)  # Rayleigh-Benard: no-slip z walls, periodic x, y
    s.zpos = BC(T.BC_FORCED_INFLOW_VARIABLE_SLIP, 0.0, 1.0)
    s.zneg = BC(T.BC_FORCED_INFLOW_VARIABLE_SLIP, 0.0, 1.0)
    return s


def poisson_bcs(flow):
    """convert_bc_to_poisson_eqn (sol_project3d.cpp:26-43)"""
    s = BCSet.uniform(T.BC_PERIODIC)
    for f in ("xpos", "xneg", "ypos", "yneg", "zpos", "zneg"):
        if getattr(flow, f).type == T.BC_FORCED_INFLOW_VARIABLE_SLIP:
            setattr(s, f, BC(T.BC_NEUMANN, 0.0, 0.0))
    return s


def project_adjoint(lu, lv, lw, flow, tol, n, h):
    """P^T in place: M^T, G^T (c = -1), pressure-fill^T, L^-1 (L = L^T),
    D^T, M^T - the transpose of MAC fill, divergence, solve, gradient, MAC
    fill taken in reverse order"""
    O.lib().orc_mac_adjoint_normal(lu.ptr(), lv.ptr(), lw.ptr(), C.byref(flow))
    _, _, _, lp = mac_grids(*n)
    lp.buf[:] = 0
    assert O.lib().orc_gradient_adjoint(lu.ptr(), lv.ptr(), lw.ptr(), lp.ptr(), -1.0, *h) == 1
    pb = poisson_bcs(flow)
    assert O.lib().orc_bc_adjoint_nocorners(lp.ptr(), C.byref(pb)) == 1
    x = HostGrid.congruent(lp)
    cfg = MGConfig.make(pb, convergence=T.CONVERGENCE_LINF)
    res, vc = C.c_double(), C.c_int()
    hist, hl = (C.c_double * 1)(), C.c_int()
    assert O.lib().orc_mg_solve(x.ptr(), lp.ptr(), *h, C.byref(cfg), tol, 15, 0, C.byref(res), C.byref(vc), hist, 0,
                                C.byref(hl)) == 1
    assert O.lib().orc_divergence_adjoint(x.ptr(), lu.ptr(), lv.ptr(), lw.ptr(), *n, *h) == 1
    O.lib().orc_mac_adjoint_normal(lu.ptr(), lv.ptr(), lw.ptr(), C.byref(flow))


@pytest.mark.parametrize("kind", ["periodic", "closed", "rb"])
def test_projection_adjoint_dot_product(kind):
    """<P u, lam> = <u, P^T lam> over the normal faces [0,n] of interior rows,
    to the pressure solves' tolerance, for periodic and walled domains"""
    flow = flow_bcs(kind)
    n, h = (16, 16, 16), (1 / 16, 1 / 16, 1 / 16)  # the multigrid needs even sizes
    u, v, w, _ = velocity(101, n=n)
    x0 = [faces(g, q, n).copy() for q, g in enumerate((u, v, w))]
    p = HostGrid.congruent(u)
    r = C.c_double()
    assert O.lib().orc_project(u.ptr(), v.ptr(), w.ptr(), C.byref(flow), *n, *h, 1e-11, p.ptr(), C.byref(r)) == 1
    lu, lv, lw, _ = mac_grids(*n)
    for q, g in enumerate((lu, lv, lw)):
        faces(g, q, n)[:] = 2 * uniform(111 + q, faces(g, q, n).shape) - 1
    y = [faces(g, q, n).copy() for q, g in enumerate((lu, lv, lw))]
    lhs = sum(float((faces(g, q, n) * yq).sum()) for q, (g, yq) in enumerate(zip((u, v, w), y)))
    project_adjoint(lu, lv, lw, flow, 1e-11, n, h)
    rhs = sum(float((faces(g, q, n) * xq).sum()) for q, (g, xq) in enumerate(zip((lu, lv, lw), x0)))
    assert abs(lhs - rhs) <= 1e-9 * max(abs(lhs), 1.0), (lhs, rhs)
