"""CPU tests that pin the oracle's second-order (linear) upwind advection,
OCU_IT_SECOND_ORDER_UPWIND, a scheme with no reference counterpart.

Pinned by hand-computed tendencies on 1-D data (both flow directions, with
and without a periodic wrap, the normal and a tangential velocity
component), by flux-form conservation on periodic random fields, by the
exact derivative of quadratics away from the boundary, and by the
convergence order on a periodic sine wave."""
import ctypes as C
import math

import numpy as np
import pytest

from opencurrent_amd import _types as T
from opencurrent_amd._types import BCSet, HostGrid
from fields import mac_grids, random_mac_velocity, uniform
import oracle_lib as O

UP2 = T.IT_SECOND_ORDER_UPWIND

# cell values at -1..4 of a 4-cell axis; hand-computed tendencies of its
# four cells under a unit velocity with h = 1:
#   +1: cell 0 has no second upstream cell below (first order there)
#   -1: cell 3 has none above
SEQ = [1.0, 2.0, 4.0, 8.0, 16.0, 32.0]
PASSIVE_WANT = {
    (1.0, False): [-1.5, -2.5, -5.0, -10.0],
    (-1.0, False): [1.0, 2.0, 4.0, 24.0],
    # periodic: cell 0's second upstream sample, cell -2, wraps to cell 2
    # (8): low face 1 + .5*(1 - 8) = -2.5, tendency -(2.5 + 2.5)
    (1.0, True): [-5.0, -2.5, -5.0, -10.0],
    # cell 3's, cell 5, wraps to cell 1 (4): high face 32 + .5*(32 - 4) = 46
    (-1.0, True): [1.0, 2.0, 4.0, 38.0],
}
# face values at -1..5 of a 4-cell axis for the component normal to it (the
# array has faces 0..4 and a ghost on each side; scaled by the sign), and
# the hand-computed self-advection tendencies at faces 0..3, h = 1
SEQN = [1.0, 2.0, 4.0, 8.0, 16.0, 32.0, 64.0]
NORMAL_WANT = {
    (1.0, False): [-6.0, -22.5, -90.0, -360.0],
    # face 3's far sample above, face 5, sits in the ghost layer
    (-1.0, False): [-4.5, -18.0, -72.0, -288.0],
    # face 0's far sample below, face -2, wraps to face 2 (8):
    # low face 1 + .5*(1 - 8) = -2.5, flux 1.5 * -2.5
    (1.0, True): [-11.25, -22.5, -90.0, -360.0],
    (-1.0, True): [-4.5, -18.0, -72.0, -288.0],
}


def L():
    return O.lib()


def _dims(axis, n=4, other=3):
    d = [other, other, other]
    d[axis] = n
    return d


def _line(axis, length, values, shape):
    """values[0..length-1] laid along `axis` of a padded box, from index 0
    (the low ghost), constant along the other axes"""
    a = np.zeros(shape[axis])
    a[:length] = values
    sh = [1, 1, 1]
    sh[axis] = shape[axis]
    return np.broadcast_to(a.reshape(sh), shape)


def _interior_line(g, axis):
    x = g.interior()
    idx = [0, 0, 0]
    idx[axis] = slice(None)
    line = x[tuple(idx)]
    assert np.all(x == np.expand_dims(line, tuple(i for i in range(3) if i != axis))), "not constant across"
    return line


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("sign", [1.0, -1.0])
@pytest.mark.parametrize("periodic", [False, True])
def test_passive_hand_values(axis, sign, periodic):
    nx, ny, nz = _dims(axis)
    u, v, w, phi = mac_grids(nx, ny, nz)
    vel = (u, v, w)[axis]
    vel.full()[:] = sign
    phi.full()[:] = _line(axis, 6, SEQ, phi.full().shape)
    d = HostGrid.congruent(phi)
    h = [.5, .25, .125]
    h[axis] = 1.0
    pmask = (1 << axis) if periodic else 0
    assert L().orc_passive_advect_ex(u.ptr(), v.ptr(), w.ptr(), phi.ptr(), d.ptr(), nx, ny, nz, *h, UP2, pmask) == 1
    np.testing.assert_array_equal(_interior_line(d, axis), PASSIVE_WANT[(sign, periodic)])
    # the other axes' periodicity changes nothing here
    d2 = HostGrid.congruent(phi)
    L().orc_passive_advect_ex(u.ptr(), v.ptr(), w.ptr(), phi.ptr(), d2.ptr(), nx, ny, nz, *h, UP2, pmask ^ 7 ^ (1 << axis))
    np.testing.assert_array_equal(d2.interior(), d.interior())


def _self(u, v, w, n, h, pmask):
    nx, ny, nz = n
    du, dv, dw = (HostGrid.congruent(g) for g in (u, v, w))
    assert L().orc_self_advect_ex(u.ptr(), v.ptr(), w.ptr(), du.ptr(), dv.ptr(), dw.ptr(), nx, ny, nz, *h, UP2,
                                  pmask) == 1
    return du, dv, dw


@pytest.mark.parametrize("comp", [0, 1, 2])
@pytest.mark.parametrize("sign", [1.0, -1.0])
@pytest.mark.parametrize("periodic", [False, True])
def test_self_normal_hand_values(comp, sign, periodic):
    """component `comp` varying along its own axis, the others zero: the
    flux of the component through the faces normal to it"""
    n = _dims(comp)
    grids = mac_grids(*n)[:3]
    X = grids[comp]
    X.full()[:] = sign * _line(comp, 7, SEQN, X.full().shape)
    h = [.5, .25, .125]
    h[comp] = 1.0
    d = _self(*grids, n, h, (1 << comp) if periodic else 0)
    got = _interior_line(d[comp], comp)[:4]
    np.testing.assert_array_equal(got, NORMAL_WANT[(sign, periodic)])
    for q in range(3):
        if q != comp:
            assert not np.any(d[q].interior()), q


@pytest.mark.parametrize("comp,axis", [(0, 1), (0, 2), (1, 0), (1, 2), (2, 0), (2, 1)])
@pytest.mark.parametrize("sign", [1.0, -1.0])
@pytest.mark.parametrize("periodic", [False, True])
def test_self_tangential_hand_values(comp, axis, sign, periodic):
    """component `comp` varying along `axis`, carried by a constant velocity
    component along `axis`: the same tendencies as a passive scalar"""
    n = _dims(axis)
    grids = mac_grids(*n)[:3]
    X, A = grids[comp], grids[axis]
    A.full()[:] = sign
    X.full()[:] = _line(axis, 6, SEQ, X.full().shape)
    h = [.5, .25, .125]
    h[axis] = 1.0
    d = _self(*grids, n, h, (1 << axis) if periodic else 0)
    line = d[comp].interior()
    idx = [0, 0, 0]
    idx[axis] = slice(0, 4)
    np.testing.assert_array_equal(line[tuple(idx)], PASSIVE_WANT[(sign, periodic)])


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_passive_quadratics_are_differentiated_exactly(axis, sign):
    """a constant velocity carries phi = x^2: away from the boundary the
    linear-upwind flux difference is exactly U * 2x (first order is off by
    U*h)"""
    n = _dims(axis, n=12, other=3)
    u, v, w, phi = mac_grids(*n)
    h = [.5, .25, .125]
    U = .75 * sign
    (u, v, w)[axis].full()[:] = U
    c = np.arange(phi.full().shape[axis]) - 1.0
    x = (c + .5) * h[axis]
    phi.full()[:] = _line(axis, len(x), x * x, phi.full().shape)
    for scheme, exact in ((UP2, True), (T.IT_FIRST_ORDER_UPWIND, False)):
        d = HostGrid.congruent(phi)
        L().orc_passive_advect_ex(u.ptr(), v.ptr(), w.ptr(), phi.ptr(), d.ptr(), *n, *h, scheme, 0)
        got = _interior_line(d, axis)[1:-1]
        want = -U * 2 * x[2:n[axis]]
        err = np.abs(got - want).max()
        if exact:
            assert err < 1e-12, err
        else:
            assert err > .5 * abs(U) * h[axis], err


@pytest.mark.parametrize("pmask", [7, 3])
def test_conservation_on_periodic_axes(pmask):
    """flux form on a periodic box: the tendencies sum to zero (each wrapped
    face flux is the same number seen from both sides), and a constant
    scalar in a divergence-free field has none"""
    nx, ny, nz = 8, 6, 10
    u, v, w, phi = random_mac_velocity(nx, ny, nz)
    bc = BCSet.uniform(T.BC_PERIODIC)
    L().orc_bc_fill_mac(u.ptr(), v.ptr(), w.ptr(), C.byref(bc))
    phi.interior()[:] = uniform(11, phi.interior().shape)
    L().orc_bc_fill(0, phi.ptr(), C.byref(bc), 1, 1, 1)
    h = (.5, .5, .5)
    d = HostGrid.congruent(phi)
    L().orc_passive_advect_ex(u.ptr(), v.ptr(), w.ptr(), phi.ptr(), d.ptr(), nx, ny, nz, *h, UP2, 7)
    assert abs(d.interior().sum()) < 1e-12 * np.abs(d.interior()).sum()
    du, dv, dw = _self(u, v, w, (nx, ny, nz), h, 7)
    for g in (du, dv, dw):
        assert abs(g.interior().sum()) < 1e-12 * np.abs(g.interior()).sum()
    if pmask != 7:
        # without the wrap on z, the z faces of the box take first order:
        # the sums no longer cancel
        d3 = HostGrid.congruent(phi)
        L().orc_passive_advect_ex(u.ptr(), v.ptr(), w.ptr(), phi.ptr(), d3.ptr(), nx, ny, nz, *h, UP2, pmask)
        assert abs(d3.interior().sum()) > 1e-6
        assert np.array_equal(d3.interior()[:, :, 2:-2], d.interior()[:, :, 2:-2])
    cst = HostGrid.congruent(phi)
    cst.full()[:] = 3.0
    p = HostGrid.congruent(u)
    r = C.c_double()
    L().orc_project(u.ptr(), v.ptr(), w.ptr(), C.byref(bc), nx, ny, nz, *h, 1e-13, p.ptr(), C.byref(r))
    L().orc_passive_advect_ex(u.ptr(), v.ptr(), w.ptr(), cst.ptr(), d.ptr(), nx, ny, nz, *h, UP2, 7)
    assert np.abs(d.interior()).max() < 1e-11


def _sine_error(axis, n, scheme):
    dims = [2, 2, 2]
    dims[axis] = n
    u, v, w, phi = mac_grids(*dims)
    h = [1.0, 1.0, 1.0]
    h[axis] = 1.0 / n
    (u, v, w)[axis].full()[:] = 1.0
    k = 2 * math.pi
    c = np.arange(phi.full().shape[axis]) - 1.0
    phi.full()[:] = _line(axis, len(c), np.sin(k * (c + .5) / n), phi.full().shape)
    d = HostGrid.congruent(phi)
    L().orc_passive_advect_ex(u.ptr(), v.ptr(), w.ptr(), phi.ptr(), d.ptr(), *dims, *h, scheme, 1 << axis)
    x = (np.arange(n) + .5) / n
    return np.abs(_interior_line(d, axis) + k * np.cos(k * x)).max()


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_second_order_convergence_on_a_periodic_sine(axis):
    e = [_sine_error(axis, n, UP2) for n in (16, 32, 64)]
    r = [e[0] / e[1], e[1] / e[2]]
    assert all(3.6 < x < 4.4 for x in r), (e, r)
    e1 = [_sine_error(axis, n, T.IT_FIRST_ORDER_UPWIND) for n in (32, 64)]
    assert 1.8 < e1[0] / e1[1] < 2.2, e1
