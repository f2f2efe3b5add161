"""CPU pins of the oracle's two-layer ghost fills and point sampling
(oracle/oracle_bc2.h; reference ocustorage/grid3dboundary.cu:122-232,
350-549 and ocustorage/grid3dsample.cu).

Properties fixed by what the fills and the interpolation are for, not by
the oracle's code: a linear field with matching Neumann gradients and a
periodic field are continued exactly two cells beyond each face; Dirichlet
ghosts reflect the boundary cell about the wall value; edges and corners
are never written; MAC no-slip / free-slip tangential ghosts mirror with
the right sign and periodic MAC fields wrap; trilinear sampling returns
cell values at cell centres, reproduces linear fields, repeats with the
period and extrapolates beyond a Neumann face with the face's slope."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from fields import uniform
from opencurrent_amd import _types as T
from opencurrent_amd._types import BC, BCSet, HostGrid


def grid2(n, dtype=np.float64, pad=(0, 0, 0)):
    return HostGrid(*n, 2, 2, 2, *pad, dtype=dtype)


def full(g):
    d = g.desc
    pre = 16 - d.gz
    return g.buf[pre:].reshape(d.pnx, d.pny, d.pnz)


def at(g, i, j, k):
    d = g.desc
    return full(g)[i + d.gx, j + d.gy, k + d.gz]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_level2_neumann_continues_a_linear_field(dtype):
    n = (6, 5, 7)
    h = (.5, .25, .125)
    b = (.375, -.75, 1.125)   # boundary values are floats: exact ones
    g = grid2(n, dtype)
    g.buf[:] = np.nan
    i, j, k = np.meshgrid(*[np.arange(-2, m + 2) for m in n], indexing="ij")
    lin = 1.0 + b[0] * (i + .5) * h[0] + b[1] * (j + .5) * h[1] + b[2] * (k + .5) * h[2]
    g.interior()[:] = lin[2:-2, 2:-2, 2:-2]
    bc = BCSet(*(BC(T.BC_NEUMANN, v, 0.0) for v in (b[0], b[0], b[1], b[1], b[2], b[2])))
    assert O.lib().orc_bc_fill_level2_nocorners(g.ptr(), C.byref(bc), *h) == 1
    F = full(g)[g.desc.gx - 2:g.desc.gx + n[0] + 2, g.desc.gy - 2:g.desc.gy + n[1] + 2,
                g.desc.gz - 2:g.desc.gz + n[2] + 2]
    tol = 1e-12 if dtype == np.float64 else 2e-6
    for ax in range(3):
        sl = [slice(2, -2)] * 3
        for layer in (0, 1, -1, -2):
            sl2 = list(sl)
            sl2[ax] = layer
            assert np.allclose(F[tuple(sl2)], lin[tuple(sl2)], rtol=0, atol=tol * 4), (ax, layer)
    # edges and corners of the two-layer frame are left alone
    assert np.isnan(F[0, 0, 2]) and np.isnan(F[1, 1, 1]) and np.isnan(F[-1, 2, -2])


def test_level2_periodic_wraps_and_dirichlet_reflects():
    n = (5, 6, 4)
    g = grid2(n)
    g.buf[:] = np.nan
    vals = 2 * uniform(9900, n) - 1
    g.interior()[:] = vals
    bc = BCSet.uniform(T.BC_PERIODIC)
    bc.zneg = BC(T.BC_DIRICHELET, .25, 0.0)
    bc.zpos = BC(T.BC_NONE, 0.0, 0.0)
    assert O.lib().orc_bc_fill_level2_nocorners(g.ptr(), C.byref(bc), 1.0, 1.0, 1.0) == 1
    for i in (-2, -1, n[0], n[0] + 1):
        assert np.array_equal([at(g, i, 2, 1)], [vals[i % n[0], 2, 1]])
    for j in (-2, -1, n[1], n[1] + 1):
        assert at(g, 3, j, 2) == vals[3, j % n[1], 2]
    # Dirichlet: the wall value is the mean of the boundary cell and its
    # first ghost; the second ghost repeats the first
    assert at(g, 1, 1, -1) == pytest.approx(2 * .25 - vals[1, 1, 0], abs=0)
    assert at(g, 1, 1, -2) == at(g, 1, 1, -1)
    # a no-fill face stays untouched
    assert np.isnan(at(g, 1, 1, n[2])) and np.isnan(at(g, 1, 1, n[2] + 1))


def mac2(n, dtype=np.float64):
    nx, ny, nz = n
    u = HostGrid(nx + 1, ny, nz, 2, 2, 2, 0, 1, 1, dtype=dtype)
    v = HostGrid(nx, ny + 1, nz, 2, 2, 2, 1, 0, 1, dtype=dtype)
    w = HostGrid(nx, ny, nz + 1, 2, 2, 2, 1, 1, 0, dtype=dtype)
    return u, v, w


def test_mac_level2_slip_and_inflow():
    n = (5, 4, 6)
    u, v, w = mac2(n)
    for q, g in enumerate((u, v, w)):
        g.interior()[:] = 2 * uniform(9910 + q, g.interior().shape) - 1
    bc = BCSet.uniform(T.BC_FORCED_INFLOW_VARIABLE_SLIP)
    bc.xneg = BC(T.BC_FORCED_INFLOW_VARIABLE_SLIP, .5, 1.0)    # inflow .5, no slip
    bc.xpos = BC(T.BC_FORCED_INFLOW_VARIABLE_SLIP, 0.0, 0.0)   # wall, free slip
    v0 = v.interior().copy()
    assert O.lib().orc_bc_fill_mac_level2(u.ptr(), v.ptr(), w.ptr(), C.byref(bc)) == 1
    nx = n[0]
    for j in range(n[1]):
        for k in range(n[2]):
            # the normal component takes the face value on the face and both ghosts
            assert at(u, 0, j, k) == at(u, -1, j, k) == at(u, -2, j, k) == .5
            assert at(u, nx, j, k) == at(u, nx + 1, j, k) == at(u, nx + 2, j, k) == 0.0
            if j == 0:
                continue  # v's face on the y wall: the later y pass sets it to the wall value
            # no slip: tangential ghosts are minus the boundary cell (wall value 0)
            assert at(v, -1, j, k) == at(v, -2, j, k) == -v0[0, j, k]
            # free slip: they copy it
            assert at(v, nx, j, k) == at(v, nx + 1, j, k) == v0[nx - 1, j, k]


def test_mac_level2_periodic_wraps():
    n = (6, 4, 5)
    u, v, w = mac2(n)
    for q, g in enumerate((u, v, w)):
        g.interior()[:] = 2 * uniform(9920 + q, g.interior().shape) - 1
    bc = BCSet.uniform(T.BC_PERIODIC)
    assert O.lib().orc_bc_fill_mac_level2(u.ptr(), v.ptr(), w.ptr(), C.byref(bc)) == 1
    nx, ny, nz = n
    for j in range(ny):
        for k in range(nz):
            for i in (-2, -1):
                assert at(u, i, j, k) == at(u, i + nx, j, k)
                assert at(w, i, j, k) == at(w, i + nx, j, k)
            for i in (0, 1, 2):
                assert at(u, nx + i, j, k) == at(u, i, j, k)
    for i in range(-1, nx + 1):
        for k in range(nz):
            assert at(v, i, ny + 1, k) == at(v, i, 1, k)
            assert at(u, i, -2, k) == at(u, i, ny - 2, k)
    for i in range(-1, nx + 1):
        for j in range(-1, ny + 1):
            assert at(w, i, j, nz + 2) == at(w, i, j, 2)
            assert at(v, i, j, -1) == at(v, i, j, nz - 1)


def sample(g, bc, pos, period, h, shift):
    px, py, pz = (np.ascontiguousarray(p, dtype=np.float32) for p in pos)
    out = np.zeros(len(px), dtype=g.buf.dtype)
    ok = O.lib().orc_sample_points(out.ctypes.data, px.ctypes.data, py.ctypes.data, pz.ctypes.data, len(px), g.ptr(),
                                   C.byref(bc), *period, *h, *shift)
    assert ok == 1
    return out


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_sampling_reproduces_cell_values_and_linear_fields(dtype):
    n = (8, 6, 5)
    h = (.25, .5, .125)
    g = HostGrid(*n, dtype=dtype)
    bc = BCSet.uniform(T.BC_NEUMANN)
    vals = 2 * uniform(9930, n) - 1
    g.interior()[:] = vals
    O.lib().orc_bc_fill(0, g.ptr(), C.byref(bc), *h)
    ii, jj, kk = (a.ravel() for a in np.meshgrid(*[np.arange(m) for m in n], indexing="ij"))
    pos = [(ii + .5) * h[0], (jj + .5) * h[1], (kk + .5) * h[2]]
    got = sample(g, bc, pos, n, h, (.5, .5, .5))
    assert np.array_equal(got, vals[ii, jj, kk].astype(np.float32).astype(dtype))
    # a linear field (ghosts continued by the matching Neumann gradient)
    b = (.4, -1.2, .7)
    X = [(np.arange(-1, m + 1) + .5) * hh for m, hh in zip(n, h)]
    lin = lambda x, y, z: .3 + b[0] * x + b[1] * y + b[2] * z
    gl = HostGrid(*n, dtype=dtype)
    full(gl)[:n[0] + 2, :n[1] + 2, :n[2] + 2] = lin(*np.meshgrid(*X, indexing="ij"))
    r = np.random.default_rng(9931)
    q = [r.uniform(.5 * hh, (m - .5) * hh, 400) for m, hh in zip(n, h)]
    got = sample(gl, bc, q, n, h, (.5, .5, .5))
    assert np.allclose(got, lin(*[np.float32(x) for x in q]), rtol=0, atol=2e-6)


def test_sampling_periodic_repeats_and_neumann_extrapolates_with_the_face_slope():
    n = (8, 8, 8)
    h = (.125,) * 3
    g = HostGrid(*n)
    bc = BCSet.uniform(T.BC_PERIODIC)
    bc.zneg = BC(T.BC_NEUMANN, .75, 0.0)
    bc.zpos = BC(T.BC_NEUMANN, -2.0, 0.0)
    g.interior()[:] = 2 * uniform(9940, n) - 1
    O.lib().orc_bc_fill(0, g.ptr(), C.byref(bc), *h)
    r = np.random.default_rng(9941)
    x, y, z = (r.uniform(0, 1, 200) for _ in range(3))
    a = sample(g, bc, (x, y, z), n, h, (.5, .5, .5))
    b = sample(g, bc, (x + 1, y - 2, z), n, h, (.5, .5, .5))
    assert np.allclose(a, b, atol=1e-5)
    # beyond the low z face the value changes by value * distance (outward)
    zs = np.array([-.3, -.2, -.1], dtype=np.float32)
    xs, ys = np.full(3, .4, np.float32), np.full(3, .6, np.float32)
    s = sample(g, bc, (xs, ys, zs), n, h, (.5, .5, .5))
    d = np.diff(s) / np.diff(zs.astype(np.float64))
    assert np.allclose(d, -.75, rtol=1e-4)
    zs = np.array([1.1, 1.2, 1.3], dtype=np.float32)
    s = sample(g, bc, (xs, ys, zs), n, h, (.5, .5, .5))
    assert np.allclose(np.diff(s) / np.diff(zs.astype(np.float64)), -2.0, rtol=1e-4)


def test_mac_sampling_of_a_uniform_flow():
    n = (6, 5, 4)
    from fields import mac_grids
    u, v, w, _ = mac_grids(*n)
    U = (.3, -.2, .7)
    for g, val in zip((u, v, w), U):
        g.buf[:] = val
    bc = BCSet.uniform(T.BC_PERIODIC)
    h = (.5, .5, .5)
    r = np.random.default_rng(9950)
    pos = [np.ascontiguousarray(r.uniform(0, m * hh, 100), dtype=np.float32) for m, hh in zip(n, h)]
    outs = [np.zeros(100) for _ in range(3)]
    ok = O.lib().orc_sample_points_mac(*(o.ctypes.data for o in outs), *(p.ctypes.data for p in pos), 100, u.ptr(),
                                       v.ptr(), w.ptr(), C.byref(bc), *h)
    assert ok == 1
    for o, val in zip(outs, U):
        assert np.allclose(o, val, atol=1e-7)
    # a face pair that mixes periodic and forced inflow leaves the negative
    # face undefined, as in the reference: the call fails
    bc.xneg = BC(T.BC_FORCED_INFLOW_VARIABLE_SLIP, 0.0, 1.0)
    assert O.lib().orc_sample_points_mac(*(o.ctypes.data for o in outs), *(p.ctypes.data for p in pos), 100,
                                         u.ptr(), v.ptr(), w.ptr(), C.byref(bc), *h) == 0
