"""CPU tests that pin the oracle itself.

The reference stores no golden vectors (its tests build inputs
procedurally and assert tolerances), so the oracle is pinned by what the
mathematics fixes: the closed-form layout numbers, the defining properties
of each ghost-cell rule, the discrete eigenpairs of the 7-point operator,
an independent sparse direct solve, conservation and exact-shift properties
of the advection stencils, and the divergence/gradient identity."""
import ctypes as C
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from opencurrent_amd import _types as T
from opencurrent_amd._types import BC, BCSet, HostGrid, MGConfig, layout
from fields import divergence_rhs, mac_grids, random_mac_velocity, trig_rhs, uniform
import oracle_lib as O


def L():
    return O.lib()


# ---------------------------------------------------------------- layout

@pytest.mark.parametrize("dims,pads,want", [
    ((64, 64, 64), (0, 0, 0), (66, 66, 80, 348495, 5376)),
    ((256, 256, 256), (0, 0, 0), (258, 258, 272, 18105423, 70464)),
    ((129, 128, 128), (0, 1, 1), (131, 131, 144, None, None)),
    ((513, 512, 256), (0, 1, 1), (515, 515, 272, None, None)),
])
def test_layout_constants(dims, pads, want):
    d = C.POINTER(T.Grid3DDesc)
    g = T.Grid3DDesc()
    L().orc_layout(*dims, 1, 1, 1, *pads, T.OCU_F64, C.byref(g))
    assert (g.pnx, g.pny, g.pnz) == want[:3]
    if want[3] is not None:
        assert g.allocated == want[3] and g.shift == want[4]
    py = layout(*dims, 1, 1, 1, *pads)
    assert (py.pnx, py.pny, py.pnz, py.allocated, py.shift) == (g.pnx, g.pny, g.pnz, g.allocated, g.shift)


def test_num_levels_and_omega():
    assert L().orc_mg_num_levels(64, 64, 64) == 5
    assert L().orc_mg_num_levels(256, 256, 256) == 7
    assert L().orc_mg_num_levels(512, 512, 256) == 7
    # 96x128x80 -> 48x64x40 -> 24x32x20 -> 12x16x10; the next halving would
    # reach nz = 5, where the reference stops (sol_mgpressure3d.cpp:84-89)
    assert L().orc_mg_num_levels(96, 128, 80) == 4
    # 8x8x4 is already the coarsest grid a halving may produce: one level
    assert L().orc_mg_num_levels(8, 8, 4) == 1
    # Yavneh's optimal red-black SOR factor for f = (1,1,1): 2/(1+sqrt(1-(2/3)^2))
    assert abs(L().orc_mg_optimal_omega(1, 1, 1) - 2 / (1 + math.sqrt(1 - 4 / 9))) < 1e-7
    assert abs(L().orc_mg_optimal_omega(1, 1, 1) - 1.1458980) < 1e-7
    assert abs(L().orc_mg_optimal_omega(.25, .25, 1) - 1.2880201) < 1e-7


# ---------------------------------------------------------- ghost rules

def bc6(xp, xn, yp, yn, zp, zn):
    s = BCSet()
    for f, spec in zip(BCSet.FACES, (xp, xn, yp, yn, zp, zn)):
        setattr(s, f, BC(spec[0], spec[1], 0.0))
    return s


def test_ghost_rules_define_the_boundary_values():
    nx, ny, nz = 6, 5, 7
    h = (.5, .25, .125)
    g = HostGrid(nx, ny, nz)
    g.buf[:] = uniform(1, g.buf.shape)
    bc = bc6((T.BC_DIRICHELET, .75), (T.BC_NEUMANN, .5), (T.BC_PERIODIC, 0), (T.BC_PERIODIC, 0),
             (T.BC_NEUMANN, -1.25), (T.BC_DIRICHELET, -.5))
    L().orc_bc_fill(0, g.ptr(), C.byref(bc), *h)
    F = g.full()
    I = lambda i, j, k: F[i + 1, j + 1, k + 1]
    for j in range(ny):
        for k in range(nz):
            # Dirichlet: the face value (mean of ghost and interior) is the BC value
            assert (I(nx, j, k) + I(nx - 1, j, k)) / 2 == .75
            # Neumann: the outward normal derivative is the BC value
            assert abs((I(0, j, k) - I(-1, j, k)) / h[0] - .5) < 1e-14
    for i in range(-1, nx + 1):
        for k in range(nz):
            assert I(i, -1, k) == I(i, ny - 1, k) and I(i, ny, k) == I(i, 0, k)  # periodic wrap
    for i in range(-1, nx + 1):
        for j in range(-1, ny + 1):
            assert abs((I(i, j, nz) - I(i, j, nz - 1)) / h[2] + 1.25) < 1e-13
            assert abs((I(i, j, -1) + I(i, j, 0)) / 2 + .5) < 1e-15


def test_dirichlet_edges_and_corners_are_reflections():
    """x, y, z passes in that order: an edge ghost is reflected twice, a
    corner three times, so with value v: edge = phi, corner = 2v - phi."""
    g = HostGrid(4, 4, 4)
    g.buf[:] = uniform(2, g.buf.shape)
    bc = BCSet.uniform(T.BC_DIRICHELET, .5)
    L().orc_bc_fill(0, g.ptr(), C.byref(bc), 1, 1, 1)
    F = g.full()
    assert F[0, 0, 2] == F[1, 1, 2]                       # edge (-1,-1,k)
    assert abs(F[0, 0, 0] - (1.0 - F[1, 1, 1])) < 1e-15    # corner (-1,-1,-1)
    g2 = HostGrid(4, 4, 4)
    g2.buf[:] = uniform(3, g2.buf.shape)
    L().orc_bc_fill(0, g2.ptr(), C.byref(BCSet.uniform(T.BC_NEUMANN)), 1, 1, 1)
    F2 = g2.full()
    assert F2[0, 0, 0] == F2[1, 1, 1] and F2[5, 0, 5] == F2[4, 1, 4]


def test_nocorner_fills_leave_edges_untouched():
    for kind in (1, 2):
        g = HostGrid(5, 4, 6)
        g.buf[:] = uniform(4, g.buf.shape)
        before = g.buf.copy()
        L().orc_bc_fill(kind, g.ptr(), C.byref(BCSet.uniform(T.BC_DIRICHELET, .25)), 1, 1, 1)
        F, B = g.full(), before.copy()[15:].reshape(g.full().shape)
        assert F[0, 0, 3] == B[0, 0, 3] and F[0, 2, 0] == B[0, 2, 0]
        assert (F[0, 1:5, 1:7] + F[1, 1:5, 1:7] == .5).all() or np.allclose(F[0, 1:5, 1:7] + F[1, 1:5, 1:7], .5)


def test_mac_fill_noslip_and_periodic():
    nx, ny, nz = 6, 5, 4
    u, v, w, _ = mac_grids(nx, ny, nz)
    for q, g in enumerate((u, v, w)):
        g.buf[:] = uniform(10 + q, g.buf.shape)
    bc = BCSet.uniform(T.BC_PERIODIC)
    bc.zneg = BC(T.BC_FORCED_INFLOW_VARIABLE_SLIP, 0.0, 1.0)
    bc.zpos = BC(T.BC_FORCED_INFLOW_VARIABLE_SLIP, 0.0, 1.0)
    assert L().orc_bc_fill_mac(u.ptr(), v.ptr(), w.ptr(), C.byref(bc)) == 1
    W = w.full()[1:-1]
    assert (W[:nx, 1:ny + 1, 1] == 0).all() and (W[:nx, 1:ny + 1, nz + 1] == 0).all()  # w = 0 on the walls
    U = u.full()
    # no slip: tangential ghost = -interior, so the wall value (average) is 0
    assert np.allclose(U[1:nx + 2, 1:ny + 1, 0] + U[1:nx + 2, 1:ny + 1, 1], 0)
    # periodic in x for u: u(nx) = u(0), u(nx+1) = u(1)
    assert (U[nx + 1, 1:ny + 1, 1:nz + 1] == U[1, 1:ny + 1, 1:nz + 1]).all()


# ------------------------------------------------ operator eigen-structure

def mode_field(n, m, kind):
    i = np.arange(n) + .5
    if kind == "neumann":
        return np.cos(math.pi * m * i / n), -4 * math.sin(math.pi * m / (2 * n)) ** 2
    if kind == "dirichlet":
        return np.sin(math.pi * m * i / n), -4 * math.sin(math.pi * m / (2 * n)) ** 2
    return np.cos(2 * math.pi * m * np.arange(n) / n), -4 * math.sin(math.pi * m / n) ** 2


@pytest.mark.parametrize("kind,bctype", [("neumann", T.BC_NEUMANN), ("dirichlet", T.BC_DIRICHELET),
                                         ("periodic", T.BC_PERIODIC)])
def test_residual_of_eigenmode(kind, bctype):
    """r = b - A u with b = 0 and u a separable discrete eigenvector equals
    -lambda u, lambda = sum_a f_a/h^2 * (-4 sin^2 ...)"""
    n = (12, 8, 16)
    m = (3, 2, 5)
    hx, hy, hz = .5, .25, .125
    h0 = min(hx, hy, hz)
    f = (h0 * h0 / hx ** 2, h0 * h0 / hy ** 2, h0 * h0 / hz ** 2)
    (ex, lx), (ey, ly), (ez, lz) = (mode_field(n[a], m[a], kind) for a in range(3))
    u = HostGrid(*n)
    u.interior()[:] = ex[:, None, None] * ey[None, :, None] * ez[None, None, :]
    L().orc_bc_fill(0, u.ptr(), C.byref(BCSet.uniform(bctype)), hx, hy, hz)
    b = HostGrid.congruent(u)
    r = HostGrid.congruent(u)
    L().orc_mg_residual(u.ptr(), b.ptr(), r.ptr(), h0, *f)
    lam = (f[0] * lx + f[1] * ly + f[2] * lz) / h0 ** 2
    want = -lam * u.interior()
    assert np.abs(r.interior() - want).max() <= 1e-12 * np.abs(want).max()


def dense_operator(n, h, bcs):
    """The cell-centred 7-point operator with ghosts eliminated, assembled
    independently: Neumann ghost = interior, Dirichlet-0 ghost = -interior,
    periodic ghost = opposite interior.  bcs: type per axis."""
    nx, ny, nz = n
    N = nx * ny * nz
    idx = lambda i, j, k: (i * ny + j) * nz + k
    rows, cols, vals = [], [], []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                c = idx(i, j, k)
                for a, (q, nq) in enumerate(((i, nx), (j, ny), (k, nz))):
                    w = 1.0 / h[a] ** 2
                    rows.append(c); cols.append(c); vals.append(-2 * w)
                    for s in (-1, 1):
                        t = [i, j, k]
                        t[a] = q + s
                        if 0 <= t[a] < nq:
                            rows.append(c); cols.append(idx(*t)); vals.append(w)
                        elif bcs[a] == T.BC_NEUMANN:
                            rows.append(c); cols.append(c); vals.append(w)
                        elif bcs[a] == T.BC_DIRICHELET:
                            rows.append(c); cols.append(c); vals.append(-w)
                        else:
                            t[a] = t[a] % nq
                            rows.append(c); cols.append(idx(*t)); vals.append(w)
    return sp.csr_matrix((vals, (rows, cols)), shape=(N, N))


def bc_axes(tx, ty, tz):
    s = BCSet()
    for f, t in zip(BCSet.FACES, (tx, tx, ty, ty, tz, tz)):
        setattr(s, f, BC(t, 0.0, 0.0))
    return s


@pytest.mark.parametrize("variant", [T.RELAX_DEVICE, T.RELAX_HOST])
@pytest.mark.parametrize("dtype_bits", [64])
def test_relax_fixed_point_is_the_direct_solution(variant, dtype_bits):
    """Red-black Gauss-Seidel sweeps (with the ghost refresh of relax())
    converge to the solution of an independently assembled sparse system."""
    n = (6, 4, 8)
    hx, hy, hz = .5, .25, .5
    axes = (T.BC_DIRICHELET, T.BC_PERIODIC, T.BC_NEUMANN)
    bc = bc_axes(*axes)
    A = dense_operator(n, (hx, hy, hz), axes)
    b = HostGrid(*n)
    b.interior()[:] = 2 * uniform(5, n) - 1
    x = spla.spsolve(A.tocsc(), b.interior().ravel()).reshape(n)
    u = HostGrid.congruent(b)
    p = O.level_params(bc, 0, hx, hy, hz)
    for _ in range(3000):
        for color in (0, 1):
            L().orc_mg_relax(u.ptr(), b.ptr(), color, C.byref(p), variant)
            L().orc_bc_fill(0, u.ptr(), C.byref(bc), hx, hy, hz)
    assert np.abs(u.interior() - x).max() <= 1e-11 * np.abs(x).max()


def test_restrict_and_prolong_closed_forms():
    r = HostGrid(8, 6, 10)
    r.buf[:] = uniform(6, r.buf.shape)
    bc_ = HostGrid(4, 3, 5)
    L().orc_mg_restrict(r.ptr(), bc_.ptr())
    assert abs(bc_.interior().sum() - r.interior().sum() / 8) < 1e-12
    want = r.interior().reshape(4, 2, 3, 2, 5, 2).mean(axis=(1, 3, 5))
    assert np.abs(bc_.interior() - want).max() < 1e-15
    # prolongation of a constant adds the constant; of a linear field adds
    # the linear field exactly, except in the reference's quirk octant
    uc = HostGrid(4, 3, 5)
    xc = (np.arange(-1, 5) + .5) * 2
    uc.full()[:] = xc[:, None, None] + 0 * uc.full()
    for quirk in (0, 1):
        uf = HostGrid(8, 6, 10)
        L().orc_mg_prolong(uc.ptr(), uf.ptr(), quirk)
        xf = np.arange(8) + .5
        err = np.abs(uf.interior() - xf[:, None, None])
        if quirk == 0:
            assert err.max() < 1e-13
        else:
            oct_mask = np.zeros(uf.interior().shape, bool)
            oct_mask[0::2, 1::2, 1::2] = True
            assert err[~oct_mask].max() < 1e-13 and err[oct_mask].max() > 0.01
    ucc = HostGrid(4, 3, 5)
    ucc.full()[:] = 2.5
    uf = HostGrid(8, 6, 10)
    L().orc_mg_prolong(ucc.ptr(), uf.ptr(), 1)
    assert np.abs(uf.interior() - 2.5).max() < 1e-15


# ----------------------------------------------------------- full solves

def test_mg_solves_eigenmode_rhs():
    n = (32, 32, 32)
    h = 1 / 32
    ex, lx = mode_field(32, 2, "neumann")
    ey, ly = mode_field(32, 3, "neumann")
    ez, lz = mode_field(32, 1, "neumann")
    mode = ex[:, None, None] * ey[None, :, None] * ez[None, None, :]
    lam = (lx + ly + lz) / h ** 2
    b = HostGrid(*n)
    b.interior()[:] = lam * mode
    u = HostGrid.congruent(b)
    ok, res, vc, hist = O.mg_solve(u, b, (h, h, h), MGConfig.make(BCSet.uniform(T.BC_NEUMANN)), 1e-10, 20)
    assert ok == 1
    x = u.interior() - u.interior().mean()
    assert np.abs(x - mode).max() < 1e-9
    assert all(hist[i + 1] < hist[i] for i in range(len(hist) - 1))


@pytest.mark.parametrize("axes", [(T.BC_NEUMANN,) * 3, (T.BC_DIRICHELET, T.BC_PERIODIC, T.BC_NEUMANN)])
def test_mg_matches_sparse_direct_solve(axes):
    n = (16, 16, 16)
    h = (1 / 16, 1 / 16, 1 / 8)
    bc = bc_axes(*axes)
    b = divergence_rhs(*n, 1 / 16)
    A = dense_operator(n, h, axes)
    rhs = b.interior().ravel()
    if all(a != T.BC_DIRICHELET for a in axes):
        A = A.tolil()
        A[0, :] = 0
        A[0, 0] = 1
        rhs = rhs.copy()
        rhs[0] = 0
        A = A.tocsc()
    x = spla.spsolve(A.tocsc(), rhs).reshape(n)
    u = HostGrid.congruent(b)
    ok, res, vc, _ = O.mg_solve(u, b, h, MGConfig.make(bc), 1e-11, 30)
    assert ok == 1
    got = u.interior()
    if all(a != T.BC_DIRICHELET for a in axes):
        got = got - got.mean()
        x = x - x.mean()
    assert np.abs(got - x).max() <= 1e-9 * np.abs(x).max()


def test_mg_all_dirichlet_zero_rhs_is_constant():
    """mgdoubletest.cpp:186-213"""
    b = HostGrid(16, 16, 16)
    u = HostGrid.congruent(b)
    u.interior()[:] = .5 - uniform(7, u.interior().shape)
    ok, res, _, _ = O.mg_solve(u, b, (1 / 16,) * 3, MGConfig.make(BCSet.uniform(T.BC_DIRICHELET, .5)), 1e-10, 15)
    assert ok == 1 and np.abs(u.interior() - .5).max() < 1e-9


def test_mixed_ir_oracle_reaches_fp64_solution():
    n = (32, 32, 32)
    h = 1 / 32
    bc = BCSet.uniform(T.BC_NEUMANN)
    b = divergence_rhs(*n, h)
    rms = np.sqrt((b.interior() ** 2).mean())
    ref = HostGrid.congruent(b)
    O.mg_solve(ref, b, (h,) * 3, MGConfig.make(bc), 1e-14 * rms, 40)
    x = HostGrid.congruent(b)
    ok, res, it, vc = O.ir_solve(x, b, (h,) * 3, MGConfig.make(bc), 1e-10 * rms, 10)
    assert ok == 1 and res <= 1e-10 * rms and 1 <= it <= 4
    xr = ref.interior() - ref.interior().mean()
    assert np.abs(x.interior() - xr).max() <= 1e-9 * np.abs(xr).max()


# ---------------------------------------------------- MAC-grid operators

def test_divergence_of_gradient_is_the_laplacian():
    nx, ny, nz = 8, 6, 10
    hx, hy, hz = .5, .25, .125
    u, v, w, p = mac_grids(nx, ny, nz)
    p.buf[:] = uniform(8, p.buf.shape)
    bc = BCSet.uniform(T.BC_PERIODIC)
    L().orc_bc_fill(0, p.ptr(), C.byref(bc), hx, hy, hz)
    # u = +grad p on the faces i<nx of every axis, then wrap the last faces
    L().orc_gradient_sub(u.ptr(), v.ptr(), w.ptr(), p.ptr(), 1.0, hx, hy, hz)
    L().orc_bc_fill_mac(u.ptr(), v.ptr(), w.ptr(), C.byref(bc))
    div = HostGrid.congruent(p)
    L().orc_divergence(u.ptr(), v.ptr(), w.ptr(), div.ptr(), nx, ny, nz, hx, hy, hz)
    lap = HostGrid.congruent(p)
    L().orc_laplacian_acc(p.ptr(), lap.ptr(), nx, ny, nz, hx, hy, hz, 1.0)
    assert np.abs(div.interior() - lap.interior()).max() <= 1e-12 * np.abs(lap.interior()).max()


def test_projection_is_divergence_free_and_idempotent():
    nx, ny, nz = 16, 8, 8
    h = (1 / 16, 1 / 8, 1 / 8)
    bc = BCSet.uniform(T.BC_PERIODIC)
    bc.zneg = BC(T.BC_FORCED_INFLOW_VARIABLE_SLIP, 0.0, 1.0)
    bc.zpos = BC(T.BC_FORCED_INFLOW_VARIABLE_SLIP, 0.0, 1.0)
    u, v, w, _ = random_mac_velocity(nx, ny, nz)
    p = HostGrid.congruent(u)
    r = C.c_double()
    assert L().orc_project(u.ptr(), v.ptr(), w.ptr(), C.byref(bc), nx, ny, nz, *h, 1e-12, p.ptr(), C.byref(r)) == 1
    div = HostGrid.congruent(p)
    L().orc_divergence(u.ptr(), v.ptr(), w.ptr(), div.ptr(), nx, ny, nz, *h)
    assert np.abs(div.interior()).max() < 1e-10
    u2, v2, w2 = u.copy(), v.copy(), w.copy()
    L().orc_project(u2.ptr(), v2.ptr(), w2.ptr(), C.byref(bc), nx, ny, nz, *h, 1e-12, p.ptr(), C.byref(r))
    for a, b_ in ((u, u2), (v, v2), (w, w2)):
        assert np.abs(a.interior() - b_.interior()).max() < 1e-10


def test_upwind_advection_is_an_exact_shift_at_cfl_one():
    nx, ny, nz = 8, 4, 4
    u, v, w, phi = mac_grids(nx, ny, nz)
    u.full()[:] = 1.0  # uniform +x velocity
    phi.interior()[:] = uniform(9, phi.interior().shape)
    bc = BCSet.uniform(T.BC_PERIODIC)
    L().orc_bc_fill(1, phi.ptr(), C.byref(bc), 1, 1, 1)
    before = phi.interior().copy()
    d = HostGrid.congruent(phi)
    assert L().orc_scalar_advect_step(u.ptr(), v.ptr(), w.ptr(), phi.ptr(), d.ptr(), C.byref(bc), nx, ny, nz,
                                      1.0, 1.0, 1.0, T.IT_FIRST_ORDER_UPWIND, 1.0) == 1
    assert np.abs(phi.interior() - np.roll(before, 1, axis=0)).max() < 1e-15


@pytest.mark.parametrize("scheme", [T.IT_FIRST_ORDER_UPWIND, T.IT_SECOND_ORDER_CENTERED])
def test_advection_conserves_periodic_mass_and_constants(scheme):
    nx, ny, nz = 8, 6, 10
    u, v, w, phi = random_mac_velocity(nx, ny, nz)
    bc = BCSet.uniform(T.BC_PERIODIC)
    L().orc_bc_fill_mac(u.ptr(), v.ptr(), w.ptr(), C.byref(bc))
    phi.interior()[:] = uniform(11, phi.interior().shape)
    L().orc_bc_fill(0, phi.ptr(), C.byref(bc), 1, 1, 1)
    d = HostGrid.congruent(phi)
    L().orc_passive_advect(u.ptr(), v.ptr(), w.ptr(), phi.ptr(), d.ptr(), nx, ny, nz, .5, .5, .5, scheme)
    assert abs(d.interior().sum()) < 1e-12 * np.abs(d.interior()).sum()
    du, dv, dw = (HostGrid.congruent(phi) for _ in range(3))
    L().orc_self_advect(u.ptr(), v.ptr(), w.ptr(), du.ptr(), dv.ptr(), dw.ptr(), nx, ny, nz, .5, .5, .5, scheme)
    assert abs(du.interior().sum()) < 1e-12 * np.abs(du.interior()).sum()
    # a constant scalar in a divergence-free field has no tendency
    c = HostGrid.congruent(phi)
    c.full()[:] = 3.0
    p = HostGrid.congruent(u)
    r = C.c_double()
    L().orc_project(u.ptr(), v.ptr(), w.ptr(), C.byref(bc), nx, ny, nz, .5, .5, .5, 1e-13, p.ptr(), C.byref(r))
    L().orc_passive_advect(u.ptr(), v.ptr(), w.ptr(), c.ptr(), d.ptr(), nx, ny, nz, .5, .5, .5, scheme)
    assert np.abs(d.interior()).max() < 1e-11


def test_rayleigh_critical_number_matches_reference_res16():
    """The oracle's NS step (advection, diffusion, thermal force, time-step
    rule, projection) reproduces the reference's measured Ra_cr of free-slip
    Rayleigh-Benard convection at resolution 16, 659.693989
    (tests/rayleigh.cpp:338), within the measurement's own resolution
    (5e-3; the discretisation error that number carries is 2.18)."""
    import rayleigh as R
    ra = R.critical_ra(R.OracleRun, 16, T.DIR_ZPOS, False, 657.0, 660.0, log=lambda s: None)
    assert abs(ra - R.RA_FREE_SLIP[16]) <= 5e-3, ra


@pytest.mark.parametrize("ts", [T.TS_FORWARD_EULER, T.TS_ADAMS_BASHFORD2])
def test_ns_time_stepping_of_a_diffusing_eigenmode(ts):
    """With zero velocity, no buoyancy and periodic faces the NS temperature
    obeys T' = kappa L T; a separable Fourier mode of the 7-point Laplacian
    (eigenvalue -sum 4/h^2 sin^2(pi m/n)) is multiplied each step by the
    amplification of the scheme: forward Euler a_{n+1} = (1 + dt mu) a_n;
    Adams-Bashforth 2 with ab = -dt^2 / (2 dt_last) (first step Euler,
    eqn_incompressns3d.cu:355-390) a_{n+1} = a_n + (dt - ab) mu a_n +
    ab mu a_{n-1}.  Unequal steps exercise dt_last."""
    from opencurrent_amd._lib import NSParams
    n, h, kappa = (16, 8, 16), (1 / 16, 1 / 8, 1 / 16), .7
    m = (1, 3, 2)
    p = NSParams()
    p.nx, p.ny, p.nz = n
    p.hx, p.hy, p.hz = h
    p.max_divergence = 1e-10
    p.temp_bc = BCSet.uniform(T.BC_PERIODIC)
    p.flow_bc = BCSet.uniform(T.BC_PERIODIC)
    p.vertical_direction = T.DIR_ZPOS
    p.gravity, p.bouyancy = 0.0, 0.0
    p.thermal_diffusion, p.viscosity = kappa, 1.0
    p.advection_scheme, p.time_step, p.cfl_factor = T.IT_FIRST_ORDER_UPWIND, ts, .5
    u, v, w, t = mac_grids(*n)
    x = [2 * math.pi * mq * (np.arange(nq) + .5) / nq for mq, nq in zip(m, n)]
    t.interior()[:] = np.cos(x[0])[:, None, None] * np.sin(x[1])[None, :, None] * np.cos(x[2] + .3)[None, None, :]
    t0 = t.interior().copy()
    mu = -kappa * sum(4 / hq ** 2 * math.sin(math.pi * mq / nq) ** 2 for hq, mq, nq in zip(h, m, n))
    ho = O.lib().orc_ns_create(T.OCU_F64, C.byref(p), u.ptr(), v.ptr(), w.ptr(), t.ptr())
    assert ho
    try:
        a_prev, a, last = None, 1.0, 0.0
        for dt in (2e-4, 2e-4, 3e-4, 1e-4, 2.5e-4):
            assert O.lib().orc_ns_advance_one_step(ho, dt) == 1
            if ts == T.TS_ADAMS_BASHFORD2 and last > 0:
                ab = -dt * dt / (2 * last)
                a_prev, a = a, a + (dt - ab) * mu * a + ab * mu * a_prev
            else:
                a_prev, a = a, a + dt * mu * a
            last = dt
            got = HostGrid.congruent(t)
            assert O.lib().orc_ns_get(ho, 3, got.ptr()) == 1
            assert np.abs(got.interior() - a * t0).max() <= 1e-13, (dt, a)
    finally:
        O.lib().orc_ns_destroy(ho)
