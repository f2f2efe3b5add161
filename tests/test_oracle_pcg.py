"""CPU pins of the oracle's Sol_PCGPressure3DDevice restatement
(oracle/oracle_pcg.h; reference ocuequation/sol_pcgpressure3d.cu:27-662).

The variable-coefficient operator, its diagonal and the dot product are
checked against a sparse matrix assembled here in numpy from the equation
-div(c grad p) with face coefficients (c_i + c_nb)/2 (not from the oracle's
code); the solver against scipy's direct solve of that matrix, and, for
c = 1, against the negated multigrid solution of the constant-coefficient
Laplacian (the reference's preconditioner relies on that sign relation)."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle_lib as O
from fields import uniform
from opencurrent_amd import _types as T
from opencurrent_amd._types import BCSet, HostGrid, MGConfig

PRECONDS = {"none": 0, "jacobi": 1, "multigrid": 2, "bfbt": 3}


def bcs(kind):
    if kind == "neumann":
        return BCSet.uniform(T.BC_NEUMANN)
    b = BCSet.uniform(T.BC_NEUMANN)
    b.xpos = T.BC(T.BC_PERIODIC, 0.0, 0.0)
    b.xneg = T.BC(T.BC_PERIODIC, 0.0, 0.0)
    return b


def assembled(c, h, periodic_x):
    """-div(c grad) on the interior: for each cell and each neighbour inside
    the box (or across a periodic x face), a face coefficient (c_i+c_nb)/2;
    homogeneous Neumann faces carry no flux."""
    n = c.shape
    N = int(np.prod(n))
    idx = np.arange(N).reshape(n)
    rows, cols, vals = [], [], []
    diag = np.zeros(n)
    for ax, ha in enumerate(h):
        w = 1.0 / (ha * ha)
        for s in (-1, 1):
            nb = np.roll(idx, -s, axis=ax)
            cn = np.roll(c, -s, axis=ax)
            valid = np.ones(n, bool)
            if not (ax == 0 and periodic_x):
                sl = [slice(None)] * 3
                sl[ax] = -1 if s == 1 else 0
                valid[tuple(sl)] = False
            a = w * .5 * (cn + c)
            diag += np.where(valid, a, 0)
            rows.append(idx[valid]); cols.append(nb[valid]); vals.append(-a[valid])
    rows.append(idx.ravel()); cols.append(idx.ravel()); vals.append(diag.ravel())
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))


def grids(n, seed, dtype=np.float64):
    c = HostGrid(*n, dtype=dtype)
    c.interior()[:] = .5 + uniform(seed, n)
    src = HostGrid(*n, dtype=dtype)
    src.interior()[:] = 2 * uniform(seed + 1, n) - 1
    return c, src


def fh(h):
    """the T values f/h^2 of the reference: h = min of the float spacings,
    f = h^2 / (hx*hx) with hx*hx a float product"""
    hm = float(min(h))
    return [(hm * hm / float(np.float32(x) * np.float32(x))) / (hm * hm) for x in h]


@pytest.mark.parametrize("kind", ["neumann", "periodic_x"])
def test_pcg_operator_and_diagonal_match_the_assembled_matrix(kind):
    n = (6, 5, 7)
    h = (np.float32(.2), np.float32(.25), np.float32(.2))
    bc = bcs(kind)
    c, src = grids(n, 9100)
    hd = [float(x) for x in h]
    for g in (c, src):
        assert O.lib().orc_bc_fill(1, g.ptr(), C.byref(bc), *hd) == 1
    fx, fy, fz = fh(h)
    dst = HostGrid(*n)
    assert O.lib().orc_pcg_apply_laplacian(dst.ptr(), src.ptr(), c.ptr(), fx, fy, fz) == 1
    A = assembled(c.interior(), [1 / np.sqrt(fx), 1 / np.sqrt(fy), 1 / np.sqrt(fz)], kind == "periodic_x")
    want = (A @ src.interior().ravel()).reshape(n)
    assert np.allclose(dst.interior(), want, rtol=1e-13, atol=1e-12 * np.abs(want).max())
    # the diagonal the Jacobi preconditioner divides by counts all six faces
    # (ghost coefficients included), not only those that carry flux
    cg = np.pad(c.interior(), 1, mode="edge")
    if kind == "periodic_x":
        cg[0, 1:-1, 1:-1], cg[-1, 1:-1, 1:-1] = c.interior()[-1], c.interior()[0]
    cc = cg[1:-1, 1:-1, 1:-1]
    diag = sum(f * .5 * (np.roll(cg, -s, axis=a)[1:-1, 1:-1, 1:-1] + cc)
               for a, f in enumerate((fx, fy, fz)) for s in (-1, 1))
    assert O.lib().orc_pcg_diag_precond(dst.ptr(), src.ptr(), c.ptr(), fx, fy, fz) == 1
    assert np.allclose(dst.interior(), src.interior() / diag, rtol=1e-14)


def test_pcg_dot_is_the_sum_of_products():
    a, b = grids((9, 4, 11), 9200)
    got = O.lib().orc_pcg_dot(a.ptr(), b.ptr())
    assert got == pytest.approx(float(np.sum(a.interior() * b.interior())), rel=1e-14)
    a32, b32 = HostGrid(9, 4, 11, dtype=np.float32), HostGrid(9, 4, 11, dtype=np.float32)
    a32.interior()[:], b32.interior()[:] = a.interior(), b.interior()
    prod = (a32.interior() * b32.interior()).astype(np.float64)   # products rounded to float
    assert O.lib().orc_pcg_dot(a32.ptr(), b32.ptr()) == np.float32(prod.sum())


def rhs_for(c, h, periodic_x, seed):
    n = c.shape
    A = assembled(c, h, periodic_x)
    b = 2 * uniform(seed, n) - 1
    b -= b.mean()   # consistent: the all-Neumann / periodic operator is singular
    return A, b


@pytest.mark.parametrize("precond", sorted(PRECONDS))
@pytest.mark.parametrize("kind", ["neumann", "periodic_x"])
def test_pcg_converges_to_the_direct_solution(precond, kind):
    """BFBT (the diagonal scaling after the multigrid cycles) is symmetric
    only when the diagonal is a constant, so it runs on a constant
    coefficient here; with a varying one CG stalls under it, in the
    reference's algorithm as in this restatement."""
    n = (16, 16, 16)
    h = (np.float32(1 / 16),) * 3
    bc = bcs(kind)
    c, _ = grids(n, 9300)
    if precond == "bfbt":
        c.interior()[:] = 2.5
    hd = [float(x) for x in h]
    fx = fh(h)[0]
    A, b = rhs_for(c.interior(), [1 / np.sqrt(fx)] * 3, kind == "periodic_x", 9301)
    rhs = HostGrid(*n)
    rhs.interior()[:] = b
    p = HostGrid(*n)
    tol = 1e-9 * np.sqrt((b * b).sum())
    ok, res, it = O.pcg_solve(p, rhs, c, hd, bc, PRECONDS[precond], tol=tol, max_iter=500)
    assert ok == 1, (precond, kind, res, it)
    x = p.interior().ravel()
    r = b.ravel() - A @ x
    # the preconditioned norm sqrt(<r, M^-1 r>) is below tol; the plain
    # residual follows within the preconditioner's scale
    assert np.linalg.norm(r) <= 1e-6 * np.linalg.norm(b), (precond, kind, it)
    # the direct solution of the singular system, pinned by its mean
    A1 = A.tolil()
    A1[0, :] = 0
    A1[0, 0] = 1
    bb = b.ravel().copy()
    bb[0] = x[0]
    xd = spla.spsolve(A1.tocsr(), bb)
    assert np.abs(x - xd).max() <= 1e-5 * np.abs(xd).max()
    if precond in ("multigrid", "bfbt"):
        assert it <= 12, it


def test_pcg_unit_coefficient_is_the_negated_multigrid_solution():
    """With c = 1 the operator is minus the multigrid's Laplacian, so PCG's
    answer is -p_MG up to the constant of the all-Neumann null space."""
    n = (16, 16, 16)
    h = 1 / 16
    bc = BCSet.uniform(T.BC_NEUMANN)
    c = HostGrid(*n)
    c.interior()[:] = 1.0
    rhs = HostGrid(*n)
    b = 2 * uniform(9400, n) - 1
    rhs.interior()[:] = b - b.mean()
    p = HostGrid(*n)
    tol = 1e-10 * float(np.sqrt((rhs.interior() ** 2).sum()))
    ok, res, it = O.pcg_solve(p, rhs, c, (h, h, h), bc, PRECONDS["bfbt"], tol=tol, max_iter=100)
    assert ok == 1
    q = HostGrid(*n)
    ok_mg, _, _, _ = O.mg_solve(q, rhs, (h, h, h), MGConfig.make(bc), 1e-13, 30)
    assert ok_mg == 1
    a = p.interior() - p.interior().mean()
    m = -(q.interior() - q.interior().mean())
    assert np.abs(a - m).max() <= 1e-8 * np.abs(m).max()
