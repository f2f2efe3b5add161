"""GPU Sol_PCGPressure3DDevice (ocuequation/sol_pcgpressure3d.cu:27-662)
against the oracle's restatement.

Kernels: the variable-coefficient operator and the Jacobi diagonal are
bit-exact against the oracle; dot products (FP64 sums of T products, in a
different order than the oracle's) agree to FP64 rounding.  Solver: every
preconditioner takes the oracle's iteration count and lands on its
solution; at 128^3 the residual is recomputed in numpy from the downloaded
pressure."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from fields import uniform
from opencurrent_amd import _lib
from opencurrent_amd import _types as T
from opencurrent_amd._types import BCSet, HostGrid
from opencurrent_amd.device import DeviceGrid

pytestmark = pytest.mark.gpu

PRECONDS = {"none": 0, "jacobi": 1, "multigrid": 2, "bfbt": 3}


def bcs(kind):
    b = BCSet.uniform(T.BC_NEUMANN)
    if kind == "periodic_x":
        b.xpos = T.BC(T.BC_PERIODIC, 0.0, 0.0)
        b.xneg = T.BC(T.BC_PERIODIC, 0.0, 0.0)
    return b


def f_div_hsq(h):
    hm = float(min(np.float32(x) for x in h))
    return [(hm * hm / float(np.float32(x) * np.float32(x))) / (hm * hm) for x in h]


def coeff_grid(n, seed, dtype, const=None):
    c = HostGrid(*n, dtype=dtype)
    c.interior()[:] = const if const is not None else .5 + uniform(seed, n)
    return c


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("kind", ["neumann", "periodic_x"])
@pytest.mark.parametrize("n", [(13, 9, 70), (16, 16, 16), (7, 20, 300)])
def test_pcg_kernels_match_oracle(gpu, n, kind, dtype):
    print(f"pcg kernels n={n} kind={kind} dtype={np.dtype(dtype).name}", flush=True)
    hip = _lib.hip()
    import torch
    bc = bcs(kind)
    h = (.1, .125, .1)
    c = coeff_grid(n, 9500, dtype)
    src = HostGrid(*n, dtype=dtype)
    src.interior()[:] = 2 * uniform(9501, n) - 1
    for g in (c, src):
        O.lib().orc_bc_fill(1, g.ptr(), C.byref(bc), *h)
    fx, fy, fz = f_div_hsq(h)
    if dtype == np.float32:
        fx, fy, fz = (float(np.float32(v)) for v in (fx, fy, fz))
    want = HostGrid.congruent(src)
    assert O.lib().orc_pcg_apply_laplacian(want.ptr(), src.ptr(), c.ptr(), fx, fy, fz) == 1
    gs, gc = DeviceGrid.from_host(src), DeviceGrid.from_host(c)
    gd = DeviceGrid.congruent(gs)
    ws = torch.zeros(hip.ocu_reduce_workspace_bytes() // 8 + 1, dtype=torch.float64, device="cuda")
    out = C.cast(ws.data_ptr() + 8 * (ws.numel() - 1), _lib.DP)
    _lib.check(hip.ocu_pcg_apply_laplacian(gd.ptr(), gs.ptr(), gc.ptr(), fx, fy, fz, ws.data_ptr(), out, None),
               "pcg laplacian")
    got = gd.to_host()
    assert np.array_equal(got.buf, want.buf)
    dot = ws[-1].item()
    ref = O.lib().orc_pcg_dot(src.ptr(), want.ptr())
    scale = float(np.abs(src.interior().astype(np.float64) * want.interior()).sum())
    assert abs(np.dtype(dtype).type(dot) - ref) <= 1e-13 * scale + (abs(ref) * 2e-7 if dtype == np.float32 else 0)
    # the dot product entry point on its own
    _lib.check(hip.ocu_grid_dot(gs.ptr(), gd.ptr(), ws.data_ptr(), out, None), "grid dot")
    assert ws[-1].item() == dot
    # Jacobi diagonal, out of place and in place
    want2 = HostGrid.congruent(src)
    assert O.lib().orc_pcg_diag_precond(want2.ptr(), src.ptr(), c.ptr(), fx, fy, fz) == 1
    gd2 = DeviceGrid.congruent(gs)
    _lib.check(hip.ocu_pcg_diag_precond(gd2.ptr(), gs.ptr(), gc.ptr(), fx, fy, fz, None), "diag")
    assert np.array_equal(gd2.to_host().buf, want2.buf)
    _lib.check(hip.ocu_pcg_diag_precond(gs.ptr(), gs.ptr(), gc.ptr(), fx, fy, fz, None), "diag in place")
    assert np.array_equal(gs.to_host().interior(), want2.interior())


CASES = [("none", "neumann", (16, 16, 16)), ("jacobi", "periodic_x", (20, 12, 18)),
         ("multigrid", "neumann", (24, 20, 18)), ("multigrid", "periodic_x", (16, 16, 16)),
         ("bfbt", "neumann", (16, 16, 16))]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("precond,kind,n", CASES)
def test_pcg_solve_matches_oracle(gpu, precond, kind, n, dtype):
    """Same iteration count and the same solution to rounding (the dot
    products' FP64 sums run in different orders)."""
    from opencurrent_amd.solvers import PCGPressure3D
    print(f"pcg solve {precond} {kind} n={n} dtype={np.dtype(dtype).name}", flush=True)
    bc = bcs(kind)
    h = (1 / n[0], 1 / n[0], 1 / n[0])
    c = coeff_grid(n, 9600, dtype, const=2.5 if precond == "bfbt" else None)
    rhs = HostGrid(*n, dtype=dtype)
    b = 2 * uniform(9601, n) - 1
    rhs.interior()[:] = b - b.mean()
    p_o = HostGrid.congruent(rhs)
    c_o = HostGrid.congruent(c)
    c_o.buf[:] = c.buf
    # the tolerance applies to sqrt(r.z), whose scale the preconditioner
    # sets: take it from the initial residual (a solve that stops at once)
    _, res0, it0 = O.pcg_solve(HostGrid.congruent(rhs), rhs, c_o.copy(), h, bc, PRECONDS[precond], tol=1e30,
                               max_iter=1)
    assert it0 == 0 and res0 > 0
    tol = (1e-8 if dtype == np.float64 else 1e-3) * res0
    ok_o, res_o, it_o = O.pcg_solve(p_o, rhs, c_o, h, bc, PRECONDS[precond], tol=tol, max_iter=400)
    grhs, gc = DeviceGrid.from_host(rhs), DeviceGrid.from_host(c)
    s = PCGPressure3D(grhs, gc, *h, bc, preconditioner=PRECONDS[precond])
    ok, res = s.solve(tol, 400)
    assert ok == bool(ok_o)
    if dtype == np.float64:
        assert ok
        assert s.iterations == it_o
        assert res == pytest.approx(res_o, rel=1e-6)
    else:
        assert abs(s.iterations - it_o) <= 1
    x, xo = s.pressure().interior().astype(np.float64), p_o.interior().astype(np.float64)
    err = np.abs((x - x.mean()) - (xo - xo.mean())).max() / np.abs(xo).max()
    assert err <= (1e-9 if dtype == np.float64 else 1e-3), err
    # the coefficient's ghosts were filled in place, as by the oracle
    assert np.array_equal(gc.to_host().buf, c_o.buf)


def neg_div_c_grad(p, c, h, periodic_x):
    """-div(c grad p) with face coefficients (c_i + c_nb)/2, homogeneous
    Neumann faces (no flux) or a periodic x pair, in numpy"""
    out = np.zeros_like(p)
    for ax, ha in enumerate(h):
        w = 1.0 / (ha * ha)
        if ax == 0 and periodic_x:
            cf = .5 * (c + np.roll(c, -1, axis=0))
            flux = w * cf * (np.roll(p, -1, axis=0) - p)          # face i+1/2
            out -= flux - np.roll(flux, 1, axis=0)
            continue
        sl_hi = [slice(None)] * 3
        sl_lo = [slice(None)] * 3
        sl_hi[ax], sl_lo[ax] = slice(1, None), slice(None, -1)
        cf = .5 * (c[tuple(sl_hi)] + c[tuple(sl_lo)])
        flux = w * cf * (p[tuple(sl_hi)] - p[tuple(sl_lo)])     # interior faces
        out[tuple(sl_lo)] -= flux
        out[tuple(sl_hi)] += flux
    return out


def test_pcg_multigrid_preconditioned_128(gpu):
    """128^3 FP64, smoothly varying coefficient, all-Neumann: the multigrid
    preconditioner converges in a handful of iterations; the residual of the
    downloaded pressure, recomputed in numpy, meets the tolerance's scale."""
    from opencurrent_amd.solvers import PCGPressure3D
    n = (128, 128, 128)
    h = (1 / 128,) * 3
    bc = BCSet.uniform(T.BC_NEUMANN)
    x = (np.arange(128) + .5) / 128
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    c = HostGrid(*n)
    c.interior()[:] = 1 + .5 * np.sin(2 * np.pi * X) * np.cos(np.pi * Y) + .25 * Z
    rhs = HostGrid(*n)
    b = 2 * uniform(9700, n) - 1
    rhs.interior()[:] = b - b.mean()
    norm_b = float(np.sqrt((rhs.interior() ** 2).sum()))
    grhs, gc = DeviceGrid.from_host(rhs), DeviceGrid.from_host(c)
    s = PCGPressure3D(grhs, gc, *h, bc, preconditioner=PCGPressure3D.MULTIGRID)
    ok, res = s.solve(1e-10 * norm_b, 100)
    assert ok, res
    assert s.iterations <= 15, s.iterations
    p = s.pressure().interior()
    hh = [float(np.float32(v)) for v in h]
    r = rhs.interior() - neg_div_c_grad(p, c.interior(), hh, False)
    assert np.sqrt((r ** 2).sum()) <= 1e-6 * norm_b
    assert np.all(np.isfinite(p))
