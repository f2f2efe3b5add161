"""Kernel-level parity: every HIP entry point against the CPU oracle on the
same seeded inputs.  The bar is bit-exact: both sides perform the same IEEE
operations in the same order (the library and the oracle are compiled
without FMA contraction), so whole buffers - interior, ghosts and padding -
must agree bit for bit."""
import ctypes as C

import numpy as np
import pytest

from opencurrent_amd import _types as T
from opencurrent_amd._types import BC, BCSet, HostGrid
from fields import uniform, mac_grids
import oracle_lib as O

pytestmark = pytest.mark.gpu


def D():
    from opencurrent_amd import device
    return device


def assert_bitexact(got, want, what=""):
    a, b = got.buf, want.buf
    assert a.dtype == b.dtype and a.shape == b.shape, what
    ia = a.view(np.uint64 if a.dtype == np.float64 else np.uint32)
    ib = b.view(np.uint64 if b.dtype == np.float64 else np.uint32)
    bad = np.nonzero(ia != ib)[0]
    if bad.size:
        n = bad[0]
        raise AssertionError(f"{what}: {bad.size} elements differ, first at {n}: got {a[n]!r} want {b[n]!r}")


def random_grid(nx, ny, nz, dtype, seed, g=1, pad=(0, 0, 0)):
    h = HostGrid(nx, ny, nz, g, g, g, *pad, dtype=dtype)
    h.buf[:] = (2 * uniform(seed, h.buf.shape) - 1).astype(dtype)
    return h


def bcset(xp, xn, yp, yn, zp, zn):
    s = BCSet()
    for f, spec in zip(BCSet.FACES, (xp, xn, yp, yn, zp, zn)):
        typ, val = spec if isinstance(spec, tuple) else (spec, 0.0)
        setattr(s, f, BC(typ, val, 0.0))
    return s


BC_CASES = {
    "neumann0": bcset(*[T.BC_NEUMANN] * 6),
    "neumann_v": bcset((T.BC_NEUMANN, .37), (T.BC_NEUMANN, -.25), (T.BC_NEUMANN, 1.5), (T.BC_NEUMANN, .1),
                       (T.BC_NEUMANN, -2.0), (T.BC_NEUMANN, .321)),
    "dirichlet": bcset(*[(T.BC_DIRICHELET, .5)] * 6),
    "mixed": bcset(T.BC_PERIODIC, T.BC_PERIODIC, (T.BC_DIRICHELET, .321), (T.BC_DIRICHELET, -1.0),
                   (T.BC_NEUMANN, .7), (T.BC_NEUMANN, 0.0)),
    "periodic": bcset(*[T.BC_PERIODIC] * 6),
    "none_faces": bcset(T.BC_NONE, (T.BC_DIRICHELET, 1.0), T.BC_NONE, (T.BC_NEUMANN, .2), T.BC_PERIODIC,
                        T.BC_PERIODIC),
}
SIZES = [(16, 12, 20), (7, 9, 5), (33, 8, 64)]
DTYPES = [np.float64, np.float32]


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("size", SIZES)
@pytest.mark.parametrize("case", list(BC_CASES))
def test_bc_fill_level1(gpu, dtype, size, case):
    h = random_grid(*size, dtype, 11)
    bc = BC_CASES[case]
    want = h.copy()
    assert O.lib().orc_bc_fill(0, want.ptr(), C.byref(bc), .3, .25, .125) == 1
    g = D().DeviceGrid.from_host(h)
    D().bc_fill_level1(g, bc, .3, .25, .125)
    assert_bitexact(g.to_host(), want, f"level1 {case} {size}")


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("case", list(BC_CASES))
def test_bc_fill_nocorners(gpu, dtype, case):
    h = random_grid(13, 10, 17, dtype, 12)
    bc = BC_CASES[case]
    want = h.copy()
    assert O.lib().orc_bc_fill(1, want.ptr(), C.byref(bc), .5, .125, .25) == 1
    g = D().DeviceGrid.from_host(h)
    D().bc_fill_nocorners(g, bc, .5, .125, .25)
    assert_bitexact(g.to_host(), want, f"nocorners {case}")


MAC_CASES = {
    "noslip": bcset(*[(T.BC_FORCED_INFLOW_VARIABLE_SLIP, 0.0)] * 6),
    "rb": None,  # periodic x/y, no-slip z (aux 1)
    "inflow": None,
    "periodic": bcset(*[T.BC_PERIODIC] * 6),
}


def mac_bc(case):
    if case == "rb":
        s = bcset(T.BC_PERIODIC, T.BC_PERIODIC, T.BC_PERIODIC, T.BC_PERIODIC, T.BC_FORCED_INFLOW_VARIABLE_SLIP,
                  T.BC_FORCED_INFLOW_VARIABLE_SLIP)
        s.zpos.aux_value = 1.0
        s.zneg.aux_value = 1.0
        return s
    if case == "inflow":
        s = bcset((T.BC_FORCED_INFLOW_VARIABLE_SLIP, .75), (T.BC_FORCED_INFLOW_VARIABLE_SLIP, .75),
                  T.BC_PERIODIC, T.BC_PERIODIC, (T.BC_FORCED_INFLOW_VARIABLE_SLIP, 0.0),
                  (T.BC_FORCED_INFLOW_VARIABLE_SLIP, 0.0))
        s.xpos.aux_value = .3
        s.zneg.aux_value = 1.0
        return s
    return MAC_CASES[case]


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("case", list(MAC_CASES))
def test_bc_fill_mac(gpu, dtype, case):
    u, v, w, _ = mac_grids(12, 9, 14, dtype)
    for q, g in enumerate((u, v, w)):
        g.buf[:] = (2 * uniform(20 + q, g.buf.shape) - 1).astype(dtype)
    bc = mac_bc(case)
    wu, wv, ww = u.copy(), v.copy(), w.copy()
    assert O.lib().orc_bc_fill_mac(wu.ptr(), wv.ptr(), ww.ptr(), C.byref(bc)) == 1
    du, dv, dw = (D().DeviceGrid.from_host(x) for x in (u, v, w))
    D().bc_fill_mac(du, dv, dw, bc)
    for got, want, n in ((du, wu, "u"), (dv, wv, "v"), (dw, ww, "w")):
        assert_bitexact(got.to_host(), want, f"mac {case} {n}")


def filled(nx, ny, nz, dtype, seed, bc, h):
    g = random_grid(nx, ny, nz, dtype, seed)
    O.lib().orc_bc_fill(0, g.ptr(), C.byref(bc), h, h, h)
    return g


RELAX_BCS = ["neumann0", "dirichlet", "mixed", "neumann_v"]


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("variant", [T.RELAX_DEVICE, T.RELAX_HOST])
@pytest.mark.parametrize("size", [(16, 16, 16), (12, 10, 7), (32, 8, 24)])
@pytest.mark.parametrize("case", RELAX_BCS)
def test_relax_colour(gpu, dtype, variant, size, case):
    bc = BC_CASES[case]
    u = filled(*size, dtype, 31, bc, 1 / 16)
    b = random_grid(*size, dtype, 32)
    for level in (0, 1):
        p = O.level_params(bc, level, 1 / 16, 1 / 12, 1 / 20)
        for color in (0, 1):
            want = u.copy()
            O.lib().orc_mg_relax(want.ptr(), b.ptr(), color, C.byref(p), variant)
            gu, gb = D().DeviceGrid.from_host(u), D().DeviceGrid.from_host(b)
            D().mg_relax(gu, gb, color, p, variant)
            assert_bitexact(gu.to_host(), want, f"relax {case} l{level} c{color}")


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("size", [
    # whole-row kernel: one, two or four cells per lane, rows that fill the
    # wave or stop short of it
    (16, 16, 16), (9, 20, 13), (64, 64, 64), (40, 30, 70), (20, 40, 66), (10, 18, 128), (12, 16, 200),
    (8, 26, 256),
    # tiled kernel: rows that do not split evenly over the lanes or exceed a wave
    (6, 14, 130), (7, 15, 101), (5, 13, 300)])
@pytest.mark.parametrize("case", ["neumann0", "dirichlet", "neumann_v"])
@pytest.mark.parametrize("first", [0, 1])
def test_relax_fused_matches_two_colours(gpu, dtype, size, case, first):
    bc = BC_CASES[case]
    u = filled(*size, dtype, 41, bc, 1 / 32)
    b = random_grid(*size, dtype, 42)
    p = O.level_params(bc, 0, 1 / 32, 1 / 32, 1 / 16)
    want = u.copy()
    O.lib().orc_mg_relax(want.ptr(), b.ptr(), first, C.byref(p), T.RELAX_DEVICE)
    O.lib().orc_mg_relax(want.ptr(), b.ptr(), 1 - first, C.byref(p), T.RELAX_DEVICE)
    gin, gb = D().DeviceGrid.from_host(u), D().DeviceGrid.from_host(b)
    gout = D().DeviceGrid.from_host(u)  # ghosts of the output are not written: start from the same ghosts
    D().mg_relax_fused(gin, gout, gb, p, bc, first)
    assert_bitexact(gout.to_host(), want, f"fused {case} {size}")


FUSED_SIZES = [(16, 16, 16), (9, 20, 13), (40, 30, 70), (12, 16, 200), (8, 26, 256), (6, 14, 130), (5, 13, 300),
               (1, 3, 2)]


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("size", FUSED_SIZES)
@pytest.mark.parametrize("case", ["neumann0", "dirichlet", "neumann_v"])
@pytest.mark.parametrize("flags", [T.FUSED_WRITE_GHOSTS, T.FUSED_ZERO_INPUT, T.FUSED_ZERO_INPUT | T.FUSED_WRITE_GHOSTS])
def test_relax_fused_bc(gpu, dtype, size, case, flags):
    """One fused iteration that also writes the level-1 ghosts of its output
    and/or starts from the cleared, ghost-filled grid without reading it:
    equal to clear_zero + level-1 fill + two colours + level-1 fill."""
    bc = BC_CASES[case]
    hx, hy, hz = 1 / 32, 1 / 20, 1 / 16
    zero = bool(flags & T.FUSED_ZERO_INPUT)
    if zero:  # the zero-input mode is for homogeneous faces (coarse levels)
        bc = homogeneous(bc)
        u = HostGrid(*size, dtype=dtype)
        O.lib().orc_bc_fill(0, u.ptr(), C.byref(bc), hx, hy, hz)
    else:
        u = filled(*size, dtype, 43, bc, 1 / 32)
    b = random_grid(*size, dtype, 44)
    p = O.level_params(bc, 1, hx, hy, hz)
    for first in (0, 1):
        want = u.copy()
        O.lib().orc_mg_relax(want.ptr(), b.ptr(), first, C.byref(p), T.RELAX_DEVICE)
        O.lib().orc_mg_relax(want.ptr(), b.ptr(), 1 - first, C.byref(p), T.RELAX_DEVICE)
        if flags & T.FUSED_WRITE_GHOSTS:
            O.lib().orc_bc_fill(0, want.ptr(), C.byref(bc), hx, hy, hz)
        # zero input is not read: hand the kernel garbage there
        gin = D().DeviceGrid.from_host(random_grid(*size, dtype, 45) if zero else u)
        gb = D().DeviceGrid.from_host(b)
        start = u.copy()
        if flags & T.FUSED_WRITE_GHOSTS:
            start.buf[:] = (2 * uniform(46, start.buf.shape) - 1).astype(dtype)  # ghosts must all be written
            start.interior()[:] = u.interior()
        gout = D().DeviceGrid.from_host(start)
        D().mg_relax_fused_bc(gin, gout, gb, p, bc, hx, hy, hz, first, flags)
        got = gout.to_host()
        if not (flags & T.FUSED_WRITE_GHOSTS):
            assert np.array_equal(got.interior(), want.interior()), f"fused {case} {size} flags={flags}"
        else:
            assert_ghost_layer_bitexact(got, want, f"fused+ghosts {case} {size} flags={flags} first={first}")
        if zero:  # u_in may be the output itself
            galias = D().DeviceGrid.from_host(start)
            D().mg_relax_fused_bc(galias, galias, gb, p, bc, hx, hy, hz, first, flags)
            a = galias.to_host()
            assert np.array_equal(a.buf.view(np.uint8), got.buf.view(np.uint8)), "aliased zero-input call"


PERIODIC_FUSED_CASES = {
    "per_xy": bcset(T.BC_PERIODIC, T.BC_PERIODIC, T.BC_PERIODIC, T.BC_PERIODIC, (T.BC_NEUMANN, .7),
                    (T.BC_NEUMANN, 0.0)),
    "per_x": bcset(T.BC_PERIODIC, T.BC_PERIODIC, (T.BC_DIRICHELET, .321), (T.BC_DIRICHELET, -1.0),
                   (T.BC_NEUMANN, .25), (T.BC_DIRICHELET, .5)),
    "per_y": bcset((T.BC_NEUMANN, .1), (T.BC_DIRICHELET, .3), T.BC_PERIODIC, T.BC_PERIODIC, (T.BC_DIRICHELET, 1.0),
                   (T.BC_NEUMANN, 0.0)),
}


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("size", [(16, 12, 20), (40, 30, 70), (8, 26, 256), (12, 16, 200), (2, 4, 33), (6, 2, 64),
                                  (4, 34, 128)])
@pytest.mark.parametrize("case", list(PERIODIC_FUSED_CASES))
@pytest.mark.parametrize("flags", [T.FUSED_WRITE_GHOSTS, 0, T.FUSED_ZERO_INPUT | T.FUSED_WRITE_GHOSTS])
def test_relax_fused_periodic_pairs(gpu, dtype, size, case, flags):
    """periodic x and/or y pairs: one fused iteration against the reference's
    two colours with the level-1 fill between them (and after them, with
    ghost writes), over the interior and the whole ghost layer"""
    bc = PERIODIC_FUSED_CASES[case]
    zero = bool(flags & T.FUSED_ZERO_INPUT)
    if zero:
        bc = homogeneous(bc)
    hx, hy, hz = 1 / 32, 1 / 20, 1 / 16
    u = HostGrid(*size, dtype=dtype)
    if not zero:
        u.interior()[:] = (2 * uniform(47, u.interior().shape) - 1).astype(dtype)
    O.lib().orc_bc_fill(0, u.ptr(), C.byref(bc), hx, hy, hz)
    b = random_grid(*size, dtype, 48)
    p = O.level_params(bc, 0, hx, hy, hz)
    # zero input is not read: hand the kernel garbage there
    uin = random_grid(*size, dtype, 50) if zero else u
    gin, gb = D().DeviceGrid.from_host(uin), D().DeviceGrid.from_host(b)
    assert D().mg_relax_fused_ok(gin, gin, gb, bc)
    for first in (0, 1):
        want = u.copy()
        O.lib().orc_mg_relax(want.ptr(), b.ptr(), first, C.byref(p), T.RELAX_DEVICE)
        O.lib().orc_bc_fill(0, want.ptr(), C.byref(bc), hx, hy, hz)
        O.lib().orc_mg_relax(want.ptr(), b.ptr(), 1 - first, C.byref(p), T.RELAX_DEVICE)
        if flags & T.FUSED_WRITE_GHOSTS:
            O.lib().orc_bc_fill(0, want.ptr(), C.byref(bc), hx, hy, hz)
        start = u.copy()
        start.buf[:] = (2 * uniform(49, start.buf.shape) - 1).astype(dtype)  # ghosts must all be written
        start.interior()[:] = u.interior()
        gout = D().DeviceGrid.from_host(start)
        D().mg_relax_fused_bc(gin, gout, gb, p, bc, hx, hy, hz, first, flags)
        got = gout.to_host()
        if flags & T.FUSED_WRITE_GHOSTS:
            assert_ghost_layer_bitexact(got, want, f"fused periodic {case} {size} first={first}")
        else:
            assert np.array_equal(got.interior(), want.interior()), f"fused periodic {case} {size} first={first}"
    assert_bitexact(gin.to_host(), uin, "input untouched")


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("cn", [(3, 5, 64), (4, 3, 128), (2, 2, 50), (5, 1, 33)])
@pytest.mark.parametrize("case", list(PERIODIC_FUSED_CASES))
def test_prolong_bc_periodic_pairs(gpu, dtype, cn, case):
    """prolongation with ghost writes on periodic x/y pairs against the
    prolongation followed by the level-1 fill, over the whole ghost layer"""
    bc = PERIODIC_FUSED_CASES[case]
    hf = (.1, .2, .05)
    uc = filled(*cn, dtype, 191, bc, .2)
    uf = random_grid(*(2 * s for s in cn), dtype, 192)
    want = uf.copy()
    O.lib().orc_mg_prolong(uc.ptr(), want.ptr(), 1)
    O.lib().orc_bc_fill(0, want.ptr(), C.byref(bc), *hf)
    guc, guf = D().DeviceGrid.from_host(uc), D().DeviceGrid.from_host(uf)
    assert D().mg_prolong_bc_ok(guc, guf, bc)
    D().mg_prolong_bc(guc, guf, bc, *hf)
    assert_ghost_layer_bitexact(guf.to_host(), want, f"prolong periodic {case} {cn}")


def test_relax_fused_periodic_rejects(gpu):
    """odd counts along a periodic axis, periodic z and rows the whole-row
    kernel does not take stay on the colour path"""
    from opencurrent_amd._lib import OcuError
    bc = PERIODIC_FUSED_CASES["per_xy"]
    p = O.level_params(bc, 0, .1, .1, .1)
    for size, b_ in (((5, 4, 16), bc), ((4, 6, 300), bc), ((4, 4, 16), BC_CASES["periodic"])):
        u, o, b = (D().DeviceGrid(*size) for _ in range(3))
        assert not D().mg_relax_fused_ok(u, o, b, b_)
        with pytest.raises(OcuError):
            D().mg_relax_fused_bc(u, o, b, p, b_, .1, .1, .1, 0, T.FUSED_WRITE_GHOSTS)


def homogeneous(bc):
    h = BCSet()
    for f in BCSet.FACES:
        b = getattr(bc, f)
        setattr(h, f, BC(b.type, 0.0, b.aux_value))
    return h


def test_zero_input_rejects_inhomogeneous_faces(gpu):
    from opencurrent_amd._lib import OcuError
    bc = BC_CASES["dirichlet"]
    u = D().DeviceGrid(8, 8, 8, dtype=np.float32)
    o = D().DeviceGrid(8, 8, 8, dtype=np.float32)
    p = O.level_params(bc, 0, .1, .1, .1)
    with pytest.raises(OcuError, match="homogeneous"):
        D().mg_relax_fused_bc(u, o, u, p, bc, .1, .1, .1, 0, T.FUSED_ZERO_INPUT)


def assert_ghost_layer_bitexact(got, want, what):
    """interior and the first ghost layer (what the level-1 fill defines)"""
    d = got.desc
    box = (slice(d.gx - 1, d.gx + d.nx + 1), slice(d.gy - 1, d.gy + d.ny + 1), slice(d.gz - 1, d.gz + d.nz + 1))
    a, b = got.full()[box], want.full()[box]
    ia = a.view(np.uint64 if a.dtype == np.float64 else np.uint32)
    ib = b.view(np.uint64 if b.dtype == np.float64 else np.uint32)
    bad = np.argwhere(ia != ib)
    if bad.size:
        i, j, k = bad[0]
        raise AssertionError(f"{what}: {len(bad)} cells differ, first at ghost-box index {(i, j, k)}: "
                             f"got {a[i, j, k]!r} want {b[i, j, k]!r}")


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("case", ["neumann0", "dirichlet", "neumann_v"])
@pytest.mark.parametrize("cn", [(8, 6, 10), (1, 2, 3), (16, 16, 32), (3, 5, 128), (4, 3, 100), (2, 9, 66),
                                (5, 4, 31)])
def test_prolong_bc(gpu, dtype, case, cn):
    bc = BC_CASES[case]
    hx, hy, hz = .25, .125, .5
    uc = filled(*cn, dtype, 63, bc, .1)
    uf = random_grid(*(2 * s for s in cn), dtype, 64)
    want = uf.copy()
    O.lib().orc_mg_prolong(uc.ptr(), want.ptr(), 1)
    O.lib().orc_bc_fill(0, want.ptr(), C.byref(bc), hx, hy, hz)
    guc, guf = D().DeviceGrid.from_host(uc), D().DeviceGrid.from_host(uf)
    D().mg_prolong_bc(guc, guf, bc, hx, hy, hz)
    assert_ghost_layer_bitexact(guf.to_host(), want, f"prolong+ghosts {case} {cn}")


def test_fused_bc_rejects_nonlocal_faces(gpu):
    from opencurrent_amd._lib import OcuError
    bc = BC_CASES["none_faces"]
    u = D().DeviceGrid(8, 8, 8, dtype=np.float32)
    o = D().DeviceGrid(8, 8, 8, dtype=np.float32)
    p = O.level_params(BC_CASES["neumann0"], 0, .1, .1, .1)
    bc2 = bcset(T.BC_NONE, T.BC_NEUMANN, T.BC_NEUMANN, T.BC_NEUMANN, T.BC_NEUMANN, T.BC_NEUMANN)
    with pytest.raises(OcuError):
        D().mg_relax_fused_bc(u, o, u, p, bc2, .1, .1, .1, 0, T.FUSED_WRITE_GHOSTS)
    with pytest.raises(OcuError):
        D().mg_prolong_bc(D().DeviceGrid(4, 4, 4, dtype=np.float32), u, bc, .1, .1, .1)


def test_relax_fused_rejects_periodic_z(gpu):
    from opencurrent_amd._lib import OcuError
    bc = BC_CASES["periodic"]
    u = D().DeviceGrid(8, 8, 8)
    o = D().DeviceGrid(8, 8, 8)
    b = D().DeviceGrid(8, 8, 8)
    p = O.level_params(bc, 0, .1, .1, .1)
    with pytest.raises(OcuError):
        D().mg_relax_fused(u, o, b, p, bc, 0)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("size", [(16, 16, 16), (32, 8, 24), (64, 64, 64), (16, 10, 256), (8, 6, 200),
                                  (12, 12, 132), (2, 2, 2), (4, 34, 130)])
def test_residual_restrict(gpu, dtype, size):
    bc = BC_CASES["mixed"]
    h = 1 / 32
    u = filled(*size, dtype, 51, bc, h)
    b = random_grid(*size, dtype, 52)
    r_want = HostGrid.congruent(u)
    O.lib().orc_mg_residual(u.ptr(), b.ptr(), r_want.ptr(), h, 1.0, .64, .25)
    cn = tuple(s // 2 for s in size)
    bc_want = HostGrid(*cn, dtype=dtype)
    O.lib().orc_mg_restrict(r_want.ptr(), bc_want.ptr())
    gu, gb = D().DeviceGrid.from_host(u), D().DeviceGrid.from_host(b)
    # residual alone
    gr = D().DeviceGrid.congruent(gu)
    D().mg_residual(gu, gb, gr, h, 1.0, .64, .25)
    assert_bitexact(gr.to_host(), r_want, "residual")
    # restriction alone
    gbc = D().DeviceGrid(*cn, dtype=dtype)
    D().mg_restrict(D().DeviceGrid.from_host(r_want), gbc)
    assert_bitexact(gbc.to_host(), bc_want, "restrict")
    # fused, with and without r, with norms
    gr2 = D().DeviceGrid.congruent(gu)
    gbc2 = D().DeviceGrid(*cn, dtype=dtype)
    s2, mx = D().mg_residual_restrict(gu, gb, gr2, gbc2, h, 1.0, .64, .25, with_norm=True)
    assert_bitexact(gbc2.to_host(), bc_want, "fused restrict")
    assert_bitexact(gr2.to_host(), r_want, "fused residual")
    ri = r_want.interior().astype(np.float64)
    assert mx == np.abs(ri).max()
    assert abs(s2 - (ri * ri).sum()) <= 1e-12 * (ri * ri).sum()
    gbc3 = D().DeviceGrid(*cn, dtype=dtype)
    D().mg_residual_restrict(gu, gb, None, gbc3, h, 1.0, .64, .25)
    assert_bitexact(gbc3.to_host(), bc_want, "fused restrict (no r)")


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("quirk", [0, 1])
@pytest.mark.parametrize("cn", [(8, 6, 10), (3, 5, 128), (4, 3, 100), (1, 1, 1)])
def test_prolong(gpu, dtype, quirk, cn):
    bc = BC_CASES["mixed"]
    uc = filled(*cn, dtype, 61, bc, .1)
    uf = random_grid(*(2 * s for s in cn), dtype, 62)
    want = uf.copy()
    O.lib().orc_mg_prolong(uc.ptr(), want.ptr(), quirk)
    guc, guf = D().DeviceGrid.from_host(uc), D().DeviceGrid.from_host(uf)
    D().mg_prolong(guc, guf, quirk)
    assert_bitexact(guf.to_host(), want, f"prolong quirk={quirk}")


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("cn", [(4, 3, 128), (5, 4, 100), (3, 5, 70), (2, 3, 96)])
@pytest.mark.parametrize("case", ["neumann0", "neumann_v", "dirichlet"])
@pytest.mark.parametrize("first,flags", [(0, T.FUSED_WRITE_GHOSTS), (1, T.FUSED_WRITE_GHOSTS), (0, 0)])
def test_relax_fused_prolong_equals_prolong_then_sweep(gpu, dtype, cn, case, first, flags):
    """the prolongation-fused sweep against ocu_mg_prolong_bc followed by the
    fused sweep: u_out bit for bit over the whole buffer (random values in
    every ghost and padding cell of both inputs), u_in untouched"""
    bc = BC_CASES[case]
    hf = (.1, .2, .05)
    fine = tuple(2 * s for s in cn)
    uc = random_grid(*cn, dtype, 181)
    uf = random_grid(*fine, dtype, 182)
    b = random_grid(*fine, dtype, 183)
    out0 = random_grid(*fine, dtype, 184)
    p = O.level_params(bc, 0, *hf)
    g = {k: D().DeviceGrid.from_host(v) for k, v in (("uc", uc), ("uf", uf), ("b", b), ("o1", out0), ("o2", out0))}
    assert D().mg_relax_fused_prolong_ok(g["uf"], g["uc"], bc)
    D().mg_relax_fused_prolong_bc(g["uf"], g["o1"], g["b"], g["uc"], p, bc, *hf, first, flags)
    assert_bitexact(g["uf"].to_host(), uf, "u_in untouched")
    D().mg_prolong_bc(g["uc"], g["uf"], bc, *hf)
    D().mg_relax_fused_bc(g["uf"], g["o2"], g["b"], p, bc, *hf, first, flags)
    assert_bitexact(g["o1"].to_host(), g["o2"].to_host(), f"fused prolong {case}")


def test_relax_fused_prolong_rejects(gpu):
    from opencurrent_amd._lib import OcuError
    bc = BC_CASES["neumann0"]
    p = O.level_params(bc, 0, .1, .1, .1)
    for fz in (32, 128):  # one or two cells per lane: the separate prolongation runs instead
        uc, uf, b, o = (D().DeviceGrid(*s) for s in ((4, 4, fz // 2), (8, 8, fz), (8, 8, fz), (8, 8, fz)))
        assert not D().mg_relax_fused_prolong_ok(uf, uc, bc)
        with pytest.raises(OcuError):
            D().mg_relax_fused_prolong_bc(uf, o, b, uc, p, bc, .1, .1, .1)
    uc, uf, b, o = (D().DeviceGrid(*s) for s in ((4, 4, 128), (8, 8, 256), (8, 8, 256), (8, 8, 256)))
    assert D().mg_relax_fused_prolong_ok(uf, uc, bc)
    assert not D().mg_relax_fused_prolong_ok(uf, uc, BC_CASES["mixed"])  # periodic faces
    uc2 = D().DeviceGrid(4, 5, 128)
    assert not D().mg_relax_fused_prolong_ok(uf, uc2, bc)
    with pytest.raises(OcuError):
        D().mg_relax_fused_prolong_bc(uf, o, b, uc, p, bc, .1, .1, .1, 0, T.FUSED_ZERO_INPUT)
    with pytest.raises(OcuError):
        D().mg_relax_fused_prolong_bc(uf, uf, b, uc, p, bc, .1, .1, .1)


def coarse_colors(iters, order):
    it_n = max(4, iters + (iters % 2)) if order == 2 else iters
    for it in range(it_n):
        for rb in (0, 1):
            yield 1 - rb if order == 1 or (order == 2 and it >= it_n // 2) else rb


# (4, 6, 4): boxes in one workgroup's LDS; (30, 28, 24): boxes of 200 KB in
# FP32 and twice that in FP64, relaxed in global memory
COARSE_SIZES = [(4, 6, 4), (30, 28, 24)]


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("case", ["neumann0", "dirichlet", "mixed", "neumann_v"])
@pytest.mark.parametrize("order", [0, 1, 2])
@pytest.mark.parametrize("cn", COARSE_SIZES)
def test_coarse_relax_reference_semantics(gpu, dtype, case, order, cn):
    """relax_on_host (sol_mgpressure3ddev.cu:134-188): host arithmetic, host
    no-corner fill after each colour, level-1 fill at the end."""
    bc = BC_CASES[case]
    h = .5
    u = filled(*cn, dtype, 71, bc, h)
    b = random_grid(*cn, dtype, 72)
    p = O.level_params(bc, 0, h, h, h)
    iters = 8
    want = u.copy()
    it_n = iters
    if order == 2:
        it_n = max(4, iters + (iters % 2))
    for it in range(it_n):
        for rb in (0, 1):
            color = rb
            if order == 1 or (order == 2 and it >= it_n // 2):
                color = 1 - rb
            O.lib().orc_mg_relax(want.ptr(), b.ptr(), color, C.byref(p), T.RELAX_HOST)
            O.lib().orc_bc_fill(2, want.ptr(), C.byref(bc), h, h, h)
    O.lib().orc_bc_fill(0, want.ptr(), C.byref(bc), h, h, h)
    gu, gb = D().DeviceGrid.from_host(u), D().DeviceGrid.from_host(b)
    D().mg_coarse_relax(gu, gb, iters, order, p, bc, h, h, h, mode=0)
    assert_bitexact(gu.to_host(), want, f"coarse {case} order {order}")


@pytest.mark.parametrize("case", ["neumann0", "dirichlet", "mixed"])
@pytest.mark.parametrize("order", [0, 2])
@pytest.mark.parametrize("cn", COARSE_SIZES)
def test_coarse_relax_double_on_float(gpu, case, order, cn):
    """mode 1 (the mixed solver's coarsest level): FP32 grids relaxed in FP64
    with the host arithmetic and a level-1 fill after each colour, rounded
    back to FP32 once at the end - the oracle's FP64 relaxation of the
    widened grids, rounded."""
    bc = BC_CASES[case]
    h = .5
    u = filled(*cn, np.float32, 73, bc, h)
    b = random_grid(*cn, np.float32, 74)
    p = O.level_params(bc, 0, h, h, h)
    iters = 6
    ud, bd = HostGrid.congruent(u, dtype=np.float64), HostGrid.congruent(b, dtype=np.float64)
    ud.buf[:], bd.buf[:] = u.buf, b.buf
    for color in coarse_colors(iters, order):
        O.lib().orc_mg_relax(ud.ptr(), bd.ptr(), color, C.byref(p), T.RELAX_HOST)
        O.lib().orc_bc_fill(0, ud.ptr(), C.byref(bc), h, h, h)
    want = u.copy()
    want.buf[:] = ud.buf.astype(np.float32)
    gu, gb = D().DeviceGrid.from_host(u), D().DeviceGrid.from_host(b)
    D().mg_coarse_relax(gu, gb, iters, order, p, bc, h, h, h, mode=1)
    assert_bitexact(gu.to_host(), want, f"coarse mode 1 {case} order {order}")


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("size", [(5, 7, 3), (9, 4, 64), (3, 5, 100), (4, 3, 256), (3, 2, 260)])
def test_lincomb_interior(gpu, dtype, size):
    """the interior-only combination (padding over 3%) on rows of several
    lengths, bit for bit, also in place"""
    a, b, r = (random_grid(*size, dtype, 870 + q) for q in range(3))
    want = r.copy()
    O.lib().orc_lincomb2(want.ptr(), .75, a.ptr(), -1.25, b.ptr())
    gr, ga, gb = (D().DeviceGrid.from_host(x) for x in (r, a, b))
    D().grid_lincomb(gr, .75, ga, -1.25, gb)
    assert_bitexact(gr.to_host(), want, "lincomb interior")
    want2 = a.copy()
    O.lib().orc_lincomb2(want2.ptr(), 1.0, a.ptr(), .5, b.ptr())
    D().grid_lincomb(ga, 1.0, ga, .5, gb)
    assert_bitexact(ga.to_host(), want2, "lincomb interior in place")


@pytest.mark.parametrize("dtype", DTYPES)
def test_grid_ops(gpu, dtype):
    # 64^3 with padding (66,66,80): > 3% waste -> interior-only update
    a = random_grid(64, 64, 64, dtype, 81)
    b = random_grid(64, 64, 64, dtype, 82)
    r = random_grid(64, 64, 64, dtype, 83)
    # a grid whose waste is under 3% (256*256*272 allocated for 254*254*270)
    # -> whole-allocation update
    a2 = random_grid(254, 254, 270, dtype, 84)
    b2 = random_grid(254, 254, 270, dtype, 85)
    r2 = random_grid(254, 254, 270, dtype, 86)
    for (ra, aa, bb) in ((r, a, b), (r2, a2, b2)):
        want = ra.copy()
        O.lib().orc_lincomb2(want.ptr(), 1.0, aa.ptr(), .125, bb.ptr())
        gr, ga, gb = (D().DeviceGrid.from_host(x) for x in (ra, aa, bb))
        D().grid_lincomb(gr, 1.0, ga, .125, gb)
        assert_bitexact(gr.to_host(), want, "lincomb2")
        want1 = ra.copy()
        O.lib().orc_lincomb1(want1.ptr(), -.75, aa.ptr())
        D().grid_lincomb(gr, -.75, ga)
        assert_bitexact(gr.to_host(), want1, "lincomb1")
    ga = D().DeviceGrid.from_host(a)
    ai = a.interior().astype(np.float64)
    assert D().grid_reduce(ga, 0) == np.abs(ai).max()
    assert D().grid_reduce(ga, 1) == ai.max()
    assert D().grid_reduce(ga, 2) == ai.min()
    assert abs(D().grid_reduce(ga, 3) - ai.sum()) <= 1e-12 * np.abs(ai).sum()
    assert abs(D().grid_reduce(ga, 4) - (ai * ai).sum()) <= 1e-12 * (ai * ai).sum()
    # conversion copy
    other = np.float32 if dtype == np.float64 else np.float64
    dst = HostGrid.congruent(a, dtype=other)
    want = dst.copy()
    O.lib().orc_copy_convert(want.ptr(), a.ptr())
    gd = D().DeviceGrid.from_host(dst)
    D().grid_copy(gd, ga)
    assert_bitexact(gd.to_host(), want, "copy_convert")


NS_SIZES = [(16, 12, 20), (9, 7, 5)]


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("size", NS_SIZES)
def test_divergence_gradient(gpu, dtype, size):
    nx, ny, nz = size
    u, v, w, t = mac_grids(nx, ny, nz, dtype)
    for q, g in enumerate((u, v, w, t)):
        g.buf[:] = (2 * uniform(90 + q, g.buf.shape) - 1).astype(dtype)
    hx, hy, hz = .5, .25, .125
    want = t.copy()
    O.lib().orc_divergence(u.ptr(), v.ptr(), w.ptr(), want.ptr(), nx, ny, nz, hx, hy, hz)
    du, dv, dw, dt = (D().DeviceGrid.from_host(x) for x in (u, v, w, t))
    D().divergence(du, dv, dw, dt, nx, ny, nz, hx, hy, hz)
    assert_bitexact(dt.to_host(), want, "divergence")
    wu, wv, ww = u.copy(), v.copy(), w.copy()
    O.lib().orc_gradient_sub(wu.ptr(), wv.ptr(), ww.ptr(), t.ptr(), -1.0, hx, hy, hz)
    dp = D().DeviceGrid.from_host(t)
    D().gradient_sub(du, dv, dw, dp, -1.0, hx, hy, hz)
    for got, want_, n in ((du, wu, "u"), (dv, wv, "v"), (dw, ww, "w")):
        assert_bitexact(got.to_host(), want_, f"gradient {n}")


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("scheme", [T.IT_FIRST_ORDER_UPWIND, T.IT_SECOND_ORDER_CENTERED])
@pytest.mark.parametrize("size", NS_SIZES)
def test_advection(gpu, dtype, scheme, size):
    nx, ny, nz = size
    u, v, w, t = mac_grids(nx, ny, nz, dtype)
    for q, g in enumerate((u, v, w, t)):
        g.buf[:] = (2 * uniform(100 + q, g.buf.shape) - 1).astype(dtype)
    hx, hy, hz = .5, .25, .125
    d = [HostGrid.congruent(t) for _ in range(4)]
    O.lib().orc_self_advect(u.ptr(), v.ptr(), w.ptr(), d[0].ptr(), d[1].ptr(), d[2].ptr(), nx, ny, nz, hx, hy, hz,
                            scheme)
    O.lib().orc_passive_advect(u.ptr(), v.ptr(), w.ptr(), t.ptr(), d[3].ptr(), nx, ny, nz, hx, hy, hz, scheme)
    du, dv, dw, dt = (D().DeviceGrid.from_host(x) for x in (u, v, w, t))
    g = [D().DeviceGrid.congruent(dt) for _ in range(4)]
    D().self_advect(du, dv, dw, g[0], g[1], g[2], nx, ny, nz, hx, hy, hz, scheme)
    D().passive_advect(du, dv, dw, dt, g[3], nx, ny, nz, hx, hy, hz, scheme)
    for q, n in enumerate(("du", "dv", "dw", "dphi")):
        assert_bitexact(g[q].to_host(), d[q], f"advection {n}")


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("pmask", [0, 3, 7, 4])
@pytest.mark.parametrize("size", NS_SIZES + [(2, 3, 1), (1, 5, 2)])
def test_advection_second_order_upwind(gpu, dtype, pmask, size):
    """second-order upwind self- and passive advection (random values in
    every ghost and padding cell, so a wrong far-sample index shows) against
    the oracle, bit for bit"""
    nx, ny, nz = size
    u, v, w, t = mac_grids(nx, ny, nz, dtype)
    for q, g in enumerate((u, v, w, t)):
        g.buf[:] = (2 * uniform(140 + q, g.buf.shape) - 1).astype(dtype)
    hx, hy, hz = .5, .25, .125
    up2 = T.IT_SECOND_ORDER_UPWIND
    d = [HostGrid.congruent(g) for g in (u, v, w, t)]
    O.lib().orc_self_advect_ex(u.ptr(), v.ptr(), w.ptr(), d[0].ptr(), d[1].ptr(), d[2].ptr(), nx, ny, nz, hx, hy,
                               hz, up2, pmask)
    O.lib().orc_passive_advect_ex(u.ptr(), v.ptr(), w.ptr(), t.ptr(), d[3].ptr(), nx, ny, nz, hx, hy, hz, up2, pmask)
    du, dv, dw, dt = (D().DeviceGrid.from_host(x) for x in (u, v, w, t))
    g = [D().DeviceGrid.from_host(HostGrid.congruent(x)) for x in (u, v, w, t)]
    D().self_advect(du, dv, dw, g[0], g[1], g[2], nx, ny, nz, hx, hy, hz, up2, pmask)
    D().passive_advect(du, dv, dw, dt, g[3], nx, ny, nz, hx, hy, hz, up2, pmask)
    for q, n in enumerate(("du", "dv", "dw", "dphi")):
        assert_bitexact(g[q].to_host(), d[q], f"advection {n}")


def test_second_order_upwind_entry_checks(gpu):
    """the reference's entry points keep their two schemes; the extended
    ones take three and a 3-bit axes mask"""
    from opencurrent_amd._lib import OcuError
    u, v, w, t = (D().DeviceGrid.from_host(x) for x in mac_grids(6, 5, 4))
    dt = D().DeviceGrid.congruent(t)
    up2 = T.IT_SECOND_ORDER_UPWIND
    with pytest.raises(OcuError):
        D().passive_advect(u, v, w, t, dt, 6, 5, 4, 1, 1, 1, up2)
    with pytest.raises(OcuError):
        D().self_advect(u, v, w, dt, dt, dt, 6, 5, 4, 1, 1, 1, up2)
    with pytest.raises(OcuError):
        D().passive_advect(u, v, w, t, dt, 6, 5, 4, 1, 1, 1, up2, 8)
    with pytest.raises(OcuError):
        D().self_advect(u, v, w, dt, dt, dt, 6, 5, 4, 1, 1, 1, 3, 0)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("scheme", [T.IT_FIRST_ORDER_UPWIND, T.IT_SECOND_ORDER_CENTERED])
@pytest.mark.parametrize("size", NS_SIZES)
def test_advection_adjoint(gpu, dtype, scheme, size):
    """the adjoint kernels against the oracle's (pinned on the CPU by the
    dot-product identity, tests/test_oracle_adjoint.py), bit for bit"""
    nx, ny, nz = size
    u, v, w, t = mac_grids(nx, ny, nz, dtype)
    for q, g in enumerate((u, v, w, t)):
        g.buf[:] = (2 * uniform(120 + q, g.buf.shape) - 1).astype(dtype)
    lam = HostGrid.congruent(t)
    lam.interior()[:] = (2 * uniform(125, lam.interior().shape) - 1).astype(dtype)
    au, av, aw, _ = mac_grids(nx, ny, nz, dtype)
    for q, g in enumerate((au, av, aw)):
        g.buf[:] = (2 * uniform(126 + q, g.buf.shape) - 1).astype(dtype)
    out = HostGrid.congruent(t)
    hx, hy, hz = .5, .25, .125
    g_au, g_av, g_aw = (D().DeviceGrid.from_host(x) for x in (au, av, aw))
    O.lib().orc_passive_advect_adjoint(u.ptr(), v.ptr(), w.ptr(), t.ptr(), lam.ptr(), out.ptr(), au.ptr(), av.ptr(),
                                       aw.ptr(), .03125, nx, ny, nz, hx, hy, hz, scheme)
    du, dv, dw, dt, dl = (D().DeviceGrid.from_host(x) for x in (u, v, w, t, lam))
    dout = D().DeviceGrid.from_host(HostGrid.congruent(t))
    D().passive_advect_adjoint(du, dv, dw, dt, dl, dout, g_au, g_av, g_aw, .03125, nx, ny, nz, hx, hy, hz, scheme)
    for got, want, n in ((dout, out, "lam_out"), (g_au, au, "adj_u"), (g_av, av, "adj_v"), (g_aw, aw, "adj_w")):
        assert_bitexact(got.to_host(), want, f"advection adjoint {n}")


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("case", ["neumann_v", "dirichlet", "mixed", "periodic", "none_faces"])
def test_ghost_adjoints(gpu, dtype, case):
    bc = BC_CASES[case]
    g = random_grid(9, 7, 13, dtype, 130)
    want = g.copy()
    O.lib().orc_bc_adjoint_nocorners(want.ptr(), C.byref(bc))
    dg = D().DeviceGrid.from_host(g)
    D().bc_adjoint_nocorners(dg, bc)
    assert_bitexact(dg.to_host(), want, f"ghost adjoint {case}")
    mbc = mac_bc("inflow") if case == "mixed" else MAC_CASES["periodic"]
    u, v, w, _ = mac_grids(8, 6, 10, dtype)
    for q, x in enumerate((u, v, w)):
        x.buf[:] = (2 * uniform(131 + q, x.buf.shape) - 1).astype(dtype)
    wu, wv, ww = u.copy(), v.copy(), w.copy()
    O.lib().orc_mac_adjoint_normal(wu.ptr(), wv.ptr(), ww.ptr(), C.byref(mbc))
    du, dv, dw = (D().DeviceGrid.from_host(x) for x in (u, v, w))
    D().mac_adjoint_normal(du, dv, dw, mbc)
    for got, want_, n in ((du, wu, "u"), (dv, wv, "v"), (dw, ww, "w")):
        assert_bitexact(got.to_host(), want_, f"mac adjoint {n}")


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("n", [(8, 6, 10), (7, 9, 5), (1, 1, 1), (33, 4, 70)])
def test_divergence_gradient_adjoints(gpu, dtype, n):
    """ocu_divergence_adjoint / ocu_gradient_adjoint bit-exact against the
    oracle over whole buffers (every ghost, padding and face included)"""
    u, v, w, lam = mac_grids(*n, dtype)
    for q, x in enumerate((u, v, w, lam)):
        x.buf[:] = (2 * uniform(141 + q, x.buf.shape) - 1).astype(dtype)
    h = (.5, .25, 1 / 3)
    wu, wv, ww = u.copy(), v.copy(), w.copy()
    assert O.lib().orc_divergence_adjoint(lam.ptr(), wu.ptr(), wv.ptr(), ww.ptr(), *n, *h) == 1
    du, dv, dw, dl = (D().DeviceGrid.from_host(x) for x in (u, v, w, lam))
    D().divergence_adjoint(dl, du, dv, dw, *n, *h)
    for got, want, nm in ((du, wu, "u"), (dv, wv, "v"), (dw, ww, "w")):
        assert_bitexact(got.to_host(), want, f"divergence adjoint {nm} {n}")
    for coef in (-1.0, .7):
        wp = lam.copy()
        assert O.lib().orc_gradient_adjoint(u.ptr(), v.ptr(), w.ptr(), wp.ptr(), coef, *h) == 1
        du, dv, dw, dp = (D().DeviceGrid.from_host(x) for x in (u, v, w, lam))
        D().gradient_adjoint(du, dv, dw, dp, coef, *h)
        assert_bitexact(dp.to_host(), wp, f"gradient adjoint {coef} {n}")
        for got, x in ((du, u), (dv, v), (dw, w)):  # read only
            assert_bitexact(got.to_host(), x, "gradient adjoint input")


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("size", [(16, 12, 20), (5, 7, 200), (3, 4, 130), (8, 6, 256), (4, 9, 64), (2, 3, 1),
                                  (6, 5, 300)])
def test_laplacian_thermal(gpu, dtype, size):
    """the Laplacian (whole rows of 1, 2 or 4 cells per lane, or tiled past
    256) and the thermal force against the oracle, bit for bit"""
    nx, ny, nz = size
    u, v, w, t = mac_grids(nx, ny, nz, dtype)
    for q, g in enumerate((u, t)):
        g.buf[:] = (2 * uniform(110 + q, g.buf.shape) - 1).astype(dtype)
    d = HostGrid.congruent(t)
    d.buf[:] = (2 * uniform(115, d.buf.shape) - 1).astype(dtype)
    want = d.copy()
    O.lib().orc_laplacian_acc(u.ptr(), want.ptr(), nx + 1, ny, nz, .5, .25, .125, .71)
    gu, gd = D().DeviceGrid.from_host(u), D().DeviceGrid.from_host(d)
    D().laplacian_acc(gu, gd, nx + 1, ny, nz, .5, .25, .125, .71)
    assert_bitexact(gd.to_host(), want, "laplacian")
    for vd in (T.DIR_ZPOS, T.DIR_XNEG, T.DIR_YPOS):
        want2 = d.copy()
        O.lib().orc_thermal_force(want2.ptr(), t.ptr(), nx, ny, nz, vd, 7.1e4)
        gd2, gt = D().DeviceGrid.from_host(d), D().DeviceGrid.from_host(t)
        D().thermal_force(gd2, gt, nx, ny, nz, vd, 7.1e4)
        assert_bitexact(gd2.to_host(), want2, f"thermal {vd}")


@pytest.mark.parametrize("size", [(16, 12, 20), (8, 6, 256), (5, 7, 200), (3, 4, 130), (4, 9, 64), (2, 3, 1)])
def test_ir_residual(gpu, size):
    """ocu_ir_residual: r = (b - bshift) - A x in FP64 with the reference
    residual's operand order, r32 = (float)(r * scale) bit for bit (numpy
    performs the same IEEE double operations in the same order), and the
    (sum r^2, max |r|, sum r, sum x) partials reduced by ocu_norm4_finalize."""
    import torch
    from opencurrent_amd import _lib
    h, fx, fy, fz, bshift, scale = 1 / 32, 1.0, .64, .25, .0123, 2.0 ** 7
    x = random_grid(*size, np.float64, 151)
    b = random_grid(*size, np.float64, 152)
    r32 = HostGrid(*size, dtype=np.float32)
    X = x.full()
    n = size
    cx, cy, cz, cd = fx / (h * h), fy / (h * h), fz / (h * h), 2 * (fx + fy + fz) / (h * h)
    xc = X[1:n[0] + 1, 1:n[1] + 1, 1:n[2] + 1]
    zm, zp = X[1:n[0] + 1, 1:n[1] + 1, 0:n[2]], X[1:n[0] + 1, 1:n[1] + 1, 2:n[2] + 2]
    ym, yp = X[1:n[0] + 1, 0:n[1], 1:n[2] + 1], X[1:n[0] + 1, 2:n[1] + 2, 1:n[2] + 1]
    xm, xp = X[0:n[0], 1:n[1] + 1, 1:n[2] + 1], X[2:n[0] + 2, 1:n[1] + 1, 1:n[2] + 1]
    r = (b.interior() - bshift) - (cz * (zm + zp) + cy * (ym + yp) + cx * (xm + xp) - cd * xc)
    gx, gb, gr = (D().DeviceGrid.from_host(g) for g in (x, b, r32))
    hip = _lib.hip()
    ws = torch.zeros(4 * 2048 + 4, dtype=torch.float64, device="cuda")
    out = ws[4 * 2048:]
    _lib.check(hip.ocu_ir_residual(gx.ptr(), gb.ptr(), bshift, gr.ptr(), scale, h, fx, fy, fz,
                                   C.cast(ws.data_ptr(), _lib.DP), None), "ir_residual")
    _lib.check(hip.ocu_norm4_finalize(C.cast(ws.data_ptr(), _lib.DP), C.cast(out.data_ptr(), _lib.DP), None),
               "norm4_finalize")
    torch.cuda.synchronize()
    got = gr.to_host().interior()
    want = (r * scale).astype(np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    s2, mx, sm, sx = (float(v) for v in out.cpu())
    assert mx == np.abs(r).max()
    assert abs(s2 - (r * r).sum()) <= 1e-12 * (r * r).sum()
    assert abs(sm - r.sum()) <= 1e-12 * np.abs(r).sum()
    assert abs(sx - xc.sum()) <= 1e-12 * np.abs(xc).sum()


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("size", [(1, 1, 1), (3, 5, 7), (17, 9, 64), (4, 3, 255), (2, 6, 300), (2, 2, 513),
                                  (100, 100, 40), (90, 100, 257)])
def test_grid_reduce_max_ops(gpu, dtype, size):
    """maxabs, max and min (rows per wave, four loads per lane in flight;
    the last two sizes have more rows than the walk has waves) over
    row lengths below, at and past one pass of 256 cells, with the
    extremes planted in the ghost layer (never read) and, in a second grid,
    in the last cell of the last row"""
    g = random_grid(*size, dtype, 901)
    full = g.full()
    full[0], full[-1] = 50.0, -50.0
    full[:, 0], full[:, -1] = -60.0, 60.0
    full[:, :, 0], full[:, :, size[2] + 1:] = 70.0, -70.0  # k = -1, and k = nz with the padding
    ai = g.interior().astype(np.float64)
    gd = D().DeviceGrid.from_host(g)
    assert D().grid_reduce(gd, 0) == np.abs(ai).max()
    assert D().grid_reduce(gd, 1) == ai.max()
    assert D().grid_reduce(gd, 2) == ai.min()
    for v in (3.0, -3.0):
        g.interior()[-1, -1, -1] = v
        gd = D().DeviceGrid.from_host(g)
        assert D().grid_reduce(gd, 0) == 3.0
        assert D().grid_reduce(gd, 1 if v > 0 else 2) == v


@pytest.mark.parametrize("size", [(16, 12, 20), (8, 6, 256), (5, 7, 200), (3, 4, 130), (4, 9, 64), (2, 3, 1),
                                  (6, 5, 300)])
def test_grid_stats3(gpu, size):
    """(sum g^2, max|g|, sum g) of an FP64 interior, row and tiled forms"""
    import torch
    from opencurrent_amd import _lib
    g = random_grid(*size, np.float64, 161)
    gg = D().DeviceGrid.from_host(g)
    ws = torch.zeros(3 * 2048 + 3, dtype=torch.float64, device="cuda")
    out = ws[3 * 2048:]
    hip = _lib.hip()
    _lib.check(hip.ocu_grid_stats3(gg.ptr(), C.cast(ws.data_ptr(), _lib.DP), None), "stats3")
    _lib.check(hip.ocu_norm3_finalize(C.cast(ws.data_ptr(), _lib.DP), C.cast(out.data_ptr(), _lib.DP), None),
               "norm3_finalize")
    s2, mx, sm = (float(v) for v in out.cpu())
    x = g.interior()
    assert mx == np.abs(x).max()
    assert abs(s2 - (x * x).sum()) <= 1e-12 * (x * x).sum()
    assert abs(sm - x.sum()) <= 1e-12 * np.abs(x).sum()


@pytest.mark.parametrize("size", [(16, 12, 20), (8, 6, 256), (5, 7, 200), (2, 3, 1), (6, 5, 300)])
def test_ir_shift_and_correct(gpu, size):
    """ocu_ir_shift gives the interior mean of x + e*inv_scale from the sum of
    x and one pass over e; ocu_ir_correct_ex applies (x + e*inv) - shift
    bit for bit over the interior and, in face mode, the six face-ghost
    layers, leaving edges, corners and padding alone."""
    import torch
    from opencurrent_amd import _lib
    hip = _lib.hip()
    x = random_grid(*size, np.float64, 171)
    e = HostGrid(*size, dtype=np.float32)
    e.buf[:] = (2 * uniform(172, e.buf.shape) - 1).astype(np.float32)
    inv = 2.0 ** -9
    xi, ei = x.interior(), e.interior()
    ws = torch.zeros(2048 + 1, dtype=torch.float64, device="cuda")
    shift_dev = ws[2048:]
    ge = D().DeviceGrid.from_host(e)
    _lib.check(hip.ocu_ir_shift(ge.ptr(), inv, float(xi.sum()), C.cast(ws.data_ptr(), _lib.DP),
                                C.cast(shift_dev.data_ptr(), _lib.DP), None), "ir_shift")
    shift = float(shift_dev.cpu()[0])
    n = float(np.prod(size))
    want_shift = (xi.sum() + inv * ei.astype(np.float64).sum()) / n
    assert abs(shift - want_shift) <= 1e-14 * (np.abs(xi).sum() + inv * np.abs(ei).sum()) / n
    X, E = x.full(), e.full()
    for faces in (0, 1):
        gx = D().DeviceGrid.from_host(x)
        sp = C.cast(shift_dev.data_ptr(), _lib.DP) if faces else None
        _lib.check(hip.ocu_ir_correct_ex(gx.ptr(), ge.ptr(), inv, sp, faces, None), "ir_correct_ex")
        got = gx.to_host().full()
        want = X.copy()
        nx, ny, nz = size
        ii, jj, kk = np.meshgrid(np.arange(X.shape[0]) - 1, np.arange(X.shape[1]) - 1, np.arange(X.shape[2]) - 1,
                                 indexing="ij")
        ghost = ((ii < 0) | (ii >= nx)).astype(int) + ((jj < 0) | (jj >= ny)) + ((kk < 0) | (kk >= nz))
        inbox = (ii <= nx) & (jj <= ny) & (kk <= nz)
        sel = (ghost == 0) if not faces else ((ghost <= 1) & inbox)
        v = X + E.astype(np.float64) * inv
        if faces:
            v = v - shift
        want[sel] = v[sel]
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), faces


RELAX_BC_CASES = {
    "mixed": BC_CASES["mixed"],  # periodic x, Dirichlet y, Neumann z (values nonzero)
    "periodic": BC_CASES["periodic"],
    "neumann_v": BC_CASES["neumann_v"],
    "dirichlet": BC_CASES["dirichlet"],
    "rb": bcset(T.BC_PERIODIC, T.BC_PERIODIC, T.BC_PERIODIC, T.BC_PERIODIC, (T.BC_NEUMANN, 0.0), (T.BC_NEUMANN, 0.0)),
}


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("case", list(RELAX_BC_CASES))
@pytest.mark.parametrize("size", [(16, 12, 20), (8, 6, 10), (2, 4, 2), (6, 2, 34)])
def test_relax_bc_equals_colour_then_fill(gpu, dtype, case, size):
    """ocu_mg_relax_bc == one colour of ocu_mg_relax followed by the level-1
    fill, over the whole buffer, for both colours in turn"""
    bc = RELAX_BC_CASES[case]
    hx, hy, hz = .25, .125, .5
    u = random_grid(*size, dtype, 171)
    O.lib().orc_bc_fill(0, u.ptr(), C.byref(bc), hx, hy, hz)
    b = random_grid(*size, dtype, 172)
    p = O.level_params(bc, 1, hx, hy, hz)
    want = u.copy()
    gu, gb = D().DeviceGrid.from_host(u), D().DeviceGrid.from_host(b)
    for color in (1, 0):
        O.lib().orc_mg_relax(want.ptr(), b.ptr(), color, C.byref(p), T.RELAX_DEVICE)
        O.lib().orc_bc_fill(0, want.ptr(), C.byref(bc), hx, hy, hz)
        D().mg_relax_bc(gu, gb, color, p, bc, hx, hy, hz)
        assert_bitexact(gu.to_host(), want, f"relax_bc {case} {size} colour {color}")


def test_relax_bc_rejects_odd_periodic_and_none_faces(gpu):
    from opencurrent_amd._lib import OcuError
    p = O.level_params(BC_CASES["periodic"], 0, .1, .1, .1)
    for size, bc in (((7, 8, 8), BC_CASES["periodic"]), ((8, 8, 8), BC_CASES["none_faces"])):
        u, b = D().DeviceGrid(*size), D().DeviceGrid(*size)
        with pytest.raises(OcuError):
            D().mg_relax_bc(u, b, 0, p, bc, .1, .1, .1)
