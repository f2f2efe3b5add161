"""Known-answer checks of the oracle's whole-grid helpers (CPU).

The GPU copy/convert and linear-combination kernels are compared with these
oracle functions bit for bit, so the oracle itself is pinned here against
numpy, independently of its own code:

* ``copy_all_data<T,S>`` (ocustorage/grid3ddev.cu:25-33, :216-233) converts
  every allocated element, ghosts and padding included: numpy ``astype``.
* ``linear_combination(a1, g1)`` (grid3ddev.cu:283-300, kernel :57-65) writes
  ``a1*g1`` over the whole allocation.
* ``linear_combination(a1, g1, a2, g2)`` (grid3ddev.cu:303-349) writes
  ``a1*g1 + a2*g2`` over the whole allocation when at most 3% of it lies
  outside the interior, and over the interior only (everything else left as
  it was) when more does.  Layouts on both sides of the 3% line are used, so
  a reversed or missing branch fails.
"""
import numpy as np
import pytest

from opencurrent_amd._types import HostGrid, layout
import fields
import oracle_lib as O


def _filled(seed, dims, dtype, pads=(0, 0, 0)):
    g = HostGrid(*dims, 1, 1, 1, *pads, dtype=dtype)
    g.buf[:] = (4 * fields.uniform(seed, g.buf.shape) - 2).astype(dtype)
    return g


def _waste(d):
    return (d.allocated - d.nx * d.ny * d.nz) / d.allocated


@pytest.mark.parametrize("src_t,dst_t", [(np.float64, np.float32), (np.float32, np.float64),
                                         (np.float32, np.float32), (np.float64, np.float64)])
def test_copy_convert_is_astype_over_the_whole_allocation(src_t, dst_t):
    src = _filled(11, (9, 7, 21), src_t, pads=(1, 0, 3))
    dst = HostGrid.congruent(src, dtype=dst_t)
    dst.buf[:] = np.nan
    assert O.lib().orc_copy_convert(dst.ptr(), src.ptr()) == 1
    want = src.buf.astype(dst_t)
    assert np.array_equal(dst.buf.view(np.uint8), want.view(np.uint8))
    # FP64 -> FP32 rounds to nearest even: a value halfway between two floats
    if src_t == np.float64 and dst_t == np.float32:
        halfway = float(np.float32(1.0)) + 2.0 ** -24  # tie between 1 and 1+2^-23
        src.buf[0] = halfway
        O.lib().orc_copy_convert(dst.ptr(), src.ptr())
        assert dst.buf[0] == np.float32(1.0)


def test_copy_convert_layout_check_is_on_the_padded_box():
    # check_layout_match (grid3d.h:84) compares pnx, pny, pnz only
    a = HostGrid(8, 8, 8)
    assert O.lib().orc_copy_convert(HostGrid(8, 8, 9).ptr(), a.ptr()) == 1  # pnz 16 both
    assert O.lib().orc_copy_convert(HostGrid(9, 8, 8).ptr(), a.ptr()) == 0


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_lincomb1_scales_every_allocated_element(dtype):
    g = _filled(12, (16, 16, 16), dtype)
    r = HostGrid.congruent(g)
    r.buf[:] = np.nan
    a1 = -0.3
    assert O.lib().orc_lincomb1(r.ptr(), a1, g.ptr()) == 1
    want = g.buf * dtype(a1)  # the coefficient is a T, the product is rounded in T
    assert want.dtype == dtype
    assert np.array_equal(r.buf, want)


def test_three_percent_rule_layouts():
    # 64^3 wastes a quarter of its allocation; 254x254x270 about 2.3%;
    # 172^2x254 (3.05%) and 176^2x254 (2.9985%) straddle the line.
    assert _waste(layout(64, 64, 64)) > 0.2
    assert _waste(layout(254, 254, 270)) < 0.03
    assert 0.0305 > _waste(layout(172, 172, 254)) > 0.03 > _waste(layout(176, 176, 254)) > 0.0299


@pytest.mark.parametrize("dims,whole,dtype", [
    ((64, 64, 64), False, np.float64), ((64, 64, 64), False, np.float32),
    ((172, 172, 254), False, np.float32), ((176, 176, 254), True, np.float32),
    ((254, 254, 270), True, np.float32), ((254, 254, 270), True, np.float64)])
def test_lincomb2_interior_or_whole_allocation(dims, whole, dtype):
    g1 = _filled(21, dims, dtype)
    g2 = _filled(22, dims, dtype)
    r = _filled(23, dims, dtype)
    before = r.buf.copy()
    a1, a2 = 0.7, -1.3
    assert O.lib().orc_lincomb2(r.ptr(), a1, g1.ptr(), a2, g2.ptr()) == 1
    comb = g1.buf * dtype(a1) + g2.buf * dtype(a2)  # two rounded products, one rounded sum
    assert comb.dtype == dtype
    if whole:
        assert np.array_equal(r.buf, comb)
    else:
        mask = np.zeros(r.buf.shape, bool)
        d = r.desc
        full = mask[16 - d.gz:].reshape(d.pnx, d.pny, d.pnz)
        full[d.gx:d.gx + d.nx, d.gy:d.gy + d.ny, d.gz:d.gz + d.nz] = True
        assert np.array_equal(r.buf[mask], comb[mask])
        assert np.array_equal(r.buf[~mask], before[~mask])


def test_lincomb2_in_place_accumulation():
    # r = 1*r + dt*d, the form the time integrators use (eqn_incompressns3d.cu:355-390)
    x = _filled(31, (12, 10, 14), np.float64)
    d = _filled(32, (12, 10, 14), np.float64)
    want = x.interior() * 1.0 + d.interior() * 0.05
    assert O.lib().orc_lincomb2(x.ptr(), 1.0, x.ptr(), 0.05, d.ptr()) == 1
    assert np.array_equal(x.interior(), want)
