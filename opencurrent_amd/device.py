"""Device grids (torch tensors with the Grid3D layout) and thin wrappers of the
kernel entry points of libocu_hip.so.

torch is used only to own device memory and to order work on its stream;
every operation below runs a hand-written HIP kernel through the C ABI.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._types import (OCU_F32, OCU_F64, BCSet, Grid3DDesc, HostGrid, MGLevelParams, dtype_code,
                     layout, np_dtype)

_torch = None


def torch():
    global _torch
    if _torch is None:
        import torch as t
        _torch = t
    return _torch


def require_gpu():
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError("opencurrent_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    _lib.hip()
    _lib.host()


def _tdtype(code):
    t = torch()
    return {OCU_F32: t.float32, OCU_F64: t.float64}[code]


class DeviceGrid:
    """A Grid3DDevice<T>-shaped allocation held by a torch tensor."""

    def __init__(self, nx, ny, nz, gx=1, gy=1, gz=1, padx=0, pady=0, padz=0, dtype=np.float64):
        code = dtype_code(dtype)
        self.desc = layout(nx, ny, nz, gx, gy, gz, padx, pady, padz, code)
        self.tensor = torch().zeros(int(self.desc.allocated), dtype=_tdtype(code), device="cuda")
        self.desc.base = self.tensor.data_ptr()

    @classmethod
    def congruent(cls, other, dtype=None):
        d = other.desc
        return cls(d.nx, d.ny, d.nz, d.gx, d.gy, d.gz, d.pnx - (d.nx + 2 * d.gx), d.pny - (d.ny + 2 * d.gy),
                   d.pnz - (d.nz + 2 * d.gz), dtype=dtype or np_dtype(d.dtype))

    @classmethod
    def from_host(cls, h):
        d = h.desc
        g = cls(d.nx, d.ny, d.nz, d.gx, d.gy, d.gz, d.pnx - (d.nx + 2 * d.gx), d.pny - (d.ny + 2 * d.gy),
                d.pnz - (d.nz + 2 * d.gz), dtype=h.buf.dtype)
        g.tensor.copy_(torch().from_numpy(h.buf))
        return g

    def to_host(self):
        torch().cuda.synchronize()
        d = self.desc
        h = HostGrid(d.nx, d.ny, d.nz, d.gx, d.gy, d.gz, d.pnx - (d.nx + 2 * d.gx), d.pny - (d.ny + 2 * d.gy),
                     d.pnz - (d.nz + 2 * d.gz), dtype=np_dtype(d.dtype))
        h.buf[:] = self.tensor.cpu().numpy()
        return h

    def full(self):
        d = self.desc
        pre = 16 - d.gz
        return self.tensor[pre:].view(d.pnx, d.pny, d.pnz)

    def interior(self):
        d = self.desc
        return self.full()[d.gx:d.gx + d.nx, d.gy:d.gy + d.ny, d.gz:d.gz + d.nz]

    def ptr(self):
        return C.byref(self.desc)


def desc_to_host(desc):
    """Copy a device grid described by `desc` (e.g. solver-owned) into a HostGrid."""
    d = desc
    h = HostGrid(d.nx, d.ny, d.nz, d.gx, d.gy, d.gz, d.pnx - (d.nx + 2 * d.gx), d.pny - (d.ny + 2 * d.gy),
                 d.pnz - (d.nz + 2 * d.gz), dtype=np_dtype(d.dtype))
    hip = _lib.hip()
    stream = _lib.host().ocu_get_compute_stream()
    _lib.check(hip.ocu_memcpy_async(h.buf.ctypes.data, d.base, h.buf.nbytes, 1, stream), "memcpy D2H")
    _lib.check(hip.ocu_stream_synchronize(stream), "stream sync")
    return h


def host_to_desc(h, desc):
    """Copy a HostGrid into the device grid described by `desc`."""
    hip = _lib.hip()
    stream = _lib.host().ocu_get_compute_stream()
    _lib.check(hip.ocu_memcpy_async(desc.base, h.buf.ctypes.data, h.buf.nbytes, 0, stream), "memcpy H2D")
    _lib.check(hip.ocu_stream_synchronize(stream), "stream sync")


def sync():
    torch().cuda.synchronize()


# ---------------------------------------------------------------- kernels
# All take DeviceGrid arguments and enqueue on the legacy default stream
# (stream=None), which torch also uses.

def bc_fill_level1(g, bc, hx, hy, hz):
    _lib.check(_lib.hip().ocu_bc_fill_level1(g.ptr(), C.byref(bc), hx, hy, hz, None), "bc_fill_level1")


def bc_fill_nocorners(g, bc, hx, hy, hz):
    _lib.check(_lib.hip().ocu_bc_fill_nocorners(g.ptr(), C.byref(bc), hx, hy, hz, None), "bc_fill_nocorners")


def bc_fill_mac(u, v, w, bc):
    _lib.check(_lib.hip().ocu_bc_fill_mac_level1(u.ptr(), v.ptr(), w.ptr(), C.byref(bc), None), "bc_fill_mac")


def mg_relax(u, b, color, params, variant):
    _lib.check(_lib.hip().ocu_mg_relax(u.ptr(), b.ptr(), color, C.byref(params), variant, None), "mg_relax")


def mg_relax_bc(u, b, color, params, bc, hx, hy, hz):
    _lib.check(_lib.hip().ocu_mg_relax_bc(u.ptr(), b.ptr(), color, C.byref(params), C.byref(bc), hx, hy, hz, None),
               "mg_relax_bc")


def mg_relax_fused(u_in, u_out, b, params, bc, first_color=0):
    return _lib.check(_lib.hip().ocu_mg_relax_fused(u_in.ptr(), u_out.ptr(), b.ptr(), C.byref(params), C.byref(bc),
                                                    first_color, None), "mg_relax_fused")


def mg_relax_fused_bc(u_in, u_out, b, params, bc, hx, hy, hz, first_color=0, flags=0):
    return _lib.check(_lib.hip().ocu_mg_relax_fused_bc(u_in.ptr(), u_out.ptr(), b.ptr(), C.byref(params),
                                                       C.byref(bc), hx, hy, hz, first_color, flags, None),
                      "mg_relax_fused_bc")


def mg_relax_fused_prolong_bc(u_in, u_out, b, uc, params, bc, hx, hy, hz, first_color=0, flags=0, quirk=1):
    return _lib.check(_lib.hip().ocu_mg_relax_fused_prolong_bc(u_in.ptr(), u_out.ptr(), b.ptr(), uc.ptr(), quirk,
                                                               C.byref(params), C.byref(bc), hx, hy, hz,
                                                               first_color, flags, None),
                      "mg_relax_fused_prolong_bc")


def mg_prolong_bc_ok(uc, uf, bc):
    return bool(_lib.hip().ocu_mg_prolong_bc_ok(uc.ptr(), uf.ptr(), C.byref(bc)))


def mg_relax_fused_ok(u_in, u_out, b, bc):
    return bool(_lib.hip().ocu_mg_relax_fused_ok(u_in.ptr(), u_out.ptr(), b.ptr(), C.byref(bc)))


def mg_relax_fused_prolong_ok(u, uc, bc):
    return bool(_lib.hip().ocu_mg_relax_fused_prolong_ok(u.ptr(), uc.ptr(), C.byref(bc)))


def mg_residual(u, b, r, h, fx, fy, fz):
    _lib.check(_lib.hip().ocu_mg_residual(u.ptr(), b.ptr(), r.ptr(), h, fx, fy, fz, None), "mg_residual")


def mg_restrict(r, bc):
    _lib.check(_lib.hip().ocu_mg_restrict(r.ptr(), bc.ptr(), None), "mg_restrict")


def mg_residual_restrict(u, b, r, bcoarse, h, fx, fy, fz, with_norm=False):
    """Fused residual+restriction; returns (sum r^2, max|r|) when with_norm."""
    hip = _lib.hip()
    t = torch()
    ws = t.zeros(2 * hip.ocu_norm_parts() + 2, dtype=t.float64, device="cuda") if with_norm else None
    wsp = C.cast(ws.data_ptr(), _lib.DP) if with_norm else None
    _lib.check(hip.ocu_mg_residual_restrict(u.ptr(), b.ptr(), r.ptr() if r is not None else None, bcoarse.ptr(),
                                            h, fx, fy, fz, wsp, None), "mg_residual_restrict")
    if not with_norm:
        return None
    out = C.cast(ws.data_ptr() + 8 * 2 * hip.ocu_norm_parts(), _lib.DP)
    _lib.check(hip.ocu_norm_finalize(wsp, out, None), "norm_finalize")
    vals = ws[-2:].cpu().numpy()
    return float(vals[0]), float(vals[1])


def mg_prolong(uc, uf, quirk=1):
    _lib.check(_lib.hip().ocu_mg_prolong(uc.ptr(), uf.ptr(), quirk, None), "mg_prolong")


def mg_prolong_bc(uc, uf, bc, hx, hy, hz, quirk=1):
    _lib.check(_lib.hip().ocu_mg_prolong_bc(uc.ptr(), uf.ptr(), quirk, C.byref(bc), hx, hy, hz, None),
               "mg_prolong_bc")


def mg_coarse_relax(u, b, iterations, order, params, bc, hx, hy, hz, mode=0):
    _lib.check(_lib.hip().ocu_mg_coarse_relax(u.ptr(), b.ptr(), iterations, order, C.byref(params), C.byref(bc),
                                              hx, hy, hz, mode, None), "mg_coarse_relax")


def grid_reduce(g, op):
    hip = _lib.hip()
    t = torch()
    ws = t.zeros(hip.ocu_reduce_workspace_bytes() // 8 + 1, dtype=t.float64, device="cuda")
    out = C.cast(ws.data_ptr() + 8 * (ws.numel() - 1), _lib.DP)
    _lib.check(hip.ocu_grid_reduce(g.ptr(), op, ws.data_ptr(), out, None), "grid_reduce")
    return float(ws[-1].item())


def grid_lincomb(r, a1, g1, a2=0.0, g2=None):
    _lib.check(_lib.hip().ocu_grid_lincomb(r.ptr(), a1, g1.ptr(), a2, g2.ptr() if g2 is not None else None, None),
               "grid_lincomb")


def grid_copy(dst, src):
    _lib.check(_lib.hip().ocu_grid_copy(dst.ptr(), src.ptr(), None), "grid_copy")


def grid_clear_ghosts(g):
    _lib.check(_lib.hip().ocu_grid_clear_ghosts(g.ptr(), None), "grid_clear_ghosts")


def grid_clear(g, value=0.0):
    _lib.check(_lib.hip().ocu_grid_clear(g.ptr(), value, None), "grid_clear")


def divergence(u, v, w, div, nx, ny, nz, hx, hy, hz):
    _lib.check(_lib.hip().ocu_divergence(u.ptr(), v.ptr(), w.ptr(), div.ptr(), nx, ny, nz, hx, hy, hz, None),
               "divergence")


def gradient_sub(u, v, w, p, coef, hx, hy, hz):
    _lib.check(_lib.hip().ocu_gradient_sub(u.ptr(), v.ptr(), w.ptr(), p.ptr(), coef, hx, hy, hz, None),
               "gradient_sub")


def self_advect(u, v, w, du, dv, dw, nx, ny, nz, hx, hy, hz, scheme, periodic_axes=None):
    """periodic_axes None: the reference's entry point (schemes 1 and 2);
    a bitmask (x=1, y=2, z=4): the extended one, which also takes
    IT_SECOND_ORDER_UPWIND."""
    if periodic_axes is None:
        _lib.check(_lib.hip().ocu_self_advect(u.ptr(), v.ptr(), w.ptr(), du.ptr(), dv.ptr(), dw.ptr(), nx, ny, nz,
                                              hx, hy, hz, scheme, None), "self_advect")
    else:
        _lib.check(_lib.hip().ocu_self_advect_ex(u.ptr(), v.ptr(), w.ptr(), du.ptr(), dv.ptr(), dw.ptr(), nx, ny,
                                                 nz, hx, hy, hz, scheme, periodic_axes, None), "self_advect")


def passive_advect(u, v, w, phi, dphi, nx, ny, nz, hx, hy, hz, scheme, periodic_axes=None):
    if periodic_axes is None:
        _lib.check(_lib.hip().ocu_passive_advect(u.ptr(), v.ptr(), w.ptr(), phi.ptr(), dphi.ptr(), nx, ny, nz,
                                                 hx, hy, hz, scheme, None), "passive_advect")
    else:
        _lib.check(_lib.hip().ocu_passive_advect_ex(u.ptr(), v.ptr(), w.ptr(), phi.ptr(), dphi.ptr(), nx, ny, nz,
                                                    hx, hy, hz, scheme, periodic_axes, None), "passive_advect")


def passive_advect_adjoint(u, v, w, phi, lam, lam_out, au, av, aw, scale, nx, ny, nz, hx, hy, hz, scheme):
    _lib.check(_lib.hip().ocu_passive_advect_adjoint(u.ptr(), v.ptr(), w.ptr(), phi.ptr(), lam.ptr(), lam_out.ptr(),
                                                     au.ptr(), av.ptr(), aw.ptr(), scale, nx, ny, nz, hx, hy, hz,
                                                     scheme, None), "passive_advect_adjoint")


def bc_adjoint_nocorners(g, bc):
    _lib.check(_lib.hip().ocu_bc_adjoint_nocorners(g.ptr(), C.byref(bc), None), "bc_adjoint_nocorners")


def mac_adjoint_normal(u, v, w, bc):
    _lib.check(_lib.hip().ocu_mac_adjoint_normal(u.ptr(), v.ptr(), w.ptr(), C.byref(bc), None), "mac_adjoint_normal")


def divergence_adjoint(lam, au, av, aw, nx, ny, nz, hx, hy, hz):
    _lib.check(_lib.hip().ocu_divergence_adjoint(lam.ptr(), au.ptr(), av.ptr(), aw.ptr(), nx, ny, nz, hx, hy, hz,
                                                 None), "divergence_adjoint")


def gradient_adjoint(au, av, aw, ap, coef, hx, hy, hz):
    _lib.check(_lib.hip().ocu_gradient_adjoint(au.ptr(), av.ptr(), aw.ptr(), ap.ptr(), coef, hx, hy, hz, None),
               "gradient_adjoint")


def laplacian_acc(phi, dphi, nx, ny, nz, hx, hy, hz, coef):
    _lib.check(_lib.hip().ocu_laplacian_acc(phi.ptr(), dphi.ptr(), nx, ny, nz, hx, hy, hz, coef, None),
               "laplacian_acc")


def thermal_force(d, t, nx, ny, nz, vertical, coef):
    _lib.check(_lib.hip().ocu_thermal_force(d.ptr(), t.ptr(), nx, ny, nz, vertical, coef, None), "thermal_force")
