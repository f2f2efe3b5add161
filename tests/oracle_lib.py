"""Loader of the CPU oracle (oracle/_build/libocu_oracle.so) for the tests.

The oracle is test infrastructure: it restates the reference algorithm on
the CPU and is only ever used to compare results against."""
import ctypes as C
import os
import subprocess
from pathlib import Path

from opencurrent_amd._types import BCSet, Grid3DDesc, MGConfig, MGLevelParams
from opencurrent_amd._lib import NSParams

ROOT = Path(__file__).resolve().parent.parent
LIB = ROOT / "oracle" / "_build" / "libocu_oracle.so"

GP = C.POINTER(Grid3DDesc)
BP = C.POINTER(BCSet)
D = C.c_double
I = C.c_int
DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)

SIGS = {
    "orc_layout": (I, [I, I, I, I, I, I, I, I, I, I, GP]),
    "orc_set_threads": (I, [I]),
    "orc_bc_fill": (I, [I, GP, BP, D, D, D]),
    "orc_bc_fill_mac": (I, [GP, GP, GP, BP]),
    "orc_lincomb1": (I, [GP, D, GP]),
    "orc_lincomb2": (I, [GP, D, GP, D, GP]),
    "orc_copy_convert": (I, [GP, GP]),
    "orc_reduce": (D, [GP, I]),
    "orc_mg_num_levels": (I, [I, I, I]),
    "orc_mg_optimal_omega": (D, [D, D, D]),
    "orc_mg_level_params": (I, [BP, I, D, D, D, C.POINTER(MGLevelParams)]),
    "orc_mg_relax": (I, [GP, GP, I, C.POINTER(MGLevelParams), I]),
    "orc_mg_residual": (I, [GP, GP, GP, D, D, D, D]),
    "orc_mg_restrict": (I, [GP, GP]),
    "orc_mg_prolong": (I, [GP, GP, I]),
    "orc_mg_solve": (I, [GP, GP, D, D, D, C.POINTER(MGConfig), D, I, I, DP, IP, DP, I, IP]),
    "orc_ir_solve": (I, [GP, GP, D, D, D, C.POINTER(MGConfig), D, I, I, I, I, I, DP, IP, IP]),
    "orc_divergence": (I, [GP, GP, GP, GP, I, I, I, D, D, D]),
    "orc_gradient_sub": (I, [GP, GP, GP, GP, D, D, D, D]),
    "orc_self_advect": (I, [GP, GP, GP, GP, GP, GP, I, I, I, D, D, D, I]),
    "orc_passive_advect": (I, [GP, GP, GP, GP, GP, I, I, I, D, D, D, I]),
    "orc_self_advect_ex": (I, [GP, GP, GP, GP, GP, GP, I, I, I, D, D, D, I, I]),
    "orc_passive_advect_ex": (I, [GP, GP, GP, GP, GP, I, I, I, D, D, D, I, I]),
    "orc_passive_advect_adjoint": (I, [GP, GP, GP, GP, GP, GP, GP, GP, GP, D, I, I, I, D, D, D, I]),
    "orc_bc_adjoint_nocorners": (I, [GP, BP]),
    "orc_mac_adjoint_normal": (I, [GP, GP, GP, BP]),
    "orc_divergence_adjoint": (I, [GP, GP, GP, GP, I, I, I, D, D, D]),
    "orc_gradient_adjoint": (I, [GP, GP, GP, GP, D, D, D, D]),
    "orc_laplacian_acc": (I, [GP, GP, I, I, I, D, D, D, D]),
    "orc_thermal_force": (I, [GP, GP, I, I, I, I, D]),
    "orc_project": (I, [GP, GP, GP, BP, I, I, I, D, D, D, D, GP, DP]),
    "orc_ns_create": (C.c_void_p, [I, C.POINTER(NSParams), GP, GP, GP, GP]),
    "orc_ns_advance_one_step": (I, [C.c_void_p, D]),
    "orc_ns_advance": (I, [C.c_void_p, D]),
    "orc_ns_max_stable_timestep": (D, [C.c_void_p]),
    "orc_ns_get": (I, [C.c_void_p, I, GP]),
    "orc_ns_destroy": (None, [C.c_void_p]),
    "orc_pcg_apply_laplacian": (I, [GP, GP, GP, D, D, D]),
    "orc_pcg_diag_precond": (I, [GP, GP, GP, D, D, D]),
    "orc_pcg_dot": (D, [GP, GP]),
    "orc_pcg_solve": (I, [GP, GP, GP, C.c_float, C.c_float, C.c_float, BP, I, I, I, I, D, I, DP, IP]),
    "orc_scalar_advect_step": (I, [GP, GP, GP, GP, GP, BP, I, I, I, D, D, D, I, D]),
    "orc_bc_fill_level2_nocorners": (I, [GP, BP, D, D, D]),
    "orc_bc_fill_mac_level2": (I, [GP, GP, GP, BP]),
    "orc_sample_points": (I, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, I, GP, BP] + [C.c_float] * 9),
    "orc_sample_points_mac": (I, [C.c_void_p] * 6 + [I, GP, GP, GP, BP] + [C.c_float] * 3),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            subprocess.run(["make", "-C", str(ROOT / "oracle")], check=True, capture_output=True)
        L = C.CDLL(str(LIB))
        for n, (r, a) in SIGS.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def mg_solve(u, b, h, cfg, tol, max_iter, mode=0):
    """Run the oracle MG solver in place on HostGrid u; returns (ok, residual, vcycles, history)."""
    hx, hy, hz = h
    res, vc, n = C.c_double(), C.c_int(), C.c_int()
    hist = (C.c_double * 1024)()
    ok = lib().orc_mg_solve(u.ptr(), b.ptr(), hx, hy, hz, C.byref(cfg), tol, max_iter, mode, C.byref(res),
                            C.byref(vc), hist, 1024, C.byref(n))
    return ok, res.value, vc.value, [hist[i] for i in range(min(n.value, 1024))]


def ir_solve(x, b, h, cfg, tol, max_ir, cycles=4, cycles_first=4, fmg_first=False, adaptive=True):
    hx, hy, hz = h
    res, it, vc = C.c_double(), C.c_int(), C.c_int()
    ok = lib().orc_ir_solve(x.ptr(), b.ptr(), hx, hy, hz, C.byref(cfg), tol, max_ir, cycles, cycles_first,
                            int(fmg_first), int(adaptive), C.byref(res), C.byref(it), C.byref(vc))
    return ok, res.value, it.value, vc.value


def level_params(bc, level, hx, hy, hz):
    p = MGLevelParams()
    lib().orc_mg_level_params(C.byref(bc), level, hx, hy, hz, C.byref(p))
    return p


def pcg_solve(pressure, rhs, coeff, h, bc, preconditioner, cycles=2, use_fmg=True, convergence=5, tol=1e-6,
              max_iter=1000):
    """The oracle's Sol_PCGPressure3DDevice solve from the guess in HostGrid
    `pressure` (overwritten); returns (ok, residual, iterations)."""
    res, it = C.c_double(), C.c_int()
    ok = lib().orc_pcg_solve(rhs.ptr(), coeff.ptr(), pressure.ptr(), h[0], h[1], h[2], C.byref(bc), preconditioner,
                             cycles, int(use_fmg), convergence, tol, max_iter, C.byref(res), C.byref(it))
    return ok, res.value, it.value
