"""ctypes bindings of the two native libraries.

``hip()``  -> libocu_hip.so (include/ocu_hip.h): the gfx950 kernels.
``host()`` -> libocu.so (include/ocu.h): the drop-in solver classes.

Both are built in-tree under ``opencurrent_amd/lib`` by ``opencurrent_amd.build``.
Loading fails loudly (``NativeLibraryMissing``) when a library is absent:
there is no fallback implementation.
"""
import ctypes as C
import os
from pathlib import Path

from ._types import BCSet, Grid3DDesc, MGConfig, MGLevelParams, NSStepParams

LIBDIR = Path(__file__).resolve().parent / "lib"

_hip = None
_host = None

P = C.c_void_p
I = C.c_int
D = C.c_double
GP = C.POINTER(Grid3DDesc)
BP = C.POINTER(BCSet)
MP = C.POINTER(MGLevelParams)
CP = C.POINTER(MGConfig)
DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)


class NativeLibraryMissing(RuntimeError):
    pass


class OcuError(RuntimeError):
    pass


# name: (restype, argtypes)
HIP_SIGS = {
    "ocu_last_error": (C.c_char_p, []),
    "ocu_get_device_count": (I, [IP]),
    "ocu_set_device": (I, [I]),
    "ocu_device_synchronize": (I, []),
    "ocu_malloc": (I, [C.POINTER(P), C.c_size_t]),
    "ocu_free": (I, [P]),
    "ocu_host_malloc": (I, [C.POINTER(P), C.c_size_t]),
    "ocu_host_free": (I, [P]),
    "ocu_host_device_pointer": (I, [C.POINTER(P), P]),
    "ocu_memcpy_async": (I, [P, P, C.c_size_t, I, P]),
    "ocu_stream_create": (I, [C.POINTER(P)]),
    "ocu_stream_destroy": (I, [P]),
    "ocu_stream_synchronize": (I, [P]),
    "ocu_event_create": (I, [C.POINTER(P)]),
    "ocu_event_destroy": (I, [P]),
    "ocu_event_record": (I, [P, P]),
    "ocu_event_synchronize": (I, [P]),
    "ocu_event_elapsed_ms": (I, [C.POINTER(C.c_float), P, P]),
    "ocu_stream_begin_capture": (I, [P]),
    "ocu_stream_end_capture": (I, [P, C.POINTER(P)]),
    "ocu_graph_launch": (I, [P, P]),
    "ocu_graph_destroy": (I, [P]),
    "ocu_grid_clear": (I, [GP, D, P]),
    "ocu_grid_copy": (I, [GP, GP, P]),
    "ocu_grid_lincomb": (I, [GP, D, GP, D, GP, P]),
    "ocu_grid_add_scalar": (I, [GP, D, P]),
    "ocu_reduce_workspace_bytes": (C.c_size_t, []),
    "ocu_grid_reduce": (I, [GP, I, P, DP, P]),
    "ocu_grid_reduce_tagged": (I, [GP, I, P, DP, C.c_ulonglong, P]),
    "ocu_grid_maxabs3_tagged": (I, [GP, GP, GP, P, DP, C.c_ulonglong, P]),
    "ocu_grid_reduce_noalloc": (I, [GP, I, DP, P]),
    "ocu_bc_fill_level1": (I, [GP, BP, D, D, D, P]),
    "ocu_bc_fill_nocorners": (I, [GP, BP, D, D, D, P]),
    "ocu_bc_fill_mac_level1": (I, [GP, GP, GP, BP, P]),
    "ocu_bc_fill_level2_nocorners": (I, [GP, BP, D, D, D, P]),
    "ocu_bc_fill_mac_level2": (I, [GP, GP, GP, BP, P]),
    "ocu_sample_points": (I, [P, P, P, P, I, GP, BP] + [C.c_float] * 9 + [P]),
    "ocu_sample_points_mac": (I, [P] * 6 + [I, GP, GP, GP, BP] + [C.c_float] * 3 + [P]),
    "ocu_mg_relax": (I, [GP, GP, I, MP, I, P]),
    "ocu_mg_relax_fused": (I, [GP, GP, GP, MP, BP, I, P]),
    "ocu_mg_relax_bc": (I, [GP, GP, I, MP, BP, D, D, D, P]),
    "ocu_mg_relax_fused_bc": (I, [GP, GP, GP, MP, BP, D, D, D, I, I, P]),
    "ocu_mg_relax_fused_prolong_bc": (I, [GP, GP, GP, GP, I, MP, BP, D, D, D, I, I, P]),
    "ocu_mg_relax_fused_prolong_ok": (I, [GP, GP, BP]),
    "ocu_mg_relax_fused_ok": (I, [GP, GP, GP, BP]),
    "ocu_mg_prolong_bc_ok": (I, [GP, GP, BP]),
    "ocu_mg_residual": (I, [GP, GP, GP, D, D, D, D, P]),
    "ocu_mg_restrict": (I, [GP, GP, P]),
    "ocu_mg_residual_restrict": (I, [GP, GP, GP, GP, D, D, D, D, DP, P]),
    "ocu_mg_residual_norm": (I, [GP, GP, GP, D, D, D, D, DP, P]),
    "ocu_norm_finalize": (I, [DP, DP, P]),
    "ocu_norm_finalize_tagged": (I, [DP, DP, C.c_ulonglong, P]),
    "ocu_mg_prolong": (I, [GP, GP, I, P]),
    "ocu_mg_prolong_bc": (I, [GP, GP, I, BP, D, D, D, P]),
    "ocu_grid_clear_ghosts": (I, [GP, P]),
    "ocu_grid_stats3": (I, [GP, DP, P]),
    "ocu_passive_advect_adjoint": (I, [GP, GP, GP, GP, GP, GP, GP, GP, GP, D, I, I, I, D, D, D, I, P]),
    "ocu_bc_adjoint_nocorners": (I, [GP, BP, P]),
    "ocu_mac_adjoint_normal": (I, [GP, GP, GP, BP, P]),
    "ocu_divergence_adjoint": (I, [GP, GP, GP, GP, I, I, I, D, D, D, P]),
    "ocu_gradient_adjoint": (I, [GP, GP, GP, GP, D, D, D, D, P]),
    "ocu_mg_bottom_vcycle": (I, [GP, GP, MP, DP, I, BP, I, I, I, I, I, P]),
    "ocu_mg_bottom_lds_bytes": (C.c_size_t, [GP, I, I]),
    "ocu_mg_bottom_lds_max": (C.c_size_t, []),
    "ocu_mg_coarse_relax": (I, [GP, GP, I, I, MP, BP, D, D, D, I, P]),
    "ocu_ir_residual": (I, [GP, GP, D, GP, D, D, D, D, D, DP, P]),
    "ocu_norm3_finalize": (I, [DP, DP, P]),
    "ocu_norm3_finalize_tagged": (I, [DP, DP, C.c_ulonglong, P]),
    "ocu_ir_correct": (I, [GP, GP, D, P]),
    "ocu_ir_correct_ex": (I, [GP, GP, D, DP, I, P]),
    "ocu_ir_shift": (I, [GP, D, D, DP, DP, P]),
    "ocu_norm4_finalize": (I, [DP, DP, P]),
    "ocu_norm4_finalize_tagged": (I, [DP, DP, C.c_ulonglong, P]),
    "ocu_grid_sum": (I, [GP, DP, DP, P]),
    "ocu_norm_parts": (I, []),
    "ocu_divergence": (I, [GP, GP, GP, GP, I, I, I, D, D, D, P]),
    "ocu_gradient_sub": (I, [GP, GP, GP, GP, D, D, D, D, P]),
    "ocu_self_advect": (I, [GP, GP, GP, GP, GP, GP, I, I, I, D, D, D, I, P]),
    "ocu_passive_advect": (I, [GP, GP, GP, GP, GP, I, I, I, D, D, D, I, P]),
    "ocu_self_advect_ex": (I, [GP, GP, GP, GP, GP, GP, I, I, I, D, D, D, I, I, P]),
    "ocu_passive_advect_ex": (I, [GP, GP, GP, GP, GP, I, I, I, D, D, D, I, I, P]),
    "ocu_laplacian_acc": (I, [GP, GP, I, I, I, D, D, D, D, P]),
    "ocu_thermal_force": (I, [GP, GP, I, I, I, I, D, P]),
    "ocu_ns_step_fused": (I, [GP] * 16 + [C.POINTER(NSStepParams), BP, P]),
    "ocu_pcg_apply_laplacian": (I, [GP, GP, GP, D, D, D, P, DP, P]),
    "ocu_pcg_diag_precond": (I, [GP, GP, GP, D, D, D, P]),
    "ocu_grid_dot": (I, [GP, GP, P, DP, P]),
}


class NSParams(C.Structure):
    """ocu_ns_params (include/ocu.h)"""
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("hx", C.c_double), ("hy", C.c_double), ("hz", C.c_double),
        ("max_divergence", C.c_double),
        ("temp_bc", BCSet), ("flow_bc", BCSet),
        ("vertical_direction", C.c_int32),
        ("gravity", C.c_double), ("bouyancy", C.c_double),
        ("thermal_diffusion", C.c_double), ("viscosity", C.c_double),
        ("advection_scheme", C.c_int32), ("time_step", C.c_int32),
        ("cfl_factor", C.c_double),
        ("mixed_projection", C.c_int32),
        ("unfused_step", C.c_int32),
    ]


class ScalarParams(C.Structure):
    """ocu_scalar_params (include/ocu.h)"""
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("hx", C.c_double), ("hy", C.c_double), ("hz", C.c_double),
        ("bc", BCSet),
        ("advection_scheme", C.c_int32),
    ]


HOST_SIGS = {
    "ocu_set_compute_stream": (I, [P]),
    "ocu_get_compute_stream": (P, []),
    "ocu_mg_create": (P, [GP, D, D, D, CP, I]),
    "ocu_mg_pressure": (I, [P, GP]),
    "ocu_mg_solve": (I, [P, D, I, DP]),
    "ocu_mg_run_vcycles": (I, [P, I]),
    "ocu_mg_run_fmg": (I, [P, I]),
    "ocu_mg_stats": (I, [P, IP, IP, DP, I, IP]),
    "ocu_mg_graph_counts": (I, [P, C.POINTER(C.c_longlong)]),
    "ocu_mg_destroy": (None, [P]),
    "ocu_mg_sweep_counts": (I, [P, C.POINTER(C.c_longlong)]),
    "ocu_mg_get_level_params": (I, [P, I, C.POINTER(MGLevelParams)]),
    "ocu_mgmixed_create": (P, [GP, D, D, D, CP, I, I, I, I]),
    "ocu_mgmixed_pressure": (I, [P, GP]),
    "ocu_mgmixed_solve": (I, [P, D, I, DP]),
    "ocu_mgmixed_stats": (I, [P, IP, IP, DP, DP, I, IP]),
    "ocu_mgmixed_destroy": (None, [P]),
    "ocu_mgmixed_sweep_counts": (I, [P, C.POINTER(C.c_longlong)]),
    "ocu_mgmixed_get_level_params": (I, [P, I, C.POINTER(MGLevelParams)]),
    "ocu_project_create": (P, [GP, GP, GP, BP, I, I, I, D, D, D, I]),
    "ocu_project_solve": (I, [P, D, DP]),
    "ocu_project_solve_divergence_only": (I, [P]),
    "ocu_project_divergence": (I, [P, GP]),
    "ocu_project_pressure": (I, [P, GP]),
    "ocu_project_destroy": (None, [P]),
    "ocu_nc_writer_open": (P, [C.c_char_p, I, I, I, C.c_float, C.c_float, C.c_float]),
    "ocu_nc_writer_define_variable": (I, [P, C.c_char_p, I, I, C.c_float, C.c_float, C.c_float]),
    "ocu_nc_writer_add_time_level": (I, [P, C.c_float, IP]),
    "ocu_nc_writer_add_data": (I, [P, C.c_char_p, GP, I]),
    "ocu_nc_writer_close": (I, [P]),
    "ocu_nc_reader_open": (P, [C.c_char_p]),
    "ocu_nc_reader_info": (I, [P, IP, IP, IP, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float), IP]),
    "ocu_nc_reader_time": (I, [P, I, C.POINTER(C.c_float)]),
    "ocu_nc_reader_variable": (I, [P, I, C.c_char_p, I, IP, IP, IP, IP]),
    "ocu_nc_reader_read": (I, [P, C.c_char_p, GP, I]),
    "ocu_nc_reader_close": (None, [P]),
    "ocu_pcg_create": (P, [GP, GP, C.c_float, C.c_float, C.c_float, BP, I, I, I, I]),
    "ocu_pcg_solve": (I, [P, D, I, DP]),
    "ocu_pcg_iterations": (I, [P]),
    "ocu_pcg_pressure": (I, [P, GP]),
    "ocu_pcg_destroy": (None, [P]),
    "ocu_ns_create": (P, [I, C.POINTER(NSParams), GP, GP, GP, GP]),
    "ocu_ns_advance_one_step": (I, [P, D]),
    "ocu_ns_advance": (I, [P, D]),
    "ocu_ns_max_stable_timestep": (I, [P, DP]),
    "ocu_ns_field": (I, [P, I, GP]),
    "ocu_ns_num_steps": (I, [P]),
    "ocu_ns_destroy": (None, [P]),
    "ocu_scalar_create": (P, [I, C.POINTER(ScalarParams), GP, GP, GP, GP]),
    "ocu_scalar_advance_one_step": (I, [P, D]),
    "ocu_scalar_advance": (I, [P, D]),
    "ocu_scalar_max_stable_timestep": (I, [P, DP]),
    "ocu_scalar_field": (I, [P, I, GP]),
    "ocu_scalar_destroy": (None, [P]),
}


def _bind(lib, sigs):
    for name, (res, args) in sigs.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args


def _load(name):
    path = LIBDIR / name
    if not path.exists():
        raise NativeLibraryMissing(
            f"{path} is missing: build it with `python -m opencurrent_amd.build` "
            "(there is no non-native fallback)")
    return C.CDLL(str(path), mode=C.RTLD_GLOBAL)


def hip():
    """libocu_hip.so with argument types bound."""
    global _hip
    if _hip is None:
        lib = _load("libocu_hip.so")
        _bind(lib, HIP_SIGS)
        _hip = lib
    return _hip


def host():
    """libocu.so (loads libocu_hip.so first)."""
    global _host
    if _host is None:
        hip()
        lib = _load("libocu.so")
        _bind(lib, HOST_SIGS)
        _host = lib
    return _host


def last_error():
    return hip().ocu_last_error().decode(errors="replace")


def check(status, what="ocu call"):
    """Raise OcuError for a negative status; return the status otherwise."""
    if status < 0:
        raise OcuError(f"{what} failed ({status}): {last_error()}")
    return status


def library_paths():
    return [str(LIBDIR / "libocu_hip.so"), str(LIBDIR / "libocu.so")]
