"""Python front-end of the drop-in solver classes (include/ocu.h).

Names follow the reference's C++ classes: MultigridPressure3D wraps
Sol_MultigridPressure3DDevice<T>, MultigridPressureMixed3D wraps
Sol_MultigridPressureMixed3DDeviceD, ProjectDivergence3D wraps
Sol_ProjectDivergence3DDevice<T>, IncompressibleNS3D wraps
Eqn_IncompressibleNS3D<T>.  Input grids are DeviceGrid objects that must stay
alive as long as the solver (the solvers keep pointers to them, as the
reference does).
"""
import ctypes as C

import numpy as np

from . import _lib
from ._types import (BCSet, CONVERGENCE_L2, Grid3DDesc, MGConfig, MGLevelParams, OCU_F32, OCU_F64, dtype_code,
                     layout)
from .device import desc_to_host, host_to_desc


MG_FUSED_RELAX, MG_LDS_BOTTOM, MG_ADAPTIVE_IR, MG_FUSED_PROLONG, MG_GRAPHS = 1, 2, 4, 8, 16


def _options(use_fused_relax, use_bottom, use_fused_prolong=True, use_graphs=False):
    return ((MG_FUSED_RELAX if use_fused_relax else 0) | (MG_LDS_BOTTOM if use_bottom else 0) |
            (MG_FUSED_PROLONG if use_fused_prolong else 0) | (MG_GRAPHS if use_graphs else 0))


def _history(fn, h, extra=()):
    cap = 4096
    hist = (C.c_double * cap)()
    n = C.c_int()
    return hist, n


class MultigridPressure3D:
    """Sol_MultigridPressure3DDevice<T> (ocuequation/sol_mgpressure3d.h:130-181)."""

    def __init__(self, rhs, hx, hy, hz, bc, nu1=2, nu2=2, convergence=CONVERGENCE_L2, sym=False,
                 disable_relax_on_host=False, use_fused_relax=True, use_bottom=True, use_fused_prolong=True,
                 use_graphs=False):
        self._rhs = rhs  # keep alive
        self.cfg = MGConfig.make(bc, nu1, nu2, convergence, sym, disable_relax_on_host)
        self.h = _lib.host().ocu_mg_create(rhs.ptr(), hx, hy, hz, C.byref(self.cfg),
                                           _options(use_fused_relax, use_bottom, use_fused_prolong, use_graphs))
        if not self.h:
            raise _lib.OcuError(f"Sol_MultigridPressure3DDevice::initialize_storage failed: {_lib.last_error()}")

    def pressure_desc(self):
        d = Grid3DDesc()
        _lib.check(_lib.host().ocu_mg_pressure(self.h, C.byref(d)), "pressure")
        return d

    def pressure(self):
        return desc_to_host(self.pressure_desc())

    def set_pressure(self, host_grid):
        host_to_desc(host_grid, self.pressure_desc())

    def solve(self, tol=1e-6, max_iter=4):
        r = C.c_double()
        ok = _lib.check(_lib.host().ocu_mg_solve(self.h, tol, max_iter, C.byref(r)), "solve")
        return bool(ok), r.value

    def run_vcycles(self, count):
        return bool(_lib.check(_lib.host().ocu_mg_run_vcycles(self.h, count), "run_vcycles"))

    def run_fmg(self, count):
        return bool(_lib.check(_lib.host().ocu_mg_run_fmg(self.h, count), "run_fmg"))

    def stats(self):
        nl, vc, n = C.c_int(), C.c_int(), C.c_int()
        hist = (C.c_double * 4096)()
        _lib.check(_lib.host().ocu_mg_stats(self.h, C.byref(nl), C.byref(vc), hist, 4096, C.byref(n)), "stats")
        return {"num_levels": nl.value, "vcycles": vc.value, "history": [hist[i] for i in range(min(n.value, 4096))]}

    def level_params(self, level):
        """MGLevelParams of a level as the solver passes them to the kernels"""
        p = MGLevelParams()
        _lib.check(_lib.host().ocu_mg_get_level_params(self.h, level, C.byref(p)), "level_params")
        return p

    def sweep_counts(self):
        """level-0 fused sweeps so far: (plain, from the zero grid, adding a prolongation)"""
        c = (C.c_longlong * 3)()
        _lib.check(_lib.host().ocu_mg_sweep_counts(self.h, c), "sweep_counts")
        return tuple(c)

    def graph_counts(self):
        """V-cycle segments recorded as HIP graphs and replayed so far: (recorded, replayed)"""
        c = (C.c_longlong * 2)()
        _lib.check(_lib.host().ocu_mg_graph_counts(self.h, c), "graph_counts")
        return tuple(c)

    def __del__(self):
        if getattr(self, "h", None):
            _lib.host().ocu_mg_destroy(self.h)
            self.h = None


class MultigridPressureMixed3D:
    """Sol_MultigridPressureMixed3DDeviceD: FP32 V-cycles inside FP64 iterative refinement."""

    def __init__(self, rhs, hx, hy, hz, bc, nu1=2, nu2=2, convergence=CONVERGENCE_L2, cycles=4,
                 cycles_first=4, fmg_first=False, use_fused_relax=True, use_bottom=True, adaptive=True,
                 use_fused_prolong=True):
        self._rhs = rhs
        self.cfg = MGConfig.make(bc, nu1, nu2, convergence)
        self.h = _lib.host().ocu_mgmixed_create(rhs.ptr(), hx, hy, hz, C.byref(self.cfg), cycles, cycles_first,
                                                int(fmg_first),
                                                _options(use_fused_relax, use_bottom, use_fused_prolong) |
                                                (MG_ADAPTIVE_IR if adaptive else 0))
        if not self.h:
            raise _lib.OcuError(f"Sol_MultigridPressureMixed3DDeviceD::initialize_storage failed: {_lib.last_error()}")

    def pressure_desc(self):
        d = Grid3DDesc()
        _lib.check(_lib.host().ocu_mgmixed_pressure(self.h, C.byref(d)), "pressure")
        return d

    def pressure(self):
        return desc_to_host(self.pressure_desc())

    def set_pressure(self, host_grid):
        host_to_desc(host_grid, self.pressure_desc())

    def solve(self, tol=1e-6, max_iter=10):
        r = C.c_double()
        ok = _lib.check(_lib.host().ocu_mgmixed_solve(self.h, tol, max_iter, C.byref(r)), "solve")
        return bool(ok), r.value

    def stats(self):
        it, vc, n = C.c_int(), C.c_int(), C.c_int()
        rms = C.c_double()
        hist = (C.c_double * 4096)()
        _lib.check(_lib.host().ocu_mgmixed_stats(self.h, C.byref(it), C.byref(vc), C.byref(rms), hist, 4096,
                                                 C.byref(n)), "stats")
        return {"ir_iterations": it.value, "vcycles": vc.value, "rhs_rms": rms.value,
                "history": [hist[i] for i in range(min(n.value, 4096))]}

    def level_params(self, level):
        """MGLevelParams of a level of the FP32 hierarchy"""
        p = MGLevelParams()
        _lib.check(_lib.host().ocu_mgmixed_get_level_params(self.h, level, C.byref(p)), "level_params")
        return p

    def sweep_counts(self):
        """level-0 fused sweeps of the FP32 hierarchy so far: (plain, from
        the zero grid, adding a prolongation)"""
        c = (C.c_longlong * 3)()
        _lib.check(_lib.host().ocu_mgmixed_sweep_counts(self.h, c), "sweep_counts")
        return tuple(c)

    def __del__(self):
        if getattr(self, "h", None):
            _lib.host().ocu_mgmixed_destroy(self.h)
            self.h = None


class ProjectDivergence3D:
    """Sol_ProjectDivergence3DDevice<T> (ocuequation/sol_project3d.h:74-93)."""

    def __init__(self, u, v, w, bc, nx, ny, nz, hx, hy, hz, mixed=False):
        self._keep = (u, v, w)
        self.bc = bc
        self.h = _lib.host().ocu_project_create(u.ptr(), v.ptr(), w.ptr(), C.byref(bc), nx, ny, nz, hx, hy, hz,
                                                int(mixed))
        if not self.h:
            raise _lib.OcuError(f"Sol_ProjectDivergence3DDevice::initialize_storage failed: {_lib.last_error()}")

    def solve(self, tol=1e-5):
        r = C.c_double()
        ok = _lib.check(_lib.host().ocu_project_solve(self.h, tol, C.byref(r)), "project solve")
        return bool(ok), r.value

    def solve_divergence_only(self):
        return bool(_lib.check(_lib.host().ocu_project_solve_divergence_only(self.h), "divergence"))

    def divergence(self):
        d = Grid3DDesc()
        _lib.check(_lib.host().ocu_project_divergence(self.h, C.byref(d)), "divergence")
        return desc_to_host(d)

    def pressure(self):
        d = Grid3DDesc()
        _lib.check(_lib.host().ocu_project_pressure(self.h, C.byref(d)), "pressure")
        return desc_to_host(d)

    def __del__(self):
        if getattr(self, "h", None):
            _lib.host().ocu_project_destroy(self.h)
            self.h = None


_INIT_PADS = ((0, 1, 1), (1, 0, 1), (1, 1, 0), (1, 1, 1))


def _check_init_grids(what, params, dt, grids):
    """The init grids must have the init_grids layouts of params' size
    (eqn_incompressns3d.h:78-86, eqn_scalaradvection3d.h init_grids) and
    one element type; the library checks the same and copies whole
    allocations."""
    nx, ny, nz = params.nx, params.ny, params.nz
    dims = ((nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1), (nx, ny, nz))
    for name, g, d, pad in zip(("u", "v", "w", "scalar"), grids, dims, _INIT_PADS):
        want = layout(*d, 1, 1, 1, *pad, dtype=dt)
        got = g.desc
        fields = ("nx", "ny", "nz", "gx", "gy", "gz", "pnx", "pny", "pnz", "dtype", "shift", "allocated")
        if any(getattr(got, f) != getattr(want, f) for f in fields):
            raise ValueError(f"{what}: init grid {name} does not have the init_grids layout for "
                             f"{(nx, ny, nz)} / dtype {dt}")


class PCGPressure3D:
    """Sol_PCGPressure3DDevice<T> (ocuequation/sol_pcgpressure3d.h:41-101):
    -div(c grad p) = rhs for device grids rhs and coeff (congruent, kept
    alive here).  preconditioner: 0 none, 1 Jacobi, 2 multigrid, 3 BFBT
    (the reference's default)."""

    NONE, JACOBI, MULTIGRID, BFBT = 0, 1, 2, 3

    def __init__(self, rhs, coeff, hx, hy, hz, bc, preconditioner=3, multigrid_cycles=2, use_fmg=True,
                 convergence=CONVERGENCE_L2):
        self._keep = (rhs, coeff)
        self.h = _lib.host().ocu_pcg_create(rhs.ptr(), coeff.ptr(), hx, hy, hz, C.byref(bc), convergence,
                                            preconditioner, multigrid_cycles, int(use_fmg))
        if not self.h:
            raise _lib.OcuError(f"Sol_PCGPressure3DDevice::initialize_storage failed: {_lib.last_error()}")

    def solve(self, tol=1e-6, max_iter=1000):
        r = C.c_double()
        ok = _lib.check(_lib.host().ocu_pcg_solve(self.h, tol, max_iter, C.byref(r)), "solve")
        return bool(ok), r.value

    @property
    def iterations(self):
        return _lib.host().ocu_pcg_iterations(self.h)

    def pressure_desc(self):
        d = Grid3DDesc()
        _lib.check(_lib.host().ocu_pcg_pressure(self.h, C.byref(d)), "pressure")
        return d

    def pressure(self):
        return desc_to_host(self.pressure_desc())

    def set_pressure(self, host_grid):
        host_to_desc(host_grid, self.pressure_desc())

    def __del__(self):
        if getattr(self, "h", None):
            _lib.host().ocu_pcg_destroy(self.h)
            self.h = None


class IncompressibleNS3D:
    """Eqn_IncompressibleNS3D<T> (ocuequation/eqn_incompressns3d.h:132-175).

    init_u/v/w/temp are HostGrids laid out like
    Eqn_IncompressibleNS3DParams::init_grids (paddings (0,1,1), (1,0,1),
    (1,1,0), (1,1,1))."""

    def __init__(self, params, init_u, init_v, init_w, init_temp):
        dt = dtype_code(init_u.buf.dtype)
        _check_init_grids("IncompressibleNS3D", params, dt, (init_u, init_v, init_w, init_temp))
        self.dtype = dt
        self.params = params
        self._init = (init_u, init_v, init_w, init_temp)
        self.h = _lib.host().ocu_ns_create(dt, C.byref(params), init_u.ptr(), init_v.ptr(), init_w.ptr(),
                                           init_temp.ptr())
        if not self.h:
            raise _lib.OcuError(f"Eqn_IncompressibleNS3D::set_parameters failed: {_lib.last_error()}")

    def advance_one_step(self, dt):
        return bool(_lib.check(_lib.host().ocu_ns_advance_one_step(self.h, dt), "advance_one_step"))

    def advance(self, dt):
        return bool(_lib.check(_lib.host().ocu_ns_advance(self.h, dt), "advance"))

    def get_max_stable_timestep(self):
        d = C.c_double()
        _lib.check(_lib.host().ocu_ns_max_stable_timestep(self.h, C.byref(d)), "max_stable_timestep")
        return d.value

    def field_desc(self, which):
        d = Grid3DDesc()
        _lib.check(_lib.host().ocu_ns_field(self.h, which, C.byref(d)), "field")
        return d

    def field(self, which):
        """0 u, 1 v, 2 w, 3 temperature, 4 pressure, copied to a HostGrid."""
        return desc_to_host(self.field_desc(which))

    @property
    def num_steps(self):
        return _lib.host().ocu_ns_num_steps(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            _lib.host().ocu_ns_destroy(self.h)
            self.h = None


class ScalarAdvection3D:
    """Eqn_ScalarAdvection3D<T> (ocuequation/eqn_scalaradvection3d.h,
    eqn_scalaradvection3d.cpp): forward-Euler passive advection of a scalar
    by a fixed MAC velocity.  init_u/v/w/phi are HostGrids laid out like
    Eqn_ScalarAdvection3DParams::init_grids."""

    def __init__(self, params, init_u, init_v, init_w, init_phi):
        dt = dtype_code(init_u.buf.dtype)
        _check_init_grids("ScalarAdvection3D", params, dt, (init_u, init_v, init_w, init_phi))
        self.dtype = dt
        self.params = params
        self.h = _lib.host().ocu_scalar_create(dt, C.byref(params), init_u.ptr(), init_v.ptr(), init_w.ptr(),
                                               init_phi.ptr())
        if not self.h:
            raise _lib.OcuError(f"Eqn_ScalarAdvection3D::set_parameters failed: {_lib.last_error()}")

    def advance_one_step(self, dt):
        return bool(_lib.check(_lib.host().ocu_scalar_advance_one_step(self.h, dt), "advance_one_step"))

    def advance(self, dt):
        return bool(_lib.check(_lib.host().ocu_scalar_advance(self.h, dt), "advance"))

    def get_max_stable_timestep(self):
        d = C.c_double()
        _lib.check(_lib.host().ocu_scalar_max_stable_timestep(self.h, C.byref(d)), "max_stable_timestep")
        return d.value

    def field(self, which):
        """0 u, 1 v, 2 w, 3 phi, 4 deriv_phidt, copied to a HostGrid."""
        d = Grid3DDesc()
        _lib.check(_lib.host().ocu_scalar_field(self.h, which, C.byref(d)), "field")
        return desc_to_host(d)

    def __del__(self):
        if getattr(self, "h", None):
            _lib.host().ocu_scalar_destroy(self.h)
            self.h = None
