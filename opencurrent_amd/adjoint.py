"""Gradients through the advection-projection path.

Forward model (the passive-scalar part of the projection-method step, as
Eqn_ScalarAdvection3D::advance_one_step chains it, eqn_scalaradvection3d.cpp):

    u0          = P(v)                         Sol_ProjectDivergence3DDevice::solve
    phi_{n+1}   = fill(phi_n + dt*A(u0, phi_n))  n = 0 .. steps-1
    L           = 1/2 * sum_interior (phi_steps - target)^2

with A the passive advection (Sol_PassiveAdvection3DDevice) and fill the
device no-corner ghost fill.  The reference has no adjoint; this one is built
from hand-written kernels of libocu_hip.so: ocu_passive_advect_adjoint,
ocu_bc_adjoint_nocorners, ocu_mac_adjoint_normal, ocu_gradient_adjoint,
ocu_divergence_adjoint.  The projection P = M (I - G B S D) M (MAC fill M,
divergence D, pressure solve S, pressure ghost fill B, gradient G) has the
adjoint P^T = M^T (I - D^T S B^T G^T) M^T with S^T = S, since the
cell-centred Poisson operator with the projection's boundary conditions is
symmetric: the backward pass runs one more multigrid solve on the adjoint
right-hand side.  On a periodic grid D^T = -G and P is self-adjoint.
The states are checkpointed every `checkpoint_every` steps; the backward pass
recomputes one segment at a time.

`AdvectProjectLoss.apply(u, v, w, problem)` exposes the pair as a
torch.autograd.Function over the flat Grid3D buffers of the MAC velocity.
The parameters are the normal faces 0..n-1 of interior rows (faces the MAC
fill overwrites get a zero gradient); the gradient is zero everywhere else.
"""
import ctypes as C

import numpy as np

from . import _lib
from . import device as D
from . import _types as T
from ._types import BC, BCSet, CONVERGENCE_LINF
from .solvers import MultigridPressure3D, ProjectDivergence3D


def _mac(nx, ny, nz, dtype):
    """u, v, w and a scalar grid laid out like Eqn_IncompressibleNS3DParams::init_grids"""
    return (D.DeviceGrid(nx + 1, ny, nz, 1, 1, 1, 0, 1, 1, dtype=dtype),
            D.DeviceGrid(nx, ny + 1, nz, 1, 1, 1, 1, 0, 1, dtype=dtype),
            D.DeviceGrid(nx, ny, nz + 1, 1, 1, 1, 1, 1, 0, dtype=dtype),
            D.DeviceGrid(nx, ny, nz, 1, 1, 1, 1, 1, 1, dtype=dtype))


def parameter_mask(g, axis):
    """bool view over g's buffer: True at the independent normal faces"""
    t = D.torch()
    m = t.zeros_like(g.tensor, dtype=t.bool)
    d = g.desc
    pre = 16 - d.gz
    full = m[pre:].view(d.pnx, d.pny, d.pnz)
    nx, ny, nz = d.nx, d.ny, d.nz
    sl = [slice(d.gx, d.gx + nx), slice(d.gy, d.gy + ny), slice(d.gz, d.gz + nz)]
    n = (nx, ny, nz)[axis]
    sl[axis] = slice((d.gx, d.gy, d.gz)[axis], (d.gx, d.gy, d.gz)[axis] + n - 1)  # faces 0..n-2 of the n faces
    full[tuple(sl)] = True
    return m


def poisson_bc(flow_bc):
    """convert_bc_to_poisson_eqn (sol_project3d.cpp:26-43): periodic stays
    periodic, forced inflow becomes homogeneous Neumann"""
    s = BCSet.uniform(T.BC_PERIODIC)
    for f in ("xpos", "xneg", "ypos", "yneg", "zpos", "zneg"):
        t = getattr(flow_bc, f).type
        if t == T.BC_FORCED_INFLOW_VARIABLE_SLIP:
            setattr(s, f, BC(T.BC_NEUMANN, 0.0, 0.0))
        elif t != T.BC_PERIODIC:
            raise _lib.OcuError(f"invalid flow boundary condition type {t}")
    return s


class _Desc:
    def __init__(self, d):
        self.desc = d

    def ptr(self):
        return C.byref(self.desc)


class AdvectProjectProblem:
    """The model above on an nx*ny*nz grid, with its adjoint.  flow_bc (MAC
    velocity: periodic or forced inflow) and phi_bc (scalar) default to
    periodic."""

    def __init__(self, nx, ny, nz, hx, hy, hz, dt, steps, scheme=T.IT_SECOND_ORDER_CENTERED, dtype=np.float64,
                 checkpoint_every=5, proj_tol=None, flow_bc=None, phi_bc=None):
        self.n = (nx, ny, nz)
        self.h = (hx, hy, hz)
        self.dt, self.steps, self.scheme, self.ce = float(dt), int(steps), int(scheme), int(checkpoint_every)
        self.dtype = np.dtype(dtype)
        # max |div| after the projection: FP32 multigrid reaches ~1e-3 on O(1/h) divergences
        self.tol = proj_tol if proj_tol is not None else (1e-12 if self.dtype == np.float64 else 1e-3)
        self.flow_bc = flow_bc.copy() if flow_bc is not None else BCSet.uniform(T.BC_PERIODIC)
        self.phi_bc = phi_bc.copy() if phi_bc is not None else BCSet.uniform(T.BC_PERIODIC)
        self.u, self.v, self.w, self.phi = _mac(nx, ny, nz, dtype)
        self.au, self.av, self.aw, _ = _mac(nx, ny, nz, dtype)
        self.proj = ProjectDivergence3D(self.u, self.v, self.w, self.flow_bc, nx, ny, nz, hx, hy, hz)
        like = self.phi
        # the adjoint pressure solve: right-hand side lam_p, the projection's
        # Poisson boundary conditions and stopping rule
        self.lam_p = D.DeviceGrid.congruent(like)
        self.mg_adj = MultigridPressure3D(self.lam_p, hx, hy, hz, poisson_bc(self.flow_bc),
                                          convergence=CONVERGENCE_LINF)
        self.phi0, self.target, self.dphi = (D.DeviceGrid.congruent(like) for _ in range(3))
        # seed = phi_steps - target of the last forward; backward works on
        # copies in lam / lam2, so it can be repeated
        self.seed = D.DeviceGrid.congruent(like)
        self.lam, self.lam2 = D.DeviceGrid.congruent(like), D.DeviceGrid.congruent(like)
        nseg = (self.steps + self.ce - 1) // self.ce
        self.ck = [D.DeviceGrid.congruent(like) for _ in range(nseg)]
        self.seg = [D.DeviceGrid.congruent(like) for _ in range(min(self.ce, self.steps))]
        self.masks = [parameter_mask(g, q) for q, g in enumerate((self.u, self.v, self.w))]
        self.loss = None
        # number of forward passes run; a gradient belongs to the forward
        # whose checkpoints it replays (see AdvectProjectLoss)
        self.generation = 0

    # -- pieces ----------------------------------------------------------
    def _project(self, proj):
        pd = T.Grid3DDesc()
        _lib.check(_lib.host().ocu_project_pressure(proj.h, C.byref(pd)), "pressure")
        _lib.check(_lib.hip().ocu_grid_clear(C.byref(pd), 0.0, None), "clear pressure")  # same start every solve
        ok, res = proj.solve(self.tol)
        if not ok:
            raise _lib.OcuError(f"projection did not converge to {self.tol:g} (residual {res:g})")

    def _project_adjoint(self):
        """(au, av, aw) <- P^T (au, av, aw), see the module doc"""
        nx, ny, nz = self.n
        D.mac_adjoint_normal(self.au, self.av, self.aw, self.flow_bc)
        self.lam_p.tensor.zero_()
        D.gradient_adjoint(self.au, self.av, self.aw, self.lam_p, -1.0, *self.h)
        D.bc_adjoint_nocorners(self.lam_p, self.mg_adj.cfg.bc)
        pd = self.mg_adj.pressure_desc()
        _lib.check(_lib.hip().ocu_grid_clear(C.byref(pd), 0.0, None), "clear adjoint pressure")
        # the forward tolerance bounds |div| of an O(1) velocity; the adjoint
        # right-hand side carries the loss's scale, so the bound scales with it
        tol = self.tol * max(1.0, D.grid_reduce(self.lam_p, 0))
        ok, res = self.mg_adj.solve(tol, 15)
        if not ok:
            raise _lib.OcuError(f"adjoint pressure solve did not converge to {tol:g} (residual {res:g})")
        D.divergence_adjoint(_Desc(pd), self.au, self.av, self.aw, nx, ny, nz, *self.h)
        D.mac_adjoint_normal(self.au, self.av, self.aw, self.flow_bc)

    def _step(self, src, dst):
        nx, ny, nz = self.n
        D.passive_advect(self.u, self.v, self.w, src, self.dphi, nx, ny, nz, *self.h, self.scheme)
        D.grid_lincomb(dst, 1.0, src, self.dt, self.dphi)
        D.bc_fill_nocorners(dst, self.phi_bc, *self.h)

    def set_scalars(self, phi0, target):
        """initial scalar and target (tensors or arrays over the interior)"""
        t = D.torch()
        for g, x in ((self.phi0, phi0), (self.target, target)):
            g.tensor.zero_()
            g.interior().copy_(t.as_tensor(np.asarray(x) if not t.is_tensor(x) else x, dtype=g.tensor.dtype))
        D.bc_fill_nocorners(self.phi0, self.phi_bc, *self.h)

    # -- forward / backward ------------------------------------------------
    def forward(self, u, v, w):
        """loss for the MAC velocity buffers u, v, w (flat tensors shaped like
        the problem's grids; only the parameter faces are read)"""
        for g, x, m in zip((self.u, self.v, self.w), (u, v, w), self.masks):
            g.tensor.zero_()
            g.tensor[m] = x.detach().to(g.tensor.dtype)[m]
        self._project(self.proj)
        D.grid_copy(self.phi, self.phi0)
        for n in range(self.steps):
            if n % self.ce == 0:
                D.grid_copy(self.ck[n // self.ce], self.phi)
            self._step(self.phi, self.phi)
        self.seed.tensor.zero_()
        D.grid_lincomb(self.seed, 1.0, self.phi, -1.0, self.target)
        D.grid_clear_ghosts(self.seed)
        self.loss = .5 * D.grid_reduce(self.seed, 4)
        self.generation += 1
        return self.loss

    def backward(self):
        """gradients of the last forward's loss w.r.t. u, v, w (flat tensors)
        and w.r.t. the initial scalar's interior; repeatable until the next
        forward"""
        if self.loss is None:
            raise RuntimeError("AdvectProjectProblem.backward: run forward first")
        nx, ny, nz = self.n
        for g in (self.au, self.av, self.aw):
            g.tensor.zero_()
        lam, lam2 = self.lam, self.lam2
        D.grid_copy(lam, self.seed)
        for s in reversed(range(len(self.ck))):
            n0, n1 = s * self.ce, min(s * self.ce + self.ce, self.steps)
            D.grid_copy(self.seg[0], self.ck[s])
            for m in range(1, n1 - n0):
                self._step(self.seg[m - 1], self.seg[m])
            for n in reversed(range(n0, n1)):
                D.bc_adjoint_nocorners(lam, self.phi_bc)
                D.passive_advect_adjoint(self.u, self.v, self.w, self.seg[n - n0], lam, lam2, self.au, self.av,
                                         self.aw, self.dt, nx, ny, nz, *self.h, self.scheme)
                lam, lam2 = lam2, lam
        D.bc_adjoint_nocorners(lam, self.phi_bc)  # the initial fill of phi0
        grad_phi0 = lam.interior().clone()
        self._project_adjoint()
        grads = []
        for g, m in zip((self.au, self.av, self.aw), self.masks):
            x = g.tensor.clone()
            x[~m] = 0
            grads.append(x)
        return grads[0], grads[1], grads[2], grad_phi0


def _function():
    torch = D.torch()

    class AdvectProjectLoss(torch.autograd.Function):
        """loss = problem.forward(u, v, w); gradients from problem.backward().
        The problem keeps one forward's checkpoints: a loss whose problem has
        run forward again since cannot be differentiated (RuntimeError);
        backward of the same loss may be repeated (retain_graph)."""

        @staticmethod
        def forward(ctx, u, v, w, problem):
            ctx.problem = problem
            loss = problem.forward(u, v, w)
            ctx.generation = problem.generation
            return torch.tensor(loss, dtype=torch.float64, device=u.device)

        @staticmethod
        def backward(ctx, g):
            if ctx.problem.generation != ctx.generation:
                raise RuntimeError("advect_project_loss: the problem ran forward again after this loss was "
                                   "computed, so its checkpoints belong to another input; use one problem per "
                                   "loss in a graph")
            gu, gv, gw, _ = ctx.problem.backward()
            s = g.to(gu.dtype)
            return (gu * s).to(gu.dtype), (gv * s).to(gv.dtype), (gw * s).to(gw.dtype), None

    return AdvectProjectLoss


_FN = None


def advect_project_loss(u, v, w, problem):
    """torch-differentiable loss of AdvectProjectProblem (see module doc)"""
    global _FN
    if _FN is None:
        _FN = _function()
    return _FN.apply(u, v, w, problem)
