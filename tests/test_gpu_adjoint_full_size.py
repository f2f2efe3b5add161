"""The gradient model at the sizes its workloads run at.

* Lock exchange, 512 x 512 x 256 cells, hx = hy = 1, hz = 0.5, no-slip
  walls, FP64: the projection's adjoint satisfies <P a, b> = <a, P^T b> to
  the pressure solves' tolerance, and a forward/backward pass of the
  advection-projection model gives a finite loss and gradient.
* Passive advection at 128^3, 20 steps after the projection, checkpoints
  every 5: the FP32 gradient tracks the FP64 one, both are finite, and the
  FP64 gradient agrees with central differences along two seeded
  directions (eps = 1e-4).
"""
import numpy as np
import pytest

from opencurrent_amd import _types as T
from opencurrent_amd._types import BC, BCSet
from fields import uniform

pytestmark = pytest.mark.gpu


def _log(s):
    print(s, flush=True)


def _velocity(p, seeds, mean=(0.0, 0.0, 0.0), amp=1.0):
    import torch
    out = []
    for q, g in enumerate((p.u, p.v, p.w)):
        x = torch.as_tensor(mean[q] + amp * (2 * uniform(seeds[q], tuple(g.tensor.shape)) - 1),
                            dtype=g.tensor.dtype, device=g.tensor.device)
        x[~p.masks[q]] = 0
        out.append(x)
    return out


def _dot(a, b):
    return sum(float((x.double() * y.double()).sum()) for x, y in zip(a, b))


def _norm(a):
    return _dot(a, a) ** .5


def test_lock_exchange_projection_adjoint_512(gpu):
    import torch
    from opencurrent_amd import adjoint
    nx, ny, nz = 512, 512, 256
    walls = BCSet.uniform(T.BC_FORCED_INFLOW_VARIABLE_SLIP)
    for f in ("xpos", "xneg", "ypos", "yneg", "zpos", "zneg"):
        setattr(walls, f, BC(T.BC_FORCED_INFLOW_VARIABLE_SLIP, 0.0, 1.0))
    p = adjoint.AdvectProjectProblem(nx, ny, nz, 1.0, 1.0, .5, dt=.25, steps=2, dtype=np.float64,
                                     checkpoint_every=1, flow_bc=walls, phi_bc=BCSet.uniform(T.BC_NEUMANN))
    a = _velocity(p, (5101, 5102, 5103))
    c = _velocity(p, (5104, 5105, 5106))
    b = [x + y for x, y in zip(a, c)]  # correlated with a, so <Pa, b> is of the order of |a||b|
    for g, y, m in zip((p.u, p.v, p.w), a, p.masks):
        g.tensor.zero_()
        g.tensor[m] = y[m]
    p._project(p.proj)
    pa = [torch.where(m, g.tensor, torch.zeros_like(g.tensor)) for g, m in zip((p.u, p.v, p.w), p.masks)]
    for g, y, m in zip((p.au, p.av, p.aw), b, p.masks):
        g.tensor.zero_()
        g.tensor[m] = y[m]
    p._project_adjoint()
    ptb = [torch.where(m, g.tensor, torch.zeros_like(g.tensor)) for g, m in zip((p.au, p.av, p.aw), p.masks)]
    lhs, rhs = _dot(pa, b), _dot(a, ptb)
    scale = _norm(a) * _norm(b)
    _log(f"<Pa,b> {lhs:.12e}  <a,P^T b> {rhs:.12e}  |a||b| {scale:.3e}  rel {abs(lhs - rhs) / scale:.2e}")
    # the 1e-12 pressure solves leave a velocity error of ~1e-10 per face
    assert abs(lhs - rhs) <= 1e-8 * scale
    assert abs(lhs) > .1 * scale

    # one forward/backward pass of the model on the same grid and walls
    x = (np.arange(nx) + .5) / nx
    y = (np.arange(ny) + .5) / ny
    z = (np.arange(nz) + .5) / nz
    r0 = (x[:, None, None] - .4) ** 2 + (y[None, :, None] - .5) ** 2 + (z[None, None, :] - .5) ** 2
    r1 = (x[:, None, None] - .45) ** 2 + (y[None, :, None] - .5) ** 2 + (z[None, None, :] - .5) ** 2
    p.set_scalars(np.exp(-r0 / .02), np.exp(-r1 / .02))
    del r0, r1
    loss = p.forward(*a)
    gu, gv, gw, gphi = p.backward()
    assert np.isfinite(loss) and loss > 0
    for g in (gu, gv, gw, gphi):
        assert bool(torch.isfinite(g).all())
    assert _norm((gu, gv, gw)) > 0


def _c4_problem(dtype):
    from opencurrent_amd import adjoint
    n = 128
    h = 1.0 / n
    p = adjoint.AdvectProjectProblem(n, n, n, h, h, h, dt=.25 * h, steps=20, dtype=dtype, checkpoint_every=5)
    x = (np.arange(n) + .5) * h
    r2 = (x[:, None, None] - .4) ** 2 + (x[None, :, None] - .5) ** 2 + (x[None, None, :] - .6) ** 2
    r2t = (x[:, None, None] - .6) ** 2 + (x[None, :, None] - .45) ** 2 + (x[None, None, :] - .5) ** 2
    p.set_scalars(np.exp(-r2 / .02), np.exp(-r2t / .02))
    return p


def test_adjoint_128_fp32_tracks_fp64_and_central_differences(gpu):
    p64, p32 = _c4_problem(np.float64), _c4_problem(np.float32)
    v = _velocity(p64, (1001, 1002, 1003))
    l64 = p64.forward(*v)
    l32 = p32.forward(*[x.float() for x in v])
    g64, g32 = p64.backward()[:3], p32.backward()[:3]
    assert np.isfinite(l64) and np.isfinite(l32)
    for g in g64 + g32:
        assert bool(g.isfinite().all())
    num = sum(float(((x.double() - y.double()) ** 2).sum()) for x, y in zip(g64, g32)) ** .5
    rel = num / _norm(g64)
    _log(f"loss fp64 {l64:.10e} fp32 {l32:.10e}; gradient rel. L2 difference {rel:.3e}")
    assert abs(l64 - l32) <= 1e-3 * l64
    assert rel <= 1e-2  # FP32 arithmetic and the FP32 projection's 1e-3 divergence tolerance
    eps = 1e-4
    for d in range(2):
        dv = _velocity(p64, (4001 + 3 * d, 4002 + 3 * d, 4003 + 3 * d))
        lp = p64.forward(*[a + eps * b for a, b in zip(v, dv)])
        lm = p64.forward(*[a - eps * b for a, b in zip(v, dv)])
        fd = (lp - lm) / (2 * eps)
        an = _dot(g64, dv)
        _log(f"direction {d}: central difference {fd:.12e}, adjoint {an:.12e}, rel {abs(fd - an) / abs(an):.2e}")
        assert abs(fd - an) <= 1e-6 * abs(an)
