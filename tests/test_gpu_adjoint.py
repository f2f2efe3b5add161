"""Gradients through advection + projection (config C4's model, no
reference counterpart): the hand-written adjoint kernels chained with
checkpointing and the self-adjoint projection, checked against central
finite differences in FP64, against themselves without checkpoints, across
precisions, and through torch.autograd."""
import numpy as np
import pytest

from opencurrent_amd import _types as T
from fields import uniform

pytestmark = pytest.mark.gpu


def A():
    from opencurrent_amd import adjoint
    return adjoint


def blob(n, h, centre, sigma):
    x = (np.arange(n[0]) + .5) * h[0]
    y = (np.arange(n[1]) + .5) * h[1]
    z = (np.arange(n[2]) + .5) * h[2]
    r2 = (x[:, None, None] - centre[0]) ** 2 + (y[None, :, None] - centre[1]) ** 2 + \
         (z[None, None, :] - centre[2]) ** 2
    return np.exp(-r2 / (2 * sigma ** 2))


def problem(n=(32, 32, 32), steps=20, scheme=T.IT_SECOND_ORDER_CENTERED, dtype=np.float64, ce=5):
    h = tuple(1.0 / x for x in n)
    p = A().AdvectProjectProblem(*n, *h, dt=.25 * h[0], steps=steps, scheme=scheme, dtype=dtype,
                                 checkpoint_every=ce)
    c0 = .3 + .4 * uniform(3002, (3,))
    c1 = .3 + .4 * uniform(3001, (3,))
    p.set_scalars(blob(n, h, c0, .1), blob(n, h, c1, .1))
    return p


def velocity(p, seeds=(1001, 1002, 1003), mean=(0.0, 0.0, 0.0), amp=1.0):
    import torch
    out = []
    for q, g in enumerate((p.u, p.v, p.w)):
        x = torch.as_tensor(mean[q] + amp * (2 * uniform(seeds[q], tuple(g.tensor.shape)) - 1),
                            dtype=g.tensor.dtype, device=g.tensor.device)
        x[~p.masks[q]] = 0
        out.append(x)
    return out


def dot(a, b):
    return sum(float((x.double() * y.double()).sum()) for x, y in zip(a, b))


@pytest.mark.parametrize("scheme,flow", [(T.IT_SECOND_ORDER_CENTERED, "random"), (T.IT_FIRST_ORDER_UPWIND, "drift")])
def test_gradient_matches_central_differences(gpu, scheme, flow):
    """8 seeded directions, eps = 1e-4, FP64, 32^3, relative error <= 1e-6.
    The centred scheme is smooth in the velocity.  The upwind flux has a kink
    wherever a face velocity changes sign, so its check uses a uniform drift
    (1, .8, .6) plus U[-.1,.1] noise: no face crosses zero within eps."""
    p = problem(scheme=scheme)
    v = velocity(p) if flow == "random" else velocity(p, mean=(1.0, .8, .6), amp=.1)
    tol = 1e-6
    p.forward(*v)
    g = p.backward()[:3]
    eps = 1e-4
    worst = 0.0
    for d in range(8):
        dv = velocity(p, seeds=(4001 + 3 * d, 4002 + 3 * d, 4003 + 3 * d))
        lp = p.forward(*[a + eps * b for a, b in zip(v, dv)])
        lm = p.forward(*[a - eps * b for a, b in zip(v, dv)])
        fd = (lp - lm) / (2 * eps)
        an = dot(g, dv)
        worst = max(worst, abs(fd - an) / abs(an))
    assert worst <= tol, worst


def test_checkpointing_gives_the_same_gradient(gpu):
    a = problem(ce=5)
    b = problem(ce=20)
    v = velocity(a)
    la, lb = a.forward(*v), b.forward(*v)
    assert la == lb
    ga, gb = a.backward(), b.backward()
    for x, y in zip(ga, gb):
        assert bool((x == y).all())


def test_projection_is_self_adjoint(gpu):
    """<P a, b> = <a, P b> on the periodic MAC grid (to the solve tolerance)"""
    import torch
    p = problem(steps=1)
    a, b = velocity(p, (5001, 5002, 5003)), velocity(p, (5004, 5005, 5006))

    def proj(x):
        for g, y, m in zip((p.u, p.v, p.w), x, p.masks):
            g.tensor.zero_()
            g.tensor[m] = y[m]
        p._project(p.proj)
        return [torch.where(m, g.tensor, torch.zeros_like(g.tensor)) for g, m in zip((p.u, p.v, p.w), p.masks)]

    pa, pb = proj(a), proj(b)
    lhs, rhs = dot(pa, b), dot(a, pb)
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    # idempotent
    ppa = proj(pa)
    assert max(float((x - y).abs().max()) for x, y in zip(ppa, pa)) <= 1e-9


def test_fp32_gradient_tracks_fp64(gpu):
    p64, p32 = problem(dtype=np.float64), problem(dtype=np.float32)
    v = velocity(p64)
    p64.forward(*v)
    p32.forward(*[x.float() for x in v])
    g64, g32 = p64.backward()[:3], p32.backward()[:3]
    num = np.sqrt(sum(float(((x.double() - y.double()) ** 2).sum()) for x, y in zip(g64, g32)))
    den = np.sqrt(sum(float((x.double() ** 2).sum()) for x in g64))
    # FP32 arithmetic plus the FP32 projection's 1e-3 divergence tolerance
    assert num / den <= 1e-2, num / den


def test_torch_autograd_function(gpu):
    import torch
    p = problem(n=(16, 16, 16), steps=6, ce=4)
    v = [x.clone().requires_grad_(True) for x in velocity(p)]
    loss = A().advect_project_loss(*v, p)
    loss.backward()
    p2 = problem(n=(16, 16, 16), steps=6, ce=4)
    l2 = p2.forward(*[x.detach() for x in v])
    g2 = p2.backward()[:3]
    assert float(loss.detach()) == l2
    for x, y in zip(v, g2):
        assert torch.equal(x.grad, y)
