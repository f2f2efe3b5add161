"""Gradients through advection + projection (config C4's model, no
reference counterpart): the hand-written adjoint kernels chained with
checkpointing and the projection's adjoint (one more pressure solve),
checked against central finite differences in FP64 on periodic and walled
domains, by the dot-product identity, against themselves without
checkpoints, across precisions, and through torch.autograd."""
import numpy as np
import pytest

from opencurrent_amd import _types as T
from opencurrent_amd._types import BC, BCSet
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


def walls(kind):
    """(flow_bc, phi_bc): a closed free-slip box with an insulated scalar, or
    Rayleigh-Benard walls (no-slip z, periodic x and y) with a Dirichlet
    scalar on z"""
    if kind == "closed":
        return BCSet.uniform(T.BC_FORCED_INFLOW_VARIABLE_SLIP), BCSet.uniform(T.BC_NEUMANN)
    flow, phi = BCSet.uniform(T.BC_PERIODIC), BCSet.uniform(T.BC_PERIODIC)
    for f in ("zpos", "zneg"):
        setattr(flow, f, BC(T.BC_FORCED_INFLOW_VARIABLE_SLIP, 0.0, 1.0))
        setattr(phi, f, BC(T.BC_DIRICHELET, 0.0, 0.0))
    return flow, phi


def problem(n=(32, 32, 32), steps=20, scheme=T.IT_SECOND_ORDER_CENTERED, dtype=np.float64, ce=5, bcs=None):
    h = tuple(1.0 / x for x in n)
    flow, phi = bcs if bcs is not None else (None, None)
    p = A().AdvectProjectProblem(*n, *h, dt=.25 * h[0], steps=steps, scheme=scheme, dtype=dtype,
                                 checkpoint_every=ce, flow_bc=flow, phi_bc=phi)
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


@pytest.mark.parametrize("kind", ["closed", "rb"])
def test_gradient_with_walls_matches_central_differences(gpu, kind):
    """as above (centred scheme, random flow, 8 directions, eps 1e-4, FP64,
    32^3, 1e-6) with walls: the projection's adjoint goes through the
    gradient and divergence adjoints and a Neumann pressure solve"""
    p = problem(bcs=walls(kind))
    v = velocity(p)
    p.forward(*v)
    g = p.backward()[:3]
    eps = 1e-4
    worst = 0.0
    for d in range(8):
        dv = velocity(p, seeds=(4101 + 3 * d, 4102 + 3 * d, 4103 + 3 * d))
        lp = p.forward(*[a + eps * b for a, b in zip(v, dv)])
        lm = p.forward(*[a - eps * b for a, b in zip(v, dv)])
        fd = (lp - lm) / (2 * eps)
        an = dot(g, dv)
        worst = max(worst, abs(fd - an) / abs(an))
    assert worst <= 1e-6, worst


@pytest.mark.parametrize("kind", ["periodic", "closed", "rb"])
def test_projection_adjoint_dot_product(gpu, kind):
    """<P a, b> = <a, P^T b> over the parameter faces, P^T from the adjoint
    kernels and one pressure solve, to the solves' tolerance"""
    import torch
    p = problem(n=(16, 16, 32), steps=1, bcs=None if kind == "periodic" else walls(kind))
    a, b = velocity(p, (5101, 5102, 5103)), velocity(p, (5104, 5105, 5106))
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
    lhs, rhs = dot(pa, b), dot(a, ptb)
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs), (lhs, rhs)


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


def test_backward_can_be_repeated(gpu):
    """backward replays the last forward's checkpoints from a kept seed, so
    a second call (retain_graph) gives the same gradient again"""
    import torch
    p = problem(n=(16, 16, 16), steps=6, ce=4)
    v = [x.clone().requires_grad_(True) for x in velocity(p)]
    loss = A().advect_project_loss(*v, p)
    loss.backward(retain_graph=True)
    first = [x.grad.clone() for x in v]
    loss.backward()
    for x, g in zip(v, first):
        assert torch.equal(x.grad, 2 * g)
    again = p.backward()[:3]
    for a, g in zip(again, first):
        assert torch.equal(a, g)


def test_two_losses_on_one_problem_are_refused(gpu):
    """a loss whose problem ran forward again holds no checkpoints of its own:
    differentiating it raises instead of returning another input's gradient"""
    p = problem(n=(16, 16, 16), steps=4, ce=2)
    v = [x.clone().requires_grad_(True) for x in velocity(p)]
    l1 = A().advect_project_loss(*v, p)
    l2 = A().advect_project_loss(*[2 * x for x in v], p)
    with pytest.raises(RuntimeError, match="ran forward again"):
        (l1 + l2).backward()
