"""Critical Rayleigh number of Rayleigh-Benard convection, measured the way
the reference's own tests measure it: tests/rayleigh.cpp:60-300 (free-slip
walls, periodic third axis) and tests/rayleighnoslip.cpp:69-336 (no-slip
top and bottom, free-slip sides).

For each Rayleigh number the FP64 Eqn_IncompressibleNS3D is started from the
conductive temperature profile plus 1e-4 velocity noise and advanced in
calls of advance(0.01) until the growth rates of the two velocity components
in the convection plane agree; Ra_cr is the zero crossing of that rate,
interpolated between the last decaying and the first growing Ra.

The reference's results at each resolution are the only numeric answers its
test suite holds for the NS step (rayleigh.cpp:338-341,
rayleighnoslip.cpp:343-345); they depend on the advection, diffusion,
thermal force, time-step rule and projection together.

The noise is this build's seeded generator, not the reference's unseeded
rand(): the measured rate is the dominant eigenvalue once the other modes
have decayed, so the start only changes the result by the stopping rule's
resolution (the two rates agree within 0.5%, ~5e-3 in Ra).
"""
import ctypes as C
import math

import numpy as np

from opencurrent_amd import _types as T
from opencurrent_amd._types import BC, BCSet, HostGrid, dtype_code
from fields import mac_grids, uniform

# rayleigh.cpp:338-341 and :344; rayleighnoslip.cpp:343-345 and :347
RA_FREE_SLIP = {16: 659.693989, 32: 658.054503, 64: 657.650149, 128: 657.541609}
RA_FREE_SLIP_EXACT = 27 * math.pi ** 4 / 4
RA_NO_SLIP = {16: 1674.296911, 32: 1699.253950, 64: 1705.590663}
RA_NO_SLIP_LITERATURE = 1707.76

_AXIS = {T.DIR_XPOS: 0, T.DIR_XNEG: 0, T.DIR_YPOS: 1, T.DIR_YNEG: 1, T.DIR_ZPOS: 2, T.DIR_ZNEG: 2}
_FACES = (("xpos", "xneg"), ("ypos", "yneg"), ("zpos", "zneg"))


def _f32(x):
    return np.float32(x)


def setup(res, ra, vertical, no_slip, seed=7001):
    """(NSParams, (u, v, w, temp) host grids) of rayleigh.cpp:64-212 /
    rayleighnoslip.cpp:69-208.  Domain sizes and spacings are floats there,
    and so here."""
    from opencurrent_amd._lib import NSParams
    vax = _AXIS[vertical]
    side = _f32(math.pi / 3.11) if no_slip else _f32(math.sqrt(2.0))
    if vax == 0:
        n = (res, res, res // 2)
        dom = (_f32(1.0), side, None)
        dom = (dom[0], dom[1], _f32(float(dom[1] if no_slip else dom[0]) * .5))
    elif vax == 1:
        n = (res, res, res // 2)
        dom = (side, _f32(1.0), _f32(float(side) * .5))
    else:
        n = (res, res // 2, res)
        dom = (side, _f32(float(side) * .5), _f32(1.0))
    h = tuple(_f32(d / _f32(k)) for d, k in zip(dom, n))

    p = NSParams()
    p.nx, p.ny, p.nz = n
    p.hx, p.hy, p.hz = (float(x) for x in h)
    p.max_divergence = 1e-12
    p.viscosity = 1.0
    p.thermal_diffusion = 1.0
    p.gravity = -1.0
    p.bouyancy = 1.0  # with all coefficients 1, Ra = the temperature difference
    p.vertical_direction = vertical
    p.advection_scheme = T.IT_FIRST_ORDER_UPWIND
    p.time_step = T.TS_FORWARD_EULER
    p.cfl_factor = .99

    flow = BCSet.uniform(T.BC_FORCED_INFLOW_VARIABLE_SLIP)  # closed, free-slip (aux 0)
    temp = BCSet.uniform(T.BC_NEUMANN)
    pos, neg = _FACES[vax]
    setattr(temp, pos, BC(T.BC_DIRICHELET, 0.0, 0.0))
    setattr(temp, neg, BC(T.BC_DIRICHELET, 0.0, 0.0))
    if no_slip:
        setattr(flow, pos, BC(T.BC_FORCED_INFLOW_VARIABLE_SLIP, 0.0, 1.0))
        setattr(flow, neg, BC(T.BC_FORCED_INFLOW_VARIABLE_SLIP, 0.0, 1.0))
    else:
        pax = 2 if vax in (0, 1) else 1  # the periodic axis
        for f in _FACES[pax]:
            setattr(temp, f, BC(T.BC_PERIODIC, 0.0, 0.0))
            setattr(flow, f, BC(T.BC_PERIODIC, 0.0, 0.0))
    hot = pos if vertical in (T.DIR_XNEG, T.DIR_YNEG, T.DIR_ZNEG) else neg
    setattr(temp, hot, BC(T.BC_DIRICHELET, float(ra), 0.0))
    p.flow_bc, p.temp_bc = flow, temp

    u, v, w, t = mac_grids(*n)
    c = (np.arange(n[vax]) + .5) * float(h[vax]) / float(dom[vax])
    y = (c if vertical in (T.DIR_XNEG, T.DIR_YNEG, T.DIR_ZNEG) else 1 - c).astype(np.float32).astype(np.float64)
    shape = [1, 1, 1]
    shape[vax] = n[vax]
    t.interior()[:] = (y * float(ra)).reshape(shape)
    for q, g in enumerate((u, v, w)):  # cells i<nx, j<ny, k<nz of each component
        g.interior()[:n[0], :n[1], :n[2]] = -1e-4 + 2e-4 * uniform(seed + q, n)
    return p, (u, v, w, t)


def _rates_pair(vertical):
    """(index of the growth-rate component, index of the check component) as
    rayleigh.cpp:245-258 pairs them"""
    vax = _AXIS[vertical]
    return {0: (1, 0), 1: (0, 1), 2: (0, 2)}[vax]


class OracleRun:
    """the oracle's Eqn_IncompressibleNS3D restatement"""

    def __init__(self, p, grids):
        import oracle_lib as O
        self.O = O
        u, v, w, t = grids
        self.h = O.lib().orc_ns_create(dtype_code(np.float64), C.byref(p), u.ptr(), v.ptr(), w.ptr(), t.ptr())
        assert self.h, "orc_ns_create"
        self.buf = [HostGrid.congruent(g) for g in (u, v, w)]

    def advance(self, dt):
        return self.O.lib().orc_ns_advance(self.h, dt) == 1

    def maxabs(self):
        out = []
        for q, g in enumerate(self.buf):
            assert self.O.lib().orc_ns_get(self.h, q, g.ptr()) == 1
            out.append(float(np.abs(g.interior()).max()))
        return out

    def close(self):
        if self.h:
            self.O.lib().orc_ns_destroy(self.h)
            self.h = None


class DeviceRun:
    """Eqn_IncompressibleNS3DDeviceD through libocu.so"""

    def __init__(self, p, grids):
        from opencurrent_amd.solvers import IncompressibleNS3D
        self.ns = IncompressibleNS3D(p, *grids)

    def advance(self, dt):
        return self.ns.advance(dt)

    def maxabs(self):
        return [float(np.abs(self.ns.field(q).interior()).max()) for q in range(3)]

    def close(self):
        self.ns = None


def growth_rate(run, vertical, no_slip, max_count=1000):
    """log growth of the lateral velocity per advance(0.01), or None
    (rayleigh.cpp:214-300; rayleighnoslip.cpp:210-306 needs five agreeing
    rates in all, counts to 5000 and keeps going past 0.1)"""
    a, b = _rates_pair(vertical)
    last = [0.0, 0.0, 0.0]
    count, under = 1, 0
    limit = 5000 if no_slip else max_count
    while True:
        count += 1
        assert run.advance(.01), "advance failed"
        m = run.maxabs()
        with np.errstate(divide="ignore", invalid="ignore"):
            ra = math.log(m[a] / last[a]) if last[a] > 0 else math.inf
            rb = math.log(m[b] / last[b]) if last[b] > 0 else math.inf
        agree = math.isfinite(ra) and ra != 0 and abs((ra - rb) / ra) < .005
        if agree:
            under += 1
            if count > 50 and (under > 5 if no_slip else True):
                return .5 * (ra + rb), count
        elif not no_slip:
            under = 0
        last = m
        if m[a] < (1e-12 if no_slip else 1e-10):
            return None, count
        if m[a] > .1 and not no_slip:
            return None, count
        if count > limit:
            return None, count


def critical_ra(make_run, res, vertical, no_slip, start, end, step=1.0, log=print):
    """rayleigh.cpp:302-324 / rayleighnoslip.cpp:310-336: walk Ra upward until
    the rate turns positive; interpolate the zero crossing"""
    lb = ub = None
    ra = start
    while ra <= end:
        run = make_run(*setup(res, ra, vertical, no_slip))
        try:
            rate, count = growth_rate(run, vertical, no_slip)
        finally:
            run.close()
        log(f"res {res} dir {vertical} Ra {ra:g}: rate {rate} after {count} advances")
        assert rate is not None, f"no growth rate at Ra {ra}"
        if rate < 0:
            lb = (ra, rate)
        else:
            ub = (ra, rate)
            break
        ra += step
    assert lb is not None and ub is not None, "Ra range does not bracket the crossing"
    return lb[0] - lb[1] / (ub[1] - lb[1])
