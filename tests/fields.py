"""Seeded synthetic inputs shared by the tests (numpy, CPU).

Random numbers come from splitmix64 of (seed << 40) ^ interior_index, so a
field is the same on every machine and for every consumer."""
import math

import numpy as np

from opencurrent_amd._types import HostGrid

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x):
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & M64
    z = x
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & M64
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & M64
    return z ^ (z >> np.uint64(31))


def uniform(seed, shape):
    """U[0,1) array of `shape`, C order over the interior indices."""
    n = int(np.prod(shape))
    idx = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = splitmix64((np.uint64(seed) << np.uint64(40)) ^ idx)
    return ((z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))).reshape(shape)


def mac_grids(nx, ny, nz, dtype=np.float64):
    """u, v, w, scalar laid out like Eqn_IncompressibleNS3DParams::init_grids."""
    u = HostGrid(nx + 1, ny, nz, 1, 1, 1, 0, 1, 1, dtype=dtype)
    v = HostGrid(nx, ny + 1, nz, 1, 1, 1, 1, 0, 1, dtype=dtype)
    w = HostGrid(nx, ny, nz + 1, 1, 1, 1, 1, 1, 0, dtype=dtype)
    t = HostGrid(nx, ny, nz, 1, 1, 1, 1, 1, 1, dtype=dtype)
    return u, v, w, t


def random_mac_velocity(nx, ny, nz, seeds=(1001, 1002, 1003), dtype=np.float64):
    u, v, w, t = mac_grids(nx, ny, nz, dtype)
    u.interior()[:] = 2 * uniform(seeds[0], u.interior().shape) - 1
    v.interior()[:] = 2 * uniform(seeds[1], v.interior().shape) - 1
    w.interior()[:] = 2 * uniform(seeds[2], w.interior().shape) - 1
    return u, v, w, t


def divergence_rhs(nx, ny, nz, h, dtype=np.float64, seeds=(1001, 1002, 1003), remove_mean=True):
    """b = div(u) of a seeded U[-1,1] MAC velocity (normal boundary faces
    random too), mean removed in FP64, on a grid with the rhs layout
    (nx,ny,nz) ghosts 1, no padding."""
    u = 2 * uniform(seeds[0], (nx + 1, ny, nz)) - 1
    v = 2 * uniform(seeds[1], (nx, ny + 1, nz)) - 1
    w = 2 * uniform(seeds[2], (nx, ny, nz + 1)) - 1
    d = (u[1:] - u[:-1]) / h + (v[:, 1:] - v[:, :-1]) / h + (w[:, :, 1:] - w[:, :, :-1]) / h
    if remove_mean:
        d = d - d.mean()
    b = HostGrid(nx, ny, nz, dtype=dtype)
    b.interior()[:] = d
    return b


def trig_rhs(nx, ny, nz, hx, hy, hz, axis=-1, dtype=np.float64):
    """RHS of the reference's multigrid tests (mgdoubletest.cpp:43-95):
    cos(x) sin(y) cos(z) at x = (i - .5) 2 pi hx, mean removed."""
    pi = math.acos(-1.0)
    k = np.arange(nz); j = np.arange(ny); i = np.arange(nx)
    z = (k - .5) * 2 * pi * hz
    y = (j - .5) * 2 * pi * hy
    x = (i - .5) * 2 * pi * hx
    if axis in (0, 1):
        z = np.zeros_like(z)
    if axis in (0, 2):
        y = np.full_like(y, pi / 2)
    if axis in (1, 2):
        x = np.zeros_like(x)
    f = np.cos(x)[:, None, None] * np.sin(y)[None, :, None] * np.cos(z)[None, None, :]
    f = f - f.sum() / (nx * ny * nz)
    b = HostGrid(nx, ny, nz, dtype=dtype)
    b.interior()[:] = f
    return b
