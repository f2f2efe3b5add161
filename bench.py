#!/usr/bin/env python3
"""Benchmark of the projection-method hot path on one MI355X.

Workload (one "step"): the 256^3 all-Neumann Poisson problem on the unit cube
(h = 1/256) whose right-hand side is the divergence of a seeded U[-1,1] MAC
velocity with its mean removed, solved from a zero guess to a relative
residual ||b - Ax||_2 / ||b||_2 <= 1e-10 by FP32 V-cycles inside FP64
iterative refinement (Sol_MultigridPressureMixed3DDeviceD).  Inputs are
resident in HBM before the timed region.

Printed (one JSON line): the solve throughput in Mcells/s (cells solved to
1e-10 per second), the FP64 multigrid solve of the same problem to the same
target for comparison, the error of the mixed solution against the FP64
one, the roofline of the dominant kernel (the level-0 FP32 fused red+black
sweep, timed with HIP events on the solver's stream) and a CPU baseline (the
oracle's OpenMP FP64 multigrid on the host cores).
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "MG time to 1e-10 rel. residual + NS step Mcells/s, FP32-refined vs FP64"
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the level-0 fused sweeps at 256^3 (plain, from the
# zero grid, adding the prolongation), measured by two rocprofv3 counter
# passes (FETCH_SIZE, WRITE_SIZE) and scaled by what the same counters read
# for a plain copy of known size; recorded with the profile summary
# profiles/level0_kernels_256_hbm_traffic.md
SWEEP_TRAFFIC_256 = {"plain": 241_500_000, "zero": 152_200_000, "pro": 254_500_000}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--rel-tol", type=float, default=1e-10)
    ap.add_argument("--cycles", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison solve")
    ap.add_argument("--kernel-reps", type=int, default=50)
    ap.add_argument("--no-extra", action="store_true", help="skip the C1/C3/C4/C5 side measurements")
    ap.add_argument("--only", choices=["c1", "c3", "c4", "c5"],
                    help="run only this side measurement and print its JSON (for profiling it alone)")
    return ap.parse_args()


def rhs(n):
    import fields
    return fields.divergence_rhs(n, n, n, 1.0 / n)


class Events:
    def __init__(self, hip, stream):
        self.hip, self.stream = hip, stream
        self.a, self.b = C.c_void_p(), C.c_void_p()
        hip.ocu_event_create(C.byref(self.a))
        hip.ocu_event_create(C.byref(self.b))

    def start(self):
        self.hip.ocu_event_record(self.a, self.stream)

    def stop_ms(self):
        self.hip.ocu_event_record(self.b, self.stream)
        self.hip.ocu_event_synchronize(self.b)
        ms = C.c_float()
        self.hip.ocu_event_elapsed_ms(C.byref(ms), self.a, self.b)
        return ms.value


def time_fused_sweeps(dims, reps, params, dtype=np.float32, variants=("plain", "zero", "pro")):
    """Average duration of the level-0 fused red+black sweep on a grid of
    `dims` cells as the solver launches it - writing the output's ghost layer
    too - back to back on the solver's stream, timed with HIP events, for
    each kind the solver runs: plain, starting from the deferred zero grid
    (u not read), and adding the pending prolongation of the coarser level
    while it reads u.  `params` are the level's relaxation parameters as the
    solver computes them.  Returns {kind: (ms per launch, algorithmic bytes
    per launch)}."""
    from opencurrent_amd import _lib
    from opencurrent_amd import _types as T
    from opencurrent_amd.device import DeviceGrid
    hip = _lib.hip()
    stream = _lib.host().ocu_get_compute_stream()
    bc = T.BCSet.uniform(T.BC_NEUMANN)
    nx, ny, nz = dims
    u = DeviceGrid(nx, ny, nz, dtype=dtype)
    o = DeviceGrid(nx, ny, nz, dtype=dtype)
    b = DeviceGrid(nx, ny, nz, dtype=dtype)
    u.tensor.uniform_(-1, 1)
    b.tensor.uniform_(-1, 1)
    uc = DeviceGrid(nx // 2, ny // 2, nz // 2, dtype=dtype)
    uc.tensor.uniform_(-1, 1)
    import torch
    torch.cuda.synchronize()
    h = (params.h, params.h, params.h)
    cells = nx * ny * nz
    w = np.dtype(dtype).itemsize
    launches = {
        # read u and b, write u, once per cell
        "plain": (lambda: hip.ocu_mg_relax_fused_bc(u.ptr(), o.ptr(), b.ptr(), C.byref(params), C.byref(bc), *h, 0,
                                                    T.FUSED_WRITE_GHOSTS, stream), 3 * w * cells),
        # u is the zero grid and is not read: read b, write u
        "zero": (lambda: hip.ocu_mg_relax_fused_bc(u.ptr(), o.ptr(), b.ptr(), C.byref(params), C.byref(bc), *h, 0,
                                                   T.FUSED_WRITE_GHOSTS | T.FUSED_ZERO_INPUT, stream), 2 * w * cells),
        # + the coarse grid, read once
        "pro": (lambda: hip.ocu_mg_relax_fused_prolong_bc(u.ptr(), o.ptr(), b.ptr(), uc.ptr(), 1, C.byref(params),
                                                          C.byref(bc), *h, 0, T.FUSED_WRITE_GHOSTS, stream),
                3 * w * cells + w * (cells // 8)),
    }
    if "pro" in variants and not hip.ocu_mg_relax_fused_prolong_ok(u.ptr(), uc.ptr(), C.byref(bc)):
        variants = tuple(v for v in variants if v != "pro")
    # a plain copy o = 1*u over the whole allocation (k_lincomb_flat: 4 B
    # read and 4 B written per element): the reference point that scales
    # the HBM counters
    for _ in range(3):
        _lib.check(hip.ocu_grid_lincomb(o.ptr(), 1.0, u.ptr(), 0.0, None, stream), "calibration copy")
    out = {}
    ev = Events(hip, stream)
    for kind in variants:
        fn, nbytes = launches[kind]
        for _ in range(5):
            _lib.check(fn(), f"fused sweep ({kind})")
        ev.start()
        for _ in range(reps):
            fn()
        out[kind] = (ev.stop_ms() / reps, nbytes)
    return out


def time_fp64_mg_64(hip, stream):
    """config C1: plain FP64 multigrid on the 64^3 Neumann problem to 1e-8
    relative residual (the reference algorithm, FMG)"""
    from opencurrent_amd import _types as T
    from opencurrent_amd import device, solvers
    n = 64
    b = rhs(n)
    rms = float(np.sqrt((b.interior() ** 2).mean()))
    gb = device.DeviceGrid.from_host(b)
    mg = solvers.MultigridPressure3D(gb, 1 / n, 1 / n, 1 / n, T.BCSet.uniform(T.BC_NEUMANN))
    pd = mg.pressure_desc()
    ev = Events(hip, stream)
    times = []
    for rep in range(6):
        hip.ocu_grid_clear(C.byref(pd), 0.0, stream)
        ev.start()
        ok, res = mg.solve(1e-8 * rms, 20)
        ms = ev.stop_ms()
        if not ok:
            raise RuntimeError("C1 solve did not converge")
        if rep:
            times.append(ms)
    ms = float(np.median(times))
    return {"workload": "FP64 MG 64^3 Neumann to 1e-8 rel. residual", "ms": round(ms, 4),
            "mcells_s": round(n ** 3 / (ms / 1e3) / 1e6, 2), "vcycles": mg.stats()["vcycles"]}


def time_ns_step(steps=100, warmup=2, n=128):
    """config C3: Boussinesq Rayleigh-Benard at n^3 (periodic x/y, no-slip z,
    Ra = 1e5, Pr = 0.71, second-order upwind advection, explicit diffusion,
    CFL 0.5, forward Euler), 100 steps, FP64 with the FP64 multigrid
    projection and FP64 with the FP32-refined projection"""
    import fields
    from opencurrent_amd import _types as T
    from opencurrent_amd import solvers
    from opencurrent_amd._lib import NSParams
    out = {}
    for name, mixed in (("fp64", 0), ("fp32_refined", 1)):
        p = NSParams()
        p.nx = p.ny = p.nz = n
        p.hx = p.hy = p.hz = 1.0 / n
        p.max_divergence = 1e-8
        tb = T.BCSet.uniform(T.BC_PERIODIC)
        tb.zneg = T.BC(T.BC_DIRICHELET, 1.0, 0.0)
        tb.zpos = T.BC(T.BC_DIRICHELET, 0.0, 0.0)
        p.temp_bc = tb
        fb = T.BCSet.uniform(T.BC_PERIODIC)
        fb.zneg = T.BC(T.BC_FORCED_INFLOW_VARIABLE_SLIP, 0.0, 1.0)
        fb.zpos = T.BC(T.BC_FORCED_INFLOW_VARIABLE_SLIP, 0.0, 1.0)
        p.flow_bc = fb
        p.vertical_direction = T.DIR_ZPOS
        p.gravity, p.bouyancy, p.viscosity, p.thermal_diffusion = -1.0, 1e5 * .71, .71, 1.0
        p.advection_scheme, p.time_step, p.cfl_factor = T.IT_SECOND_ORDER_UPWIND, T.TS_FORWARD_EULER, .5
        p.mixed_projection = mixed
        u, v, w, t = fields.mac_grids(n, n, n)
        z = (np.arange(n) + .5) / n
        t.interior()[:] = (1 - z)[None, None, :] + 1e-3 * (2 * fields.uniform(2001, t.interior().shape) - 1)
        if not mixed:
            p_fp64, init = p, (u, v, w, t)
        ns = solvers.IncompressibleNS3D(p, u, v, w, t)
        for _ in range(warmup):
            ns.advance_one_step(ns.get_max_stable_timestep())
        import torch
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            if not ns.advance_one_step(ns.get_max_stable_timestep()):
                raise RuntimeError(f"NS step failed ({name})")
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        out[name] = {"ms_per_step": round(dt * 1e3, 3), "mcells_s": round(n ** 3 / dt / 1e6, 2)}
        del ns
    out["workload"] = (f"Rayleigh-Benard {n}^3, Ra 1e5, Pr 0.71, second-order upwind, CFL 0.5, FE, "
                       f"projection to 1e-8 Linf, {steps} steps")
    out["speedup_fp32_refined"] = round(out["fp64"]["ms_per_step"] / out["fp32_refined"]["ms_per_step"], 3)
    out["cpu_oracle_fp64"] = cpu_ns_step(n, p_fp64, init)
    return out


def cpu_ns_step(n, p, init, steps=3):
    """the oracle's FP64 NS step (the reference's step restated, OpenMP) on
    the same problem, on the host cores: a few steps after one warm-up"""
    import oracle_lib as O
    from opencurrent_amd._types import dtype_code
    u, v, w, t = init
    cores = O.lib().orc_set_threads(0)
    h = O.lib().orc_ns_create(dtype_code(np.float64), C.byref(p), u.ptr(), v.ptr(), w.ptr(), t.ptr())
    if not h:
        return None
    try:
        O.lib().orc_ns_advance_one_step(h, O.lib().orc_ns_max_stable_timestep(h))
        t0 = time.perf_counter()
        for _ in range(steps):
            O.lib().orc_ns_advance_one_step(h, O.lib().orc_ns_max_stable_timestep(h))
        dt = (time.perf_counter() - t0) / steps
    finally:
        O.lib().orc_ns_destroy(h)
    return {"ms_per_step": round(dt * 1e3, 1), "mcells_s": round(n ** 3 / dt / 1e6, 2),
            "cores": cores, "kind": "port",
            "sample": f"{steps} steps after one warm-up"}


def time_adjoint(n=128, steps=20):
    """config C4: 20 passive-advection steps after a projection, loss vs a
    target blob, gradient w.r.t. the initial velocity with checkpoints every
    5 steps (forward + backward, FP32)"""
    import torch
    import fields
    from opencurrent_amd import adjoint
    h = 1.0 / n
    p = adjoint.AdvectProjectProblem(n, n, n, h, h, h, dt=.25 * h, steps=steps, dtype=np.float32,
                                     checkpoint_every=5)
    x = (np.arange(n) + .5) * h
    r2 = (x[:, None, None] - .4) ** 2 + (x[None, :, None] - .5) ** 2 + (x[None, None, :] - .6) ** 2
    r2t = (x[:, None, None] - .6) ** 2 + (x[None, :, None] - .45) ** 2 + (x[None, None, :] - .5) ** 2
    p.set_scalars(np.exp(-r2 / .02), np.exp(-r2t / .02))
    vel = []
    for q, g in enumerate((p.u, p.v, p.w)):
        t = torch.as_tensor(2 * fields.uniform(1001 + q, tuple(g.tensor.shape)) - 1, dtype=g.tensor.dtype,
                            device=g.tensor.device)
        t[~p.masks[q]] = 0
        vel.append(t)
    p.forward(*vel)
    p.backward()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss = p.forward(*vel)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    p.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    return {"workload": f"passive advection {n}^3 x {steps} steps + projection, gradient w.r.t. initial velocity, "
                        f"checkpoints every 5, FP32",
            "forward_ms": round((t1 - t0) * 1e3, 2), "backward_ms": round((t2 - t1) * 1e3, 2), "loss": loss}


def _device_used():
    import torch
    free, total = torch.cuda.mem_get_info()
    return total - free


def time_lockex(n=(512, 512, 256), steps=10):
    """config C5: lock exchange in a closed no-slip box of 512 x 512 x 256
    cells (hx = hy = 1, hz = 0.5, so the multigrid runs with (fx,fy,fz) =
    (.25,.25,1)), heavy fluid for x < L/2, gravity along y, nu = kappa = 0.1;
    FP64 equations projected by FP32 V-cycles in FP64 refinement, dt =
    min(stable step, 0.1).  Then the advection-projection gradient model on
    the same grid and walls: 10 steps forward and backward, FP64.  Reports
    the device memory each holds (measured) against the 288 GB of HBM."""
    import torch
    import fields
    from opencurrent_amd import _types as T
    from opencurrent_amd import adjoint, solvers
    from opencurrent_amd._lib import NSParams
    nx, ny, nz = n
    p = NSParams()
    p.nx, p.ny, p.nz = n
    p.hx, p.hy, p.hz = 1.0, 1.0, .5
    p.max_divergence = 1e-8
    walls = T.BCSet.uniform(T.BC_FORCED_INFLOW_VARIABLE_SLIP)
    for f in ("xpos", "xneg", "ypos", "yneg", "zpos", "zneg"):
        setattr(walls, f, T.BC(T.BC_FORCED_INFLOW_VARIABLE_SLIP, 0.0, 1.0))
    p.flow_bc = walls
    p.temp_bc = T.BCSet.uniform(T.BC_NEUMANN)
    p.vertical_direction = T.DIR_YPOS
    p.gravity, p.bouyancy, p.viscosity, p.thermal_diffusion = -1.0, 1.0, .1, .1
    p.advection_scheme, p.time_step, p.cfl_factor = T.IT_FIRST_ORDER_UPWIND, T.TS_FORWARD_EULER, 1.0
    p.mixed_projection = 1
    u, v, w, t = fields.mac_grids(nx, ny, nz)
    noise = 1e-3 * (2 * fields.uniform(5001, (ny, nz)) - 1)
    t.interior()[:] = np.where(np.arange(nx) < nx // 2, -1.0, 1.0)[:, None, None]
    t.interior()[nx // 2 - 1:nx // 2 + 1] += noise[None]
    torch.cuda.synchronize()
    base = _device_used()
    ns = solvers.IncompressibleNS3D(p, u, v, w, t)
    del u, v, w, t
    torch.cuda.synchronize()
    ns_bytes = _device_used() - base
    ns.advance_one_step(min(ns.get_max_stable_timestep(), .1))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        if not ns.advance_one_step(min(ns.get_max_stable_timestep(), .1)):
            raise RuntimeError("lock-exchange step failed")
    torch.cuda.synchronize()
    ns_s = (time.perf_counter() - t0) / steps
    del ns
    torch.cuda.empty_cache()

    flow = walls
    base = _device_used()
    ap = adjoint.AdvectProjectProblem(nx, ny, nz, 1.0, 1.0, .5, dt=.25, steps=steps, dtype=np.float64,
                                      checkpoint_every=5, flow_bc=flow, phi_bc=T.BCSet.uniform(T.BC_NEUMANN))
    x = (np.arange(nx) + .5) / nx
    y = (np.arange(ny) + .5) / ny
    z = (np.arange(nz) + .5) / nz
    ap.set_scalars(np.tanh(8 * (x - .5))[:, None, None] * np.ones((1, ny, nz)),
                   np.tanh(8 * (x - .45))[:, None, None] + .1 * np.sin(2 * np.pi * y)[None, :, None]
                   * np.cos(np.pi * z)[None, None, :])
    vel = []
    for q, g in enumerate((ap.u, ap.v, ap.w)):
        tt = torch.as_tensor(2 * fields.uniform(1001 + q, tuple(g.tensor.shape)) - 1, dtype=g.tensor.dtype,
                             device=g.tensor.device)
        tt[~ap.masks[q]] = 0
        vel.append(tt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ap.forward(*vel)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ap.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    adj_bytes = _device_used() - base
    del ap, vel
    torch.cuda.empty_cache()
    cells = nx * ny * nz
    return {"workload": f"lock exchange {nx}x{ny}x{nz}, hx=hy=1, hz=0.5, no-slip walls, FP64 NS with the "
                        f"FP32-refined projection to 1e-8 Linf; advection-projection gradient over {steps} steps, "
                        f"FP64, checkpoints every 5",
            "ns_ms_per_step": round(ns_s * 1e3, 2), "ns_mcells_s": round(cells / ns_s / 1e6, 1),
            "ns_device_gb": round(ns_bytes / 1e9, 2),
            "adjoint_forward_ms": round((t1 - t0) * 1e3, 1), "adjoint_backward_ms": round((t2 - t1) * 1e3, 1),
            "adjoint_device_gb": round(adj_bytes / 1e9, 2), "hbm_gb": 288}


def host_cpu():
    """nproc, the cores this process may run on, and the CPU model (lscpu)"""
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    model = None
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except Exception:
        pass
    return {"nproc": os.cpu_count(), "usable": usable, "model": model}


def cpu_baseline(n, rel_tol, budget_s=25.0):
    """The oracle's FP64 multigrid (reference FMG schedule, OpenMP over
    i-planes) on the host cores, same problem and target, repeated while the
    budget allows.  Threads: every core this process may use, capped by
    OMP_NUM_THREADS when the environment sets it (the GPU host allots each
    GPU's jobs a share of its cores and sets it to that share)."""
    import oracle_lib as O
    from opencurrent_amd import _types as T
    cpu = host_cpu()
    cores = cpu["usable"]
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        cores = min(cores, int(os.environ["OMP_NUM_THREADS"]))
    cores = O.lib().orc_set_threads(cores)
    b = rhs(n)
    rms = float(np.sqrt((b.interior() ** 2).mean()))
    cfg = T.MGConfig.make(T.BCSet.uniform(T.BC_NEUMANN))
    times = []
    t_all = time.perf_counter()
    while True:
        u = T.HostGrid.congruent(b)
        t0 = time.perf_counter()
        ok, res, vc, _ = O.mg_solve(u, b, (1.0 / n,) * 3, cfg, rel_tol * rms, 40)
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_all > budget_s * 0.5 or len(times) >= 5:
            break
    t = float(np.median(times))
    return {"value": round(n ** 3 / t / 1e6, 3), "unit": "Mcells/s", "cores": cores,
            "kind": "port", "host": cpu,
            "sample": f"oracle FP64 FMG (reference schedule, OpenMP, {cores} threads) on the same {n}^3 problem to "
                      f"{rel_tol:g} rel. residual, {len(times)} solve(s), median {t:.2f} s, {vc} V-cycles"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (the multi-GPU sharding is a "
                                   "design note, not built)"}))
        sys.exit(2)
    import torch
    from opencurrent_amd import _lib
    from opencurrent_amd import _types as T
    from opencurrent_amd import device, solvers
    device.require_gpu()
    hip = _lib.hip()
    stream = _lib.host().ocu_get_compute_stream()
    if args.only:
        side = {"c1": lambda: time_fp64_mg_64(hip, stream), "c3": time_ns_step, "c4": time_adjoint,
                "c5": time_lockex}
        print(json.dumps({args.only: side[args.only]()}))
        return

    n = args.n
    cells = n ** 3
    h = 1.0 / n
    bc = T.BCSet.uniform(T.BC_NEUMANN)
    b_host = rhs(n)
    rms_b = float(np.sqrt((b_host.interior() ** 2).mean()))
    tol = args.rel_tol * rms_b
    gb = device.DeviceGrid.from_host(b_host)
    torch.cuda.synchronize()

    # ---- mixed precision (the measured workload)
    mx = solvers.MultigridPressureMixed3D(gb, h, h, h, bc, cycles=args.cycles, cycles_first=args.cycles)
    zero = T.HostGrid.congruent(b_host)
    xdesc = mx.pressure_desc()

    def solve_mixed():
        hip.ocu_grid_clear(C.byref(xdesc), 0.0, stream)
        ok, res = mx.solve(tol, 20)
        if not ok:
            raise RuntimeError(f"mixed solve did not reach {args.rel_tol:g}: {res / rms_b:.3e}")
        return res

    for _ in range(args.warmup):
        solve_mixed()
    hip.ocu_stream_synchronize(stream)
    torch.cuda.synchronize()
    ev = Events(hip, stream)
    counts0 = mx.sweep_counts()
    t0 = time.perf_counter()
    ev.start()
    for _ in range(args.steps):
        res = solve_mixed()
    gpu_ms = ev.stop_ms()
    counts1 = mx.sweep_counts()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = gpu_ms / args.steps
    st = mx.stats()
    x_mixed = mx.pressure().interior().copy()

    out = {
        "metric": METRIC,
        "value": round(cells * args.steps / (gpu_ms / 1e3) / 1e6, 2),
        "unit": "Mcells/s (256^3 cells solved to 1e-10 rel. residual per second)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "wall_ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64 (FP32 V-cycles, FP64 refinement)",
        "data": "synthetic (seeded divergence of a U[-1,1] MAC velocity, mean removed)",
        "config": {"workload": f"mixed-precision MG Poisson {n}^3, unit cube, Neumann faces, to "
                               f"{args.rel_tol:g} relative residual (FP32 V-cycles in FP64 IR)",
                   "grid": [n, n, n], "levels": 7 if n == 256 else None, "nu1": 2, "nu2": 2,
                   "v_cycles_per_refinement": f"{args.cycles} in the first refinement iteration, then as many as "
                                              f"the observed contraction needs (at most {args.cycles})"},
        "rel_residual": res / rms_b,
        "ir_iterations": st["ir_iterations"],
        "vcycles": st["vcycles"],
    }

    # ---- FP64 multigrid on the same problem (reference algorithm, FMG)
    if not args.no_fp64:
        mg = solvers.MultigridPressure3D(gb, h, h, h, bc)
        pdesc = mg.pressure_desc()

        def solve_fp64():
            hip.ocu_grid_clear(C.byref(pdesc), 0.0, stream)
            ok, r = mg.solve(tol, 40)
            if not ok:
                raise RuntimeError("FP64 solve did not converge")
            return r

        for _ in range(args.warmup):
            solve_fp64()
        ev.start()
        reps64 = max(1, min(args.steps, 5))
        for _ in range(reps64):
            solve_fp64()
        ms64 = ev.stop_ms() / reps64
        x64 = mg.pressure().interior()
        x64 = x64 - x64.mean()
        err = float(np.abs(x_mixed - x64).max() / np.abs(x64).max())
        out["fp64_ms_per_step"] = round(ms64, 4)
        out["fp64_mcells_s"] = round(cells / (ms64 / 1e3) / 1e6, 2)
        out["speedup_vs_fp64"] = round(ms64 / ms_step, 3)
        out["fp64_vcycles"] = mg.stats()["vcycles"]
        out["max_rel_err_vs_fp64"] = err
        del mg

    # ---- roofline: the level-0 FP32 fused sweeps (the smoother).  The
    # solver runs three kinds; per solve it launches `counts` of each
    # (measured from its counters over the timed solves), so the dominant
    # kernel is the kind with the most launch time per solve.
    counts = [(b - a) / args.steps for a, b in zip(counts0, counts1)]
    sweeps = time_fused_sweeps((n, n, n), args.kernel_reps, mx.level_params(0))
    kinds = {"plain": 0, "zero": 1, "pro": 2}
    names = {"plain": "k_relax_fused_rows<float,4,1,12,false,false,false> (level-0 red+black sweep, ghost writes)",
             "zero": "k_relax_fused_rows<float,4,1,12,true,false,false> (level-0 sweep from the zero grid)",
             "pro": "k_relax_fused_rows<float,4,1,12,false,true,false> (level-0 sweep adding the 128^3 "
                    "prolongation)"}
    per_solve_us = {k: counts[kinds[k]] * v[0] * 1e3 for k, v in sweeps.items()}
    dom = max(per_solve_us, key=per_solve_us.get)
    kms, kbytes = sweeps[dom]
    achieved = kbytes / (kms / 1e3) / 1e9
    traffic = SWEEP_TRAFFIC_256.get(dom) if n == 256 else None
    out["roofline"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                       "traffic_source": "rocprofv3 FETCH_SIZE and WRITE_SIZE per launch (separate passes), each "
                                         "scaled by what it reads for a plain copy of known size "
                                         "(profiles/level0_kernels_256_hbm_traffic.md)" if traffic else None,
                       "kernel": names[dom], "avg_launch_us": round(kms * 1e3, 2), "alg_bytes_per_launch": kbytes,
                       "launches_per_solve": round(counts[kinds[dom]], 2),
                       "us_per_solve": round(per_solve_us[dom], 1)}
    # the smoother figure: all level-0 sweeps of a solve, algorithmic
    # bytes over time, weighted by how often the solver runs each kind
    tot_b = sum(counts[kinds[k]] * v[1] for k, v in sweeps.items())
    tot_s = sum(counts[kinds[k]] * v[0] / 1e3 for k, v in sweeps.items())
    out["roofline"]["level0_smoother"] = {
        "frac": round(tot_b / tot_s / 1e9 / HBM_PEAK_GBS, 4) if tot_s else None,
        "us_per_solve": round(tot_s * 1e6, 1),
        "kinds": {k: {"launches_per_solve": round(counts[kinds[k]], 2), "avg_launch_us": round(v[0] * 1e3, 2),
                      "alg_bytes_per_launch": v[1],
                      "frac": round(v[1] / (v[0] / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                      "traffic": SWEEP_TRAFFIC_256.get(k) if n == 256 else None}
                  for k, v in sweeps.items()}}
    # the same plain sweep on the 512 x 512 x 256 lock-exchange grid, whose
    # 805 MB working set does not fit the 256 MB last-level cache
    big = (512, 512, 256)
    gb_big = device.DeviceGrid(*big, dtype=np.float32)
    mg_big = solvers.MultigridPressure3D(gb_big, 1.0, 1.0, .5, bc)
    sb = time_fused_sweeps(big, max(5, args.kernel_reps // 5), mg_big.level_params(0), variants=("plain",))
    del mg_big, gb_big
    bms, bbytes = sb["plain"]
    out["roofline"]["beyond_llc"] = {
        "kernel": names["plain"] + " on 512x512x256", "avg_launch_us": round(bms * 1e3, 2),
        "alg_bytes_per_launch": bbytes, "frac": round(bbytes / (bms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}

    if not args.no_extra:
        for key, fn in (("c1_fp64_mg_64", lambda: time_fp64_mg_64(hip, stream)), ("c3_ns_step", time_ns_step),
                        ("c4_adjoint", time_adjoint), ("c5_lockex", time_lockex)):
            t0 = time.perf_counter()
            out[key] = fn()
            print(f"[bench] {key} done in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, args.rel_tol)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
