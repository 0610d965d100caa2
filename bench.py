#!/usr/bin/env python3
"""Benchmark: FP64 HLLC + minmod-MUSCL SSP-RK2 step of 2-D Euler at 4096^2.

Workload (one GPU): 4096 x 4096 cells on [0,1]^2, gamma = 1.4, HLLC flux,
minmod MUSCL on primitives, 2-stage SSP-RK2, CFL 0.4, transmissive sides,
seeded four-quadrant Riemann problem (SplitMix64 seed 0). One step = both
RK stages over the whole grid with the device-side CFL time step.

Prints one JSON line: Gcell-updates/s over the timed steps (each step is one
launch of the one-pass step kernel, which also updates the device clock), the
roofline of that kernel (160 algorithmic bytes per cell-update / its average
launch duration in the timed region, from HIP events on the launch stream,
against the 8 TB/s HBM peak), a CPU baseline (the oracle on a bounded sample
of the same workload, timed on this host), the oracle parity of the first
steps of the same workload, the batched-ensemble and adjoint legs, plus the
ideal-MHD HLLD Orszag-Tang run at 2048^2 and the exact-Riemann Sod run at
200^2.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0           # MI355X HBM3E, GB/s
BYTES_PER_CELL_UPDATE = 160.0   # SSP-RK2: stage 1 reads 32 B + writes 32 B; stage 2 reads 64 B + writes 32 B


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--kernel-reps", type=int, default=60)
    ap.add_argument("--copy-ref-mib", type=int, default=0,
                    help="also run one plain device copy of this many MiB (reference for HBM counters)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batched-ensemble, adjoint, MHD and Sod legs")
    ap.add_argument("--only-extras", action="store_true",
                    help="run only the batched-ensemble, adjoint, MHD and Sod legs (for profiling them)")
    ap.add_argument("--only-leg", choices=["batched", "adjoint", "mhd", "sod", "reflayout", "io", "mhd_adjoint"],
                    default=None,
                    help="run only one of those legs (for profiling it)")
    return ap.parse_args()


METRIC = "Gcell-updates/s FP64 HLLC 4096\u00b2 (% HBM peak); batched\u00f7serial; adjoint\u00f7forward"


def traffic_from_profiles():
    """HBM bytes per one-pass step launch from the committed rocprofv3 PMC summary."""
    path = os.path.join(ROOT, "profiles", "euler_rk2_hllc_muscl_4096_pmc.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    except Exception:
        return None


PARITY_STEPS = 25


def cpu_baseline(n, seed):
    """Oracle (OpenMP) on a bounded sample: the first PARITY_STEPS SSP-RK2
    steps of the full grid. Returns the baseline record and the oracle's state
    after those steps."""
    import numpy as np
    from oracle import oracle as O
    from golikov_2d_amd import problems as PB

    rows = n
    og = O.make_grid(n, rows, dy=1.0 / n)
    U = PB.four_quadrant(og, seed=seed)
    sc = O.Scheme(O.PHYS_EULER, O.FLUX_HLLC, 2, 0.4)
    steps = PARITY_STEPS
    t0 = time.perf_counter()
    O.run(og, sc, 2, U, nsteps=steps)
    dt = time.perf_counter() - t0
    assert np.all(np.isfinite(U))
    threads = O.num_threads()
    # one thread: one step of the same grid
    U1 = PB.four_quadrant(og, seed=seed)
    O.set_num_threads(1)
    try:
        t0 = time.perf_counter()
        O.run(og, sc, 2, U1, nsteps=1)
        dt1 = time.perf_counter() - t0
    finally:
        O.set_num_threads(threads)
    return {
        "value": round(n * rows * steps / dt / 1e9, 6),
        "unit": "Gcell-updates/s",
        "cores": threads,
        "nproc": os.cpu_count(),
        "affinity_cpus": len(os.sched_getaffinity(0)),
        "kind": "port",
        "sample": f"oracle (C, OpenMP, -O3 -ffp-contract=off): the first {steps} SSP-RK2 steps of the same "
                  f"{n}x{rows} workload, {dt:.2f} s wall",
        "one_thread": round(n * rows / dt1 / 1e9, 6),
        "one_thread_sample": f"the first SSP-RK2 step of the same workload on 1 thread, {dt1:.2f} s wall",
    }, U


def cpu_leg(g_kw, scheme, U, integrator, nsteps=0, t_end=0.0):
    """The oracle on all host threads over a given run: Gcell-updates/s."""
    from oracle import oracle as O
    og = O.make_grid(**g_kw)
    t0 = time.perf_counter()
    steps, _ = O.run(og, O.Scheme(*scheme), integrator, U, nsteps=nsteps, t_end=t_end)
    dt = time.perf_counter() - t0
    return {"value": round(og.nx * og.ny * steps / dt / 1e9, 6), "unit": "Gcell-updates/s",
            "cores": O.num_threads(), "kind": "port", "steps": steps, "wall_s": round(dt, 3)}


def batched_vs_serial(seed=0, count=256, steps=50):
    """Ensemble of `count` C1-style problems (sides log-uniform on [64, 1024]):
    one tuned, graph-captured batched advance against a serial loop of
    per-problem solvers, same steps."""
    import torch
    from golikov_2d_amd import api
    from golikov_2d_amd import problems as PB

    dev = torch.device("cuda:0")
    sizes = PB.ensemble_sizes(seed, count=count, lo=64, hi=1024)
    grids = [api.make_grid(nx, ny) for nx, ny in sizes]
    sc = api.make_scheme(api.EULER, api.FLUX_HLLC, 2, api.SSPRK2, 0.4)
    init = [torch.from_numpy(PB.four_quadrant(g, seed=seed * 1000 + p)).to(dev) for p, g in enumerate(grids)]
    cells = sum(g.nx * g.ny for g in grids)

    ens_U = [u.clone() for u in init]
    e = api.Ensemble(grids, sc, ens_U)
    e.start()
    rows, blocks, _ = e.tune()
    e.advance(steps, use_graph=True)  # capture + warm-up
    for u, u0 in zip(ens_U, init):
        u.copy_(u0)
    e.start()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e.advance(steps, use_graph=True)
    torch.cuda.synchronize()
    t_batched = time.perf_counter() - t0

    ser_U = [u.clone() for u in init]
    solvers = [api.Solver(g, sc) for g in grids]
    for s_, u in zip(solvers, ser_U):  # warm-up pass
        s_.start(u)
        s_.advance(u, 1)
    for u, u0 in zip(ser_U, init):
        u.copy_(u0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s_, u in zip(solvers, ser_U):
        s_.start(u)
        s_.advance(u, steps)
    torch.cuda.synchronize()
    t_serial = time.perf_counter() - t0
    same = all(torch.equal(a, b) for a, b in zip(ens_U, ser_U))
    for s_ in solvers:
        s_.close()
    e.close()
    return {
        "value": round(t_serial / t_batched, 3),
        "unit": "x (serial wall time / batched wall time)",
        "batched_ms": round(1e3 * t_batched, 3),
        "serial_ms": round(1e3 * t_serial, 3),
        "batched_gcells": round(cells * steps / t_batched / 1e9, 3),
        "problems": count,
        "steps": steps,
        "cells": cells,
        "tiling": {"rows": rows, "blocks": blocks},
        "bitwise_equal_to_serial": same,
        "workload": f"{count} problems, sides log-uniform on [64, 1024] (seed {seed}), C1-style four-quadrant "
                    "states, HLLC + minmod, SSP-RK2, own CFL dt each; batched = one persistent launch of the "
                    "one-pass step for all steps (XCD-affine work lists, tuned strips), replayed from a hipGraph",
    }


def mhd_orszag_tang(n=2048, steps=100, seed=0, reps=10, with_cpu=True):
    """Ideal MHD Orszag-Tang vortex (gamma 5/3, periodic, seeded 1e-3 Fourier
    perturbations), HLLD + minmod MUSCL, SSP-RK2, CFL 0.3: Gcell-updates/s
    over `steps` solver steps, and the two stage kernels' HIP-event durations
    against the 280 algorithmic bytes per cell-update (7 fields: stage 1
    reads and writes 56 B, stage 2 reads 112 B and writes 56 B)."""
    import torch
    from golikov_2d_amd import api
    from golikov_2d_amd import problems as PB

    dev = torch.device("cuda:0")
    g = api.make_grid(n, n, gamma=5.0 / 3.0, bc=(api.BC_PERIODIC,) * 4)
    sc = api.make_scheme(api.MHD, api.FLUX_HLLD, 2, api.SSPRK2, 0.3)
    U = torch.from_numpy(PB.orszag_tang(g, seed=seed, amp=1e-3)).to(dev)
    s = api.Solver(g, sc)
    s.start(U)
    s.advance(U, 5)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    s.advance(U, steps)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    _, _, dt_last = s.time()
    s.close()
    stream = torch.cuda.current_stream()
    S = torch.empty_like(U)
    O = torch.empty_like(U)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * reps + 1)]
    api.stage(g, sc, U, None, S, 0.0, 1.0, dt_last)
    api.stage(g, sc, S, U, O, 0.5, 0.5, dt_last)
    ev[0].record(stream)
    for i in range(reps):
        api.stage(g, sc, U, None, S, 0.0, 1.0, dt_last)
        ev[2 * i + 1].record(stream)
        api.stage(g, sc, S, U, O, 0.5, 0.5, dt_last)
        ev[2 * i + 2].record(stream)
    torch.cuda.synchronize()
    t1 = sum(ev[2 * i].elapsed_time(ev[2 * i + 1]) for i in range(reps)) / reps
    t2 = sum(ev[2 * i + 1].elapsed_time(ev[2 * i + 2]) for i in range(reps)) / reps
    cells = n * n
    achieved = 280.0 * cells / ((t1 + t2) / 1e3) / 1e9
    cpu = None
    if with_cpu:
        from oracle import oracle as O
        cpu = cpu_leg(dict(nx=n, ny=n, gamma=5.0 / 3.0, bc=(O.BC_PERIODIC,) * 4),
                      (O.PHYS_MHD, O.FLUX_HLLD, 2, 0.3), PB.orszag_tang(g, seed=seed, amp=1e-3), 2, nsteps=3)
        cpu["sample"] = "oracle (C, OpenMP): the first 3 SSP-RK2 steps of the same Orszag-Tang workload"
    return {
        "value": round(cells * steps / (ms / 1e3) / 1e9, 4),
        "unit": "Gcell-updates/s",
        "ms_per_step": round(ms / steps, 4),
        "stage_kernel_ms": [round(t1, 4), round(t2, 4)],
        "bytes_per_cell_update": 280,
        "achieved_GBs": round(achieved, 1),
        "hbm_frac": round(achieved / HBM_PEAK_GBS, 4),
        "workload": f"ideal MHD Orszag-Tang {n}x{n}, gamma 5/3, periodic, HLLD + minmod MUSCL, SSP-RK2, CFL 0.3, "
                    f"{steps} steps (seed {seed})",
        "cpu_baseline": cpu,
    }


def sod_godunov(n=200, with_cpu=True):
    """Sod shock tube along x on n x n cells, exact Riemann (Godunov) flux,
    first order, forward Euler, CFL 0.4, transmissive, run to t = 0.2 (the
    second of two identical runs is timed)."""
    import torch
    from golikov_2d_amd import api
    from golikov_2d_amd import problems as PB

    dev = torch.device("cuda:0")
    g = api.make_grid(n, n)
    sc = api.make_scheme(api.EULER, api.FLUX_EXACT, 1, api.FORWARD_EULER, 0.4)
    U0 = torch.from_numpy(PB.sod_x(g)).to(dev)
    for rep in range(2):
        U = U0.clone()
        s = api.Solver(g, sc)
        s.start(U)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        t_end, nsteps = s.run_to(U, 0.2, chunk=64)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        s.close()
    cpu = None
    if with_cpu:
        from oracle import oracle as O
        cpu = cpu_leg(dict(nx=n, ny=n), (O.PHYS_EULER, O.FLUX_EXACT, 1, 0.4), PB.sod_x(g), 1, t_end=0.2)
        cpu["sample"] = "oracle (C, OpenMP): the whole run to t = 0.2"
    return {
        "value": round(n * n * nsteps / wall / 1e9, 4),
        "unit": "Gcell-updates/s",
        "cpu_baseline": cpu,
        "steps": nsteps,
        "wall_ms": round(1e3 * wall, 3),
        "t_end": t_end,
        "workload": f"Sod shock tube {n}x{n} (x-directed), exact Riemann flux, first order, forward Euler, "
                    "CFL 0.4, run to t = 0.2",
    }


def reference_layout(n_mhd=2048, n_euler=4096, steps=20, seed=0, with_cpu=True):
    """The drop-in entry points driven the way the reference's step loop drives
    its kernels: one reference-layout step (AoS primitives, lagged face time
    step) plus gd2_funk_time per step, ping-ponging the state buffers.
    MHD: gd2_kernel_tvd (MUSCL, HLLD = method 3) on the Orszag-Tang state;
    Euler: gd2_kernel_tvd_old (MUSCL, HLLC = method 1) on the four-quadrant
    state. Algorithmic bytes per cell-update: the AoS state read and written
    once (MHD 56 + 56 B, Euler 32 + 32 B)."""
    import torch
    from golikov_2d_amd import api
    from golikov_2d_amd import problems as PB

    dev = torch.device("cuda:0")
    out = {}
    for kind, n in (("mhd_kernel_tvd", n_mhd), ("euler_kernel_tvd_old", n_euler), ("mhd_wind_kernel_tvd", 0),
                    ("mhd_wind_kernel_gkr", 0)):
        mhd = kind.startswith("mhd")
        wind = kind.startswith("mhd_wind")
        gkr = kind == "mhd_wind_kernel_gkr"
        method = 3 if gkr else (2 if wind else 3)
        if wind:
            # the reference's production run: 3226 x 1792, symmetry / outflow /
            # axis / inflow sides, frozen inner sphere, axisymmetric sources,
            # HLLC-MHD (method 2), stellar-wind initial state; Kernel_GKR with
            # its own frozen radius 2.0 and HLLD
            nx, ny = PB.WIND_NX, PB.WIND_NY
            wkw = dict(PB.wind_grid_args(), r_frozen=2.0) if gkr else PB.wind_grid_args()
            g = api.make_grid(nx, ny, **wkw)
            sh0, uh0, bh0 = PB.stellar_wind_prim(g)
            s, u, b = (torch.from_numpy(a).to(dev) for a in (sh0, uh0, bh0))
        else:
            nx = ny = n
            g = api.make_grid(n, n, gamma=5.0 / 3.0 if mhd else 1.4,
                              bc=(api.BC_PERIODIC,) * 4 if mhd else (api.BC_TRANSMISSIVE,) * 4)
            U0 = PB.orszag_tang(g, seed=seed, amp=1e-3) if mhd else PB.four_quadrant(g, seed=seed)
            phys = api.MHD if mhd else api.EULER
            s, u, b = api.cons_to_prim(g, phys, torch.from_numpy(U0).to(dev))
        s2, u2 = torch.empty_like(s), torch.empty_like(u)
        b2 = torch.empty_like(b) if mhd else None
        T = torch.tensor([1e7], dtype=torch.float64, device=dev)
        T_do = torch.tensor([1e-8], dtype=torch.float64, device=dev)  # the reference's first step
        TT = torch.zeros(1, dtype=torch.float64, device=dev)
        i = torch.zeros(1, dtype=torch.int32, device=dev)

        def step():
            nonlocal s, u, b, s2, u2, b2
            if gkr:
                api.kernel_gkr(g, s, u, b, s2, u2, b2, T, T_do, 0, method)
                b, b2 = b2, b
            elif mhd:
                api.kernel_tvd(g, s, u, b, s2, u2, b2, T, T_do, 0, method)
                b, b2 = b2, b
            else:
                api.kernel_tvd_old(g, s, u, s2, u2, T, T_do, 1)
            s, s2 = s2, s
            u, u2 = u2, u
            api.funk_time(T, T_do, TT, i)

        for _ in range(5):
            step()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(steps):
            step()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        bpc = 112 if mhd else 64
        cells = nx * ny
        if gkr:
            work = (f"gd2_kernel_gkr (GKR predictor-corrector, HLLD) + gd2_funk_time, stellar wind {nx}x{ny}, "
                    "its own sides, frozen sphere r < 2.0, axisymmetric")
        elif wind:
            work = (f"gd2_kernel_tvd (MUSCL + HLLC-MHD, method 2) + gd2_funk_time, stellar wind {nx}x{ny}, "
                    f"symmetry/outflow/axis/inflow sides, frozen sphere r < {PB.WIND_DDIST}, axisymmetric")
        elif mhd:
            work = f"gd2_kernel_tvd (MUSCL + HLLD, method 3) + gd2_funk_time, Orszag-Tang {n}x{n}, periodic"
        else:
            work = f"gd2_kernel_tvd_old (MUSCL + HLLC, method 1) + gd2_funk_time, four-quadrant {n}x{n}, transmissive"
        r = {"value": round(cells / (ms / 1e3) / 1e9, 4), "unit": "Gcell-updates/s", "ms_per_step": round(ms, 4),
             "bytes_per_cell_update": bpc, "hbm_frac": round(bpc * cells / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
             "finite": bool(torch.isfinite(s).all()), "workload": work + f", {steps} steps after 5"}
        if with_cpu:
            from oracle import oracle as O
            if wind:
                og = O.make_grid(nx, ny, **wkw)
            else:
                og = O.make_grid(n, n, gamma=g.gamma, bc=tuple(g.bc))
            sh, uh = s.cpu().numpy().copy(), u.cpu().numpy().copy()
            Tdo = float(T_do.item())
            t0 = time.perf_counter()
            if gkr:
                O.ref_gkr_step(og, sh, uh, b.cpu().numpy().copy(), 1e7, Tdo, method)
            elif mhd:
                O.ref_mhd_step("kernel_tvd", og, sh, uh, b.cpu().numpy().copy(), 1e7, Tdo, 0, method)
            else:
                O.ref_euler_step("kernel_tvd_old", og, sh, uh, 1e7, Tdo, 1)
            dt = time.perf_counter() - t0
            r["cpu_baseline"] = {"value": round(cells / dt / 1e9, 6), "unit": "Gcell-updates/s",
                                 "cores": O.num_threads(), "kind": "port",
                                 "sample": "oracle (C, OpenMP): one step of the same reference-layout workload"}
        out[kind] = r
    return out


def io_formats(with_cpu=True, sample_rows=112):
    """The reference's on-disk formats written by the library's host writers
    at the production size (stellar-wind state, 3226 x 1792, host buffers):
    text state / restart, save_b3.dat, Tecplot every 3rd cell; and the text
    restart read back. MB/s of file per second of wall time, best of two. The
    oracle's one-cell-at-a-time stdio writer (the reference's ofstream loop)
    is timed beside it on the first sample_rows rows."""
    import shutil
    import tempfile

    import numpy as np
    from golikov_2d_amd import api
    from golikov_2d_amd import problems as PB

    nx, ny = PB.WIND_NX, PB.WIND_NY
    kw = PB.wind_grid_args()
    g = api.make_grid(nx, ny, **kw)
    s, u, b = PB.stellar_wind_prim(g)
    d = tempfile.mkdtemp(prefix="gd2_io_")
    out = {"workload": f"stellar-wind state {nx}x{ny} (host buffers), files under {tempfile.gettempdir()}",
           "threads": min(16, os.cpu_count() or 1) if not os.environ.get("GD2_IO_THREADS")
           else int(os.environ["GD2_IO_THREADS"])}
    try:
        def timed(fn, path):
            best = None
            for _ in range(2):
                t0 = time.perf_counter()
                fn(path)
                dt = time.perf_counter() - t0
                best = dt if best is None else min(best, dt)
            return best, os.path.getsize(path)

        for name, fn in (("state_text", lambda p: api.write_state_text(g, s, u, b, p)),
                         ("save_b3", lambda p: api.write_state_b3(g, s, u, b, p)),
                         ("tecplot_nmin3", lambda p: api.write_tecplot(g, s, u, b, p, nmin=3, solution_time=0.0))):
            path = os.path.join(d, name)
            t, size = timed(fn, path)
            out[name] = {"MB_per_s": round(size / t / 1e6, 1), "seconds": round(t, 4), "bytes": size}
        path = os.path.join(d, "state_text")
        t0 = time.perf_counter()
        s2, _, _ = api.read_state_text(g, path)
        t = time.perf_counter() - t0
        out["read_state_text"] = {"MB_per_s": round(os.path.getsize(path) / t / 1e6, 1), "seconds": round(t, 4),
                                  "rho_max_rel_err": float(np.max(np.abs(s2[..., 0] - s[..., 0]) / s[..., 0]))}
        if with_cpu:
            from oracle import oracle as O
            og = O.make_grid(nx, sample_rows, **kw)
            ss, us, bs = (np.ascontiguousarray(a[:sample_rows]) for a in (s, u, b))
            path = os.path.join(d, "oracle_text")
            t0 = time.perf_counter()
            O.write_state_text(og, ss, us, bs, path)
            t = time.perf_counter() - t0
            out["cpu_baseline"] = {"value": round(os.path.getsize(path) / t / 1e6, 1), "unit": "MB/s", "cores": 1,
                                   "kind": "port", "sample": f"oracle stdio text writer, first {sample_rows} rows"}
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return out


def adjoint_vs_forward(n=1024, steps=200, every=20, seed=0, reps=3):
    """Differentiable rollout (LibTorch autograd Function): backward (recompute
    from checkpoints + stage VJPs) against the forward rollout, both timed."""
    import torch
    from golikov_2d_amd import adjoint as AD
    from golikov_2d_amd import api
    from golikov_2d_amd import problems as PB

    dev = torch.device("cuda:0")
    g = api.make_grid(n, n)
    sc = api.make_scheme(api.EULER, api.FLUX_HLLC, 2, api.SSPRK2, 0.4)
    quad = PB.four_quadrant(g, seed=seed)
    U0 = torch.from_numpy(PB.smooth_perturbed_euler(g, quad, seed + 100, 1e-3)).to(dev)
    Ut = torch.from_numpy(PB.smooth_perturbed_euler(g, quad, seed + 200, 1e-2)).to(dev)
    dts, _ = AD.dt_schedule(g, sc, U0, steps)
    with torch.no_grad():
        rho_t = AD.rollout(g, sc, Ut, dts, every=steps + 1)[0].clone()
    fwd, bwd = [], []
    for _ in range(reps + 1):
        X = U0.clone().requires_grad_(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        UT = AD.rollout(g, sc, X, dts, every=every)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        AD.density_loss(UT, rho_t).backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        fwd.append(t1 - t0)
        bwd.append(t2 - t1)
    f, b = sorted(fwd[1:])[reps // 2], sorted(bwd[1:])[reps // 2]  # medians after one warm-up
    return {
        "value": round(b / f, 3),
        "unit": "x (backward wall time / forward wall time)",
        "forward_ms": round(1e3 * f, 3),
        "backward_ms": round(1e3 * b, 3),
        "n": n,
        "steps": steps,
        "checkpoint_every": every,
        "workload": f"{n}x{n} four-quadrant state (seed {seed}) with a 1e-3 smooth perturbation, HLLC + minmod, "
                    f"SSP-RK2, {steps} recorded CFL steps, loss 1/2 ||rho(T) - rho_target||^2",
    }


def mhd_adjoint_vs_forward(n=1024, steps=100, every=20, seed=0, reps=3):
    """The MHD (HLLD) differentiable rollout: backward against forward, as for
    the Euler adjoint, on the Orszag-Tang state."""
    import torch
    from golikov_2d_amd import adjoint as AD
    from golikov_2d_amd import api
    from golikov_2d_amd import problems as PB

    dev = torch.device("cuda:0")
    g = api.make_grid(n, n, gamma=5.0 / 3.0, bc=(api.BC_PERIODIC,) * 4)
    sc = api.make_scheme(api.MHD, api.FLUX_HLLD, 2, api.SSPRK2, 0.3)
    U0 = torch.from_numpy(PB.orszag_tang(g, seed=seed, amp=1e-3)).to(dev)
    Ut = torch.from_numpy(PB.orszag_tang(g, seed=seed + 1, amp=1e-2)).to(dev)
    dts, _ = AD.dt_schedule(g, sc, U0, steps)
    with torch.no_grad():
        rho_t = AD.rollout(g, sc, Ut, dts, every=steps + 1)[0].clone()
    fwd, bwd = [], []
    for _ in range(reps + 1):
        X = U0.clone().requires_grad_(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        UT = AD.rollout(g, sc, X, dts, every=every)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        AD.density_loss(UT, rho_t).backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        fwd.append(t1 - t0)
        bwd.append(t2 - t1)
    f, b = sorted(fwd[1:])[reps // 2], sorted(bwd[1:])[reps // 2]
    return {
        "value": round(b / f, 3),
        "unit": "x (backward wall time / forward wall time)",
        "forward_ms": round(1e3 * f, 3),
        "backward_ms": round(1e3 * b, 3),
        "n": n,
        "steps": steps,
        "checkpoint_every": every,
        "workload": f"{n}x{n} Orszag-Tang (seed {seed}, 1e-3 perturbation), HLLD + minmod, Powell term, SSP-RK2, "
                    f"{steps} recorded CFL steps, loss 1/2 ||rho(T) - rho_target||^2",
    }


def copy_reference(mib):
    """One plain device-to-device copy of `mib` MiB: the reference dispatch
    that scales the HBM byte counters."""
    import torch
    src = torch.ones(mib * (1 << 20) // 8, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def copy_ceiling_gbs(mib=1024, reps=5):
    """Measured HBM ceiling: a plain device-to-device copy of `mib` MiB on the
    current stream, (bytes read + bytes written) / time, best of `reps`."""
    import torch
    src = torch.ones(mib * (1 << 20) // 8, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(2):
        dst.copy_(src)
    s = torch.cuda.current_stream()
    best = float("inf")
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        dst.copy_(src)
        e1.record(s)
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1))
    del src, dst
    return 2.0 * mib * (1 << 20) / (best / 1e3) / 1e9


CLOCK_WARMUP_MS = 100.0


def clock_warmup(api, g, sc, U):
    """Run the solver's step on a scratch copy of U for CLOCK_WARMUP_MS of
    wall time (untimed work before the measured run); returns (ms, steps)."""
    import torch
    W = U.clone()
    s = api.Solver(g, sc)
    s.start(W)
    steps = 0
    t0 = time.perf_counter()
    while (time.perf_counter() - t0) * 1e3 < CLOCK_WARMUP_MS:
        s.advance(W, 20)
        torch.cuda.synchronize()
        steps += 20
    ms = (time.perf_counter() - t0) * 1e3
    s.close()
    del W
    return ms, steps


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (one MI355X per run; "
              "a grid would split into row strips with a halo exchange per step)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from golikov_2d_amd import api
    from golikov_2d_amd import problems as PB

    if args.only_leg:
        if args.copy_ref_mib > 0:
            copy_reference(args.copy_ref_mib)
        legs = {"batched": lambda: batched_vs_serial(args.seed), "adjoint": lambda: adjoint_vs_forward(seed=args.seed),
                "mhd": lambda: mhd_orszag_tang(seed=args.seed, with_cpu=not args.no_cpu),
                "sod": lambda: sod_godunov(with_cpu=not args.no_cpu),
                "reflayout": lambda: reference_layout(seed=args.seed, with_cpu=not args.no_cpu),
                "io": lambda: io_formats(with_cpu=not args.no_cpu),
                "mhd_adjoint": lambda: mhd_adjoint_vs_forward(seed=args.seed)}
        print(json.dumps({args.only_leg: legs[args.only_leg]()}), flush=True)
        return
    if args.only_extras:
        print(json.dumps({"batched_vs_serial": batched_vs_serial(args.seed),
                          "adjoint_vs_forward": adjoint_vs_forward(seed=args.seed),
                          "mhd_orszag_tang_2048": mhd_orszag_tang(seed=args.seed, with_cpu=not args.no_cpu),
                          "sod_godunov_200": sod_godunov(with_cpu=not args.no_cpu),
                  "reference_layout": reference_layout(seed=args.seed, with_cpu=not args.no_cpu),
                  "io_formats": io_formats(with_cpu=not args.no_cpu),
                  "mhd_adjoint_vs_forward": mhd_adjoint_vs_forward(seed=args.seed)}), flush=True)
        return

    n = args.n
    dev = torch.device("cuda:0")
    g = api.make_grid(n, n, gamma=1.4)
    sc = api.make_scheme(api.EULER, api.FLUX_HLLC, 2, api.SSPRK2, 0.4)
    if not api.one_pass_preferred(g, sc):
        raise SystemExit("the 4096^2 step is expected to take the one-pass kernel")
    U0 = PB.four_quadrant(g, seed=args.seed)
    U = torch.from_numpy(U0).to(dev)
    stream = torch.cuda.current_stream()
    copy_gbs = copy_ceiling_gbs()
    # Engine clock: after idle the GPU runs its first tens of milliseconds of
    # work at a lower clock (measured on this kernel: the same cycle count per
    # launch, 2.1 GHz right after idle against 2.4 GHz once busy), which a few
    # warm-up steps do not cover. So before the warm-up steps the same step
    # kernel runs on a scratch copy of the state for CLOCK_WARMUP_MS (untimed;
    # the timed region is unchanged: K steps of the run itself).
    warm_ms, warm_steps = clock_warmup(api, g, sc, U)

    solver = api.Solver(g, sc)
    solver.start(U)
    solver.advance(U, args.warmup)
    torch.cuda.synchronize()
    # timed region: K steps, each one launch of the one-pass step kernel (the
    # kernel takes its dt from the device clock and updates it), back to back
    # on the launch stream; an odd K adds one device copy back into U at the end
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    solver.advance(U, args.steps)
    end.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(end)
    t_sim, nsteps, dt_last = solver.time()
    cells = n * n
    value = cells * args.steps / (ms / 1e3) / 1e9

    # for comparison, with HIP events on the launch stream, outside the timed
    # region: the one-pass step kernel back to back, then the two stage kernels
    # of the stage-by-stage step
    S = torch.empty_like(U)
    dtd = torch.tensor([dt_last], dtype=torch.float64, device=dev)
    w = torch.zeros(1, dtype=torch.int64, device=dev)
    reps = 2 * max(1, args.kernel_reps // 2)
    for i in range(4):  # warm launches
        src, dst = (U, S) if i % 2 == 0 else (S, U)
        api.ssprk2_step(g, sc, src, dst, dt_dev=dtd, wmax_bits=w)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for i in range(reps):
        src, dst = (U, S) if i % 2 == 0 else (S, U)
        api.ssprk2_step(g, sc, src, dst, dt_dev=dtd, wmax_bits=w)
    e1.record(stream)
    torch.cuda.synchronize()
    step_kernel_ms = e0.elapsed_time(e1) / reps
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * reps + 1)]
    for _ in range(2):
        api.stage(g, sc, U, None, S, 0.0, 1.0, dt_dev=dtd)
        api.stage(g, sc, S, U, U, 0.5, 0.5, dt_dev=dtd, wmax_bits=w)
    ev[0].record(stream)
    for i in range(reps):
        api.stage(g, sc, U, None, S, 0.0, 1.0, dt_dev=dtd)
        ev[2 * i + 1].record(stream)
        api.stage(g, sc, S, U, U, 0.5, 0.5, dt_dev=dtd, wmax_bits=w)
        ev[2 * i + 2].record(stream)
    torch.cuda.synchronize()
    t1 = sum(ev[2 * i].elapsed_time(ev[2 * i + 1]) for i in range(reps)) / reps
    t2 = sum(ev[2 * i + 1].elapsed_time(ev[2 * i + 2]) for i in range(reps)) / reps
    del S

    # the dominant kernel's average launch duration inside the timed region:
    # the region holds exactly K launches of it and nothing else (even K)
    if args.steps % 2 == 0:
        kernel_ms = ms / args.steps
        kernel_src = (f"timed region / {args.steps}: {args.steps} launches of the one-pass step kernel back to "
                      "back between two HIP events on the launch stream, nothing else enqueued")
    else:
        kernel_ms = step_kernel_ms
        kernel_src = f"{reps} launches back to back after the timed region (odd K: the region ends with a copy)"
    bytes_per_launch = BYTES_PER_CELL_UPDATE * cells  # one launch advances every cell one full step
    achieved = bytes_per_launch / (kernel_ms / 1e3) / 1e9
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic_from_profiles(),
        "kernel": "euler_rk2_kernel<2,HLLC> (one-pass SSP-RK2 step)",
        "kernel_ms": round(kernel_ms, 4),
        "kernel_ms_source": kernel_src,
        "kernel_ms_back_to_back": round(step_kernel_ms, 4),
        "bytes_per_launch": bytes_per_launch,
        "two_pass_stage_ms": [round(t1, 4), round(t2, 4)],
        "copy_ceiling": round(copy_gbs, 1),
        "frac_of_copy_ceiling": round(achieved / copy_gbs, 4),
    }
    if kernel_ms > ms / args.steps * (1 + 1e-9):
        raise SystemExit(f"kernel time {kernel_ms} ms exceeds the step time {ms / args.steps} ms")

    Uh = U.cpu().numpy()
    rho = Uh[0]
    p = 0.4 * (Uh[3] - 0.5 * (Uh[1] ** 2 + Uh[2] ** 2) / rho)
    healthy = bool(np.isfinite(Uh).all() and rho.min() > 0 and p.min() > 0)

    # parity on the benchmarked workload: a fresh run of the first
    # PARITY_STEPS steps (one-pass kernel, fused clock) against the oracle's
    # state after the same steps, computed by the CPU baseline leg below
    Up = torch.from_numpy(U0).to(dev)
    psolver = api.Solver(g, sc)
    psolver.start(Up)
    psolver.advance(Up, PARITY_STEPS)
    gpu_after = Up.cpu().numpy()
    _, psteps, _ = psolver.time()
    psolver.close()
    del Up

    if args.copy_ref_mib > 0:
        copy_reference(args.copy_ref_mib)

    cpu = None
    parity = {"steps": PARITY_STEPS, "gpu_steps": psteps, "checked": False}
    if not args.no_cpu:
        cpu, oracle_after = cpu_baseline(n, args.seed)
        err = np.abs(gpu_after - oracle_after).reshape(4, -1).sum(1) / np.abs(oracle_after).reshape(4, -1).sum(1)
        parity.update({"checked": True, "rel_l1_per_field": [float(f"{e:.3e}") for e in err],
                       "gate": 1e-10, "ok": bool(np.all(err <= 1e-10)) and psteps == PARITY_STEPS,
                       "against": "oracle (C) after the same steps from the same seeded state"})
        healthy = healthy and parity["ok"]
    extras = None
    if not args.no_extras:
        extras = {"batched_vs_serial": batched_vs_serial(args.seed),
                  "adjoint_vs_forward": adjoint_vs_forward(seed=args.seed),
                  "mhd_orszag_tang_2048": mhd_orszag_tang(seed=args.seed, with_cpu=not args.no_cpu),
                  "sod_godunov_200": sod_godunov(with_cpu=not args.no_cpu),
                  "reference_layout": reference_layout(seed=args.seed, with_cpu=not args.no_cpu),
                  "io_formats": io_formats(with_cpu=not args.no_cpu),
                  "mhd_adjoint_vs_forward": mhd_adjoint_vs_forward(seed=args.seed)}

    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"2-D Euler {n}x{n}, HLLC + minmod MUSCL, SSP-RK2, CFL 0.4, transmissive, "
                               f"seeded four-quadrant Riemann problem (seed {args.seed})",
                   "cells": cells, "parallelism": "single GPU",
                   "clock_warmup": {"ms": round(warm_ms, 1), "steps": warm_steps,
                                    "what": "untimed: the same step kernel on a scratch copy of the state before "
                                            "the warm-up steps, to bring the engine clock up from idle"}},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity": parity,
        "extras": extras,
        "healthy": healthy,
        "sim_time": t_sim,
        "sim_steps": nsteps,
    }
    print(json.dumps(out), flush=True)
    solver.close()


if __name__ == "__main__":
    main()
