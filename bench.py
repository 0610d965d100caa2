#!/usr/bin/env python3
"""Benchmark: FP64 HLLC + minmod-MUSCL SSP-RK2 step of 2-D Euler at 4096^2.

Workload (one GPU): 4096 x 4096 cells on [0,1]^2, gamma = 1.4, HLLC flux,
minmod MUSCL on primitives, 2-stage SSP-RK2, CFL 0.4, transmissive sides,
seeded four-quadrant Riemann problem (SplitMix64 seed 0). One step = both
RK stages over the whole grid with the device-side CFL time step.

Prints one JSON line: Gcell-updates/s over the timed steps, the roofline of
the stage kernel (algorithmic bytes / its HIP-event-timed launch duration
against the 8 TB/s HBM peak) and a CPU baseline (the oracle on a bounded
sample of the same workload, timed on this host).
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
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--copy-ref-mib", type=int, default=0,
                    help="also run one plain device copy of this many MiB (reference for HBM counters)")
    return ap.parse_args()


def traffic_from_profiles():
    """HBM bytes per stage launch from the committed rocprofv3 PMC summary."""
    path = os.path.join(ROOT, "profiles", "euler_stage_hllc_muscl_4096_pmc.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    except Exception:
        return None


def cpu_baseline(n, seed):
    """Oracle (OpenMP) on a bounded sample: 3 SSP-RK2 steps of the full grid."""
    import numpy as np
    from oracle import oracle as O
    from golikov_2d_amd import problems as PB

    rows = n
    og = O.make_grid(n, rows, dy=1.0 / n)
    U = PB.four_quadrant(og, seed=seed)
    sc = O.Scheme(O.PHYS_EULER, O.FLUX_HLLC, 2, 0.4)
    steps = 3
    t0 = time.perf_counter()
    O.run(og, sc, 2, U, nsteps=steps)
    dt = time.perf_counter() - t0
    assert np.all(np.isfinite(U))
    return {
        "value": round(n * rows * steps / dt / 1e9, 6),
        "unit": "Gcell-updates/s",
        "cores": O.num_threads(),
        "kind": "port",
        "sample": f"oracle (C, OpenMP, -O3 -ffp-contract=off): the first {steps} SSP-RK2 steps of the same "
                  f"{n}x{rows} workload, {dt:.2f} s wall",
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (this build runs on one "
              "MI355X; the multi-GPU design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from golikov_2d_amd import api
    from golikov_2d_amd import problems as PB

    n = args.n
    dev = torch.device("cuda:0")
    g = api.make_grid(n, n, gamma=1.4)
    sc = api.make_scheme(api.EULER, api.FLUX_HLLC, 2, api.SSPRK2, 0.4)
    U0 = PB.four_quadrant(g, seed=args.seed)
    U = torch.from_numpy(U0).to(dev)
    solver = api.Solver(g, sc)
    solver.start(U)
    solver.advance(U, args.warmup)
    torch.cuda.synchronize()

    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    solver.advance(U, args.steps)
    end.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(end)
    t_sim, nsteps, dt_last = solver.time()
    cells = n * n
    value = cells * args.steps / (ms / 1e3) / 1e9

    Uh = U.cpu().numpy()
    rho = Uh[0]
    p = 0.4 * (Uh[3] - 0.5 * (Uh[1] ** 2 + Uh[2] ** 2) / rho)
    healthy = bool(np.isfinite(Uh).all() and rho.min() > 0 and p.min() > 0)

    # stage-kernel duration with HIP events on the launch stream
    stream = torch.cuda.current_stream()
    S = torch.empty_like(U)
    dtd = torch.tensor([dt_last], dtype=torch.float64, device=dev)
    w = torch.zeros(1, dtype=torch.int64, device=dev)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    t1 = t2 = 0.0
    for _ in range(args.kernel_reps):
        ev[0].record(stream)
        api.stage(g, sc, U, None, S, 0.0, 1.0, dt_dev=dtd)
        ev[1].record(stream)
        api.stage(g, sc, S, U, U, 0.5, 0.5, dt_dev=dtd, wmax_bits=w)
        ev[2].record(stream)
        torch.cuda.synchronize()
        t1 += ev[0].elapsed_time(ev[1])
        t2 += ev[1].elapsed_time(ev[2])
    t1 /= args.kernel_reps
    t2 /= args.kernel_reps
    stage_ms = 0.5 * (t1 + t2)
    achieved = BYTES_PER_CELL_UPDATE / 2.0 * cells / (stage_ms / 1e3) / 1e9  # GB/s per average launch
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic_from_profiles(),
        "kernel": "euler_tile_kernel<2,HLLC,SoA> (stage 1 and stage 2 averaged)",
        "stage1_ms": round(t1, 4),
        "stage2_ms": round(t2, 4),
        "bytes_per_launch": BYTES_PER_CELL_UPDATE / 2.0 * cells,
    }

    if args.copy_ref_mib > 0:
        nel = args.copy_ref_mib * (1 << 20) // 8
        src = torch.ones(nel, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(n, args.seed)

    out = {
        "metric": "Gcell-updates/s FP64 HLLC 4096^2 (% HBM peak)",
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
                   "cells": cells, "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "healthy": healthy,
        "sim_time": t_sim,
        "sim_steps": nsteps,
    }
    print(json.dumps(out), flush=True)
    solver.close()


if __name__ == "__main__":
    main()
