// gd2_ensemble.hpp -- batched ensembles of independent Euler problems with
// uneven grid sizes advanced together: one work list of (problem, tile)
// entries, one launch per RK stage for the whole ensemble (persistent waves
// walk the list), per-problem device clocks, the step sequence captured in a
// hipGraph, and a tuner that picks the tile height and the number of resident
// waves by timing candidate launches.
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <utility>
#include <vector>

#include "gd2_euler.hpp"

namespace gd2 {

struct EnsProblem {
    EulerArgs args;  // grid and physics of the problem (state pointers set per stage)
    double* U;       // caller-owned state (4 SoA planes)
    double* S;       // library-owned stage-1 scratch (4 SoA planes)
    Clock* clk;
};

struct EnsTile {
    int problem, tx, ty, pad;
};

template <int ORDER, int FLUX, int STAGE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(FLUX == GD2_FLUX_HLLC ? 3 : 1))) void ens_stage_kernel(const EnsProblem* __restrict__ probs,
                                                       const EnsTile* __restrict__ tiles, int ntiles, int slot) {
    __shared__ TileLds lds;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const EnsTile tl = tiles[t];
        const EnsProblem& P = probs[tl.problem];
        EulerArgs A = P.args;
        A.dt_dev = &P.clk->dt;
        if constexpr (STAGE == 1) {
            A.Uin = P.U;
            A.Ubase = nullptr;
            A.Uout = P.S;
            A.a = 0.0;
            A.b = 1.0;
            A.wmax_bits = nullptr;
            euler_tile<ORDER, FLUX, kSoA, false, false>(A, tl.tx, tl.ty, lds);
        } else {
            A.Uin = P.S;
            A.Ubase = P.U;
            A.Uout = P.U;
            A.a = 0.5;
            A.b = 0.5;
            A.wmax_bits = &P.clk->w[clock_next(slot)];
            euler_tile<ORDER, FLUX, kSoA, true, true>(A, tl.tx, tl.ty, lds);
        }
    }
}

__global__ void ens_clock_kernel(EnsProblem* probs, int n, int cur) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) clock_begin(probs[p].clk, cur, 0.0);
}

__global__ void ens_clock_start_kernel(Clock* clks, int n, double cfl) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) {
        Clock* c = clks + p;
        c->t = 0.0;
        c->dt = 0.0;
        c->cfl = cfl;
        c->w[0] = 0ull;
        c->w[1] = 0ull;
        c->w[2] = 0ull;
        c->step = 0;
    }
}

}  // namespace gd2
