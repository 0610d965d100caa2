// gd2_ensemble.hpp -- batched ensembles of independent Euler problems with
// uneven grid sizes advanced together.
//
// SSP-RK2 steps run in ONE persistent launch for up to kEnsMaxSteps steps.
// The work list is (step, one-pass tile) in step-major order, taken from a
// counter by resident waves; a tile of step s of problem p first waits until
// every tile of step s-1 of p has finished (a per-problem completion
// counter), because its halo rows and its dt (from the problem's wave-speed
// maximum) come from that step. Every wait is for items taken from the list
// earlier, by waves that are running, so the launch always drains; each wait
// is also bounded by a spin limit that flags an error instead of hanging.
//
// XCD-affine queues: the problems are packed into eight bins of equal work,
// one work list per XCD, and a wave takes items only from the list of the
// XCD it runs on (its XCC_ID). A problem's state, halos and counters then
// stay within one L2: a tile publishes its rows by waiting for its stores to
// complete (they are in the shared L2) and a waiting tile invalidates its
// CU's L1, with no L2 write-back per tile. When the bins cannot be balanced
// (one problem larger than an XCD's share) or the launch is too small to
// give every XCD waves, one list serves the whole device with agent-scope
// release and acquire.
//
// The clocks: each launch keeps a per-problem history of the wave-speed
// maximum, one slot per step (slot 0 from the problem's clock, the others
// cleared before the launch); step s takes dt from slot s and folds into
// slot s+1, by atomic maximum only. A one-thread-per-problem kernel after
// the launch replays dt and t step by step into the problem's clock, in the
// order a single problem's clock takes them (bitwise the same t).
//
// The state ping-pongs between the caller's buffer and a library-owned one
// (even step counts end in the caller's buffer; an odd count first takes
// one step as two stage launches, which write in place). The launches can be
// captured in a hipGraph; a tuner picks the strip height and the number of
// resident waves by timing candidate launches.
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <utility>
#include <vector>

#include "gd2_euler.hpp"
#include "gd2_rk2.hpp"

namespace gd2 {

struct EnsProblem {
    EulerArgs args;  // grid and physics of the problem (state pointers set per stage)
    double* U;       // caller-owned state (4 SoA planes)
    double* S;       // library-owned second state (4 SoA planes): stage-1 output or ping-pong partner
    Clock* clk;
    double clk_cfl;  // the scheme's CFL number
    int rk2_rows;    // strip height of the one-pass tiles
    int rk2_tiles;   // one-pass tiles of the problem (completions per step)
};

struct EnsTile {
    int problem, tx, ty, pad;
};

constexpr int kEnsQueues = 8;        // XCDs of the device: one work list each
constexpr int kEnsMaxSteps = 64;     // steps of one persistent launch
constexpr int kEnsHist = kEnsMaxSteps + 1;
// sync words of a persistent launch: [0, 8) work-list heads, [8] spin-limit
// flag, [9 + p] tiles of problem p finished in this launch, [9 + n + p] the
// tiles problem p has in it (checked by the host: every list drained)
constexpr int kEnsFlag = kEnsQueues;
constexpr int kEnsSyncHead = kEnsQueues + 1;
constexpr unsigned kEnsSpinLimit = 1u << 20;  // polls of ~0.3 us or more: far beyond any legitimate wait

// tile-list offsets of the work lists (off[8] = tiles per step); one list
// (off[1] = ... = off[8]) when the launch is not XCD-affine
struct EnsQueues {
    int off[kEnsQueues + 1];
};

// The persistent loop's own control is free of lane-divergent branches: its
// atomics run on every lane with wave-uniform operands (lane 0 adds 1, the
// others 0; the same maximum on every lane), which the compiler turns into
// one atomic per wave, and its waits poll on every lane with a wave-uniform
// exit. (A lane-0 branch next to the loop's back edge lets the compiler
// restructure the loop so that lanes 1..63 start the next item without lane
// 0 and never leave it; a 64-bit maximum through a generic pointer stays 64
// atomics on one address.)

typedef __attribute__((address_space(1))) unsigned long long ens_gu64;

__device__ __forceinline__ void ens_fold_max(unsigned long long* dst, double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    // readfirstlane returns int: convert through unsigned (no sign extension)
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((unsigned)b);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
    const unsigned long long u = (unsigned long long)lo | ((unsigned long long)hi << 32);
    __hip_atomic_fetch_max((ens_gu64*)dst, u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned long long ens_load_u64(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wait until *ctr >= need, then acquire: invalidate this CU's caches (agent
// scope). The polls are relaxed device-coherent loads with a sleep between
// them (an acquiring poll would invalidate the caches on every poll, which
// with many waiting waves slows the working ones several times over). On
// reaching the spin limit raise *flag and stop waiting; once the flag is up
// no wave waits again (the launch then drains at once; the host reports it).
__device__ __forceinline__ void ens_wait(const unsigned* ctr, unsigned need, unsigned* flag) {
    unsigned spins = 0;
    for (;;) {
        const unsigned v =
            __builtin_amdgcn_readfirstlane(__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        if (v >= need) break;
        const unsigned f =
            __builtin_amdgcn_readfirstlane(__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        if (++spins >= kEnsSpinLimit || f != 0u) {
            __hip_atomic_fetch_or(flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
        __builtin_amdgcn_s_sleep(8);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// Persistent one-pass SSP-RK2 steps of the whole ensemble. Step s reads U for
// even s and S for odd s and writes the other, with dt = cfl / hist[s] and
// its wave-speed maximum folded into hist[s+1] (hist: kEnsHist words per
// problem). XCD: the lists are XCD-affine (see above). dry: timing runs of
// the tuner; every step reads U and writes S with dt from hist[0], no fold.
template <int ORDER, int FLUX, bool XCD>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu((FLUX == GD2_FLUX_EXACT) ? 1 : 2))) void ens_rk2_kernel(
    const EnsProblem* __restrict__ probs, const EnsTile* __restrict__ tiles, const EnsQueues q, int nsteps,
    unsigned long long* hist, unsigned* sync, int dry) {
    __shared__ Rk2Lds lds;
    const int qi = XCD ? (int)(__builtin_amdgcn_s_getreg((15 << 11) | 20) & 7u) : 0;  // XCC_ID
    const int base = q.off[qi];
    const unsigned nt = (unsigned)(q.off[qi + 1] - base);
    const unsigned total = nt * (unsigned)nsteps;
    const unsigned one = threadIdx.x == 0 ? 1u : 0u;
    for (;;) {
        const unsigned item = __builtin_amdgcn_readfirstlane(
            __hip_atomic_fetch_add(sync + qi, one, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        if (item >= total) break;
        const int s = (int)(item / nt);
        const EnsTile tl = tiles[base + (int)(item - (unsigned)s * nt)];
        const EnsProblem& P = probs[tl.problem];
        unsigned* done = sync + kEnsSyncHead + tl.problem;
        if (s > 0) ens_wait(done, (unsigned)s * (unsigned)P.rk2_tiles, sync + kEnsFlag);
        unsigned long long* h = hist + (long long)tl.problem * kEnsHist;
        EulerArgs A = P.args;
        A.rows = P.rk2_rows;
        A.a = 0.5;
        A.b = 0.5;
        A.clk = nullptr;
        A.dt_dev = nullptr;
        A.dt = P.clk_cfl / __longlong_as_double((long long)ens_load_u64(h + (dry ? 0 : s)));
        const bool odd = !dry && (s & 1);
        A.Uin = odd ? P.S : P.U;
        A.Uout = odd ? P.U : P.S;
        A.Ustage = nullptr;
        A.wmax_bits = nullptr;
        const double w = euler_rk2_tile<ORDER, FLUX, false, false>(A, tl.tx, tl.ty, lds);
        if (!dry) ens_fold_max(h + s + 1, w);
        if constexpr (XCD) {
            // the tile's stores complete in the L2 every reader of them shares
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_s_waitcnt(0);
            __hip_atomic_fetch_add(done, one, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_fetch_add(done, one, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// clears the work-list heads, the completion counters and the wave-speed
// histories (slot 0 from each clock's current slot); the spin-limit flag
// stays until the host has read it
__global__ void ens_launch_reset_kernel(const EnsProblem* __restrict__ probs, int n, int nsteps, int slot,
                                        unsigned long long* hist, unsigned* sync) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kEnsSyncHead + n && i != kEnsFlag) sync[i] = 0u;
    if (i < n) {
        sync[kEnsSyncHead + n + i] = (unsigned)nsteps * (unsigned)probs[i].rk2_tiles;
        unsigned long long* h = hist + (long long)i * kEnsHist;
        h[0] = probs[i].clk->w[slot];
        for (int k = 1; k <= nsteps; ++k) h[k] = 0ull;
    }
}

// after a launch of nsteps steps from clock slot `slot`: each problem's clock
// takes the steps one by one (dt, t = t + dt, the step count), and its slots
// are left as a single problem's clock leaves them
__global__ void ens_launch_finish_kernel(const EnsProblem* __restrict__ probs, int n, int nsteps, int slot,
                                         const unsigned long long* hist) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Clock* c = probs[i].clk;
    const unsigned long long* h = hist + (long long)i * kEnsHist;
    double t = c->t, dt = c->dt;
    for (int k = 0; k < nsteps; ++k) {
        dt = c->cfl / __longlong_as_double((long long)h[k]);
        t = t + dt;
    }
    c->t = t;
    c->dt = dt;
    c->step += nsteps;
    const int cur = (slot + nsteps) % 3;
    c->w[cur] = h[nsteps];
    c->w[clock_next(cur)] = 0ull;
    c->w[clock_next(clock_next(cur))] = 0ull;
}

// two-stage SSP-RK2 stage launches (an odd step count's first step)
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
