// gd2_torch.cpp -- LibTorch autograd Function over the libgd2 C ABI: an
// SSP-RK2 rollout of the Euler HLLC scheme or the MHD HLLC-MHD / HLLD scheme
// (MUSCL or first order) with a prescribed dt schedule, differentiable with
// respect to the initial conserved state.
//
// Forward: nsteps SSP-RK2 steps (gd2_ssprk2_step, or two gd2_stage launches
// where the library prefers them: small grids), keeping a copy of the state
// every `every` steps, and every step's states of the last segment. Backward: per segment, from the last one down, the segment's
// states (each step's input U^n and its stage-1 state U^(1)) are recomputed
// from its checkpoint on a second stream while the previous segment's
// adjoint runs, then the step adjoint
//     lam1 = lam/2 + (dt/2) J_L(U^(1))^T lam
//     lam  = lam/2 + lam1 + dt J_L(U^n)^T lam1
// runs backwards through the segment (two gd2_stage_vjp launches per step).
// The dt schedule is data, not a function of the state: derivatives are
// those of the map with the schedule held fixed.
//
// No reference counterpart (the reference has no adjoint).
#include <hip/hip_runtime_api.h>
#include <torch/extension.h>

#include <string>
#include <utility>
#include <vector>

#include "gd2.h"

namespace {

using torch::Tensor;
using torch::autograd::AutogradContext;
using torch::autograd::variable_list;

struct Setup {
    gd2_grid g;
    gd2_scheme sc;
    int nfields;
    bool one_pass;  // SSP-RK2 steps through gd2_ssprk2_step
};

Setup make_setup(const std::vector<double>& geom, const std::vector<int64_t>& bc,
                 const std::vector<int64_t>& scheme) {
    TORCH_CHECK(geom.size() == 9, "geom = (nx, ny, dx, dy, x0, y0, gamma, krit, cfl)");
    TORCH_CHECK(bc.size() == 4, "bc = (west, east, south, north)");
    TORCH_CHECK(scheme.size() == 3, "scheme = (physics, flux, order)");
    Setup s{};
    s.g.nx = static_cast<int>(geom[0]);
    s.g.ny = static_cast<int>(geom[1]);
    s.g.dx = geom[2];
    s.g.dy = geom[3];
    s.g.x0 = geom[4];
    s.g.y0 = geom[5];
    s.g.gamma = geom[6];
    s.g.krit = geom[7];
    for (int i = 0; i < 4; ++i) s.g.bc[i] = static_cast<int>(bc[i]);
    s.sc.physics = static_cast<int>(scheme[0]);
    s.sc.flux = static_cast<int>(scheme[1]);
    s.sc.order = static_cast<int>(scheme[2]);
    s.sc.integrator = GD2_SSPRK2;
    s.sc.cfl = geom[8];
    const bool euler = s.sc.physics == GD2_EULER && s.sc.flux == GD2_FLUX_HLLC;
    const bool mhd = s.sc.physics == GD2_MHD && (s.sc.flux == GD2_FLUX_HLLC_MHD || s.sc.flux == GD2_FLUX_HLLD);
    TORCH_CHECK(euler || mhd, "differentiable rollouts cover Euler + HLLC and MHD + HLLC-MHD / HLLD");
    s.nfields = gd2_nfields(s.sc.physics);
    const int pref = gd2_one_pass_preferred(&s.g, &s.sc);
    TORCH_CHECK(pref >= 0, "gd2_one_pass_preferred failed");
    s.one_pass = pref == 1;
    return s;
}

void check(int rc, const char* what) {
    TORCH_CHECK(rc == 0, what, " returned ", rc, " (hip error ", gd2_last_hip_error(), ")");
}

void check_state(const Setup& s, const Tensor& U, const char* name) {
    TORCH_CHECK(U.is_cuda(), name, " must be a device tensor");
    TORCH_CHECK(U.scalar_type() == torch::kFloat64, name, " must be float64");
    TORCH_CHECK(U.is_contiguous(), name, " must be contiguous");
    TORCH_CHECK(U.dim() == 3 && U.size(0) == s.nfields && U.size(1) == s.g.ny && U.size(2) == s.g.nx, name,
                " must have shape (nfields, ny, nx)");
}

double* dp(const Tensor& t) { return t.data_ptr<double>(); }

void stage1(const Setup& s, const Tensor& U, Tensor& U1, double dt, void* st) {
    check(gd2_stage(&s.g, &s.sc, dp(U), nullptr, dp(U1), 0.0, 1.0, dt, nullptr, nullptr, st), "gd2_stage");
}

void stage2(const Setup& s, const Tensor& U1, const Tensor& Un, Tensor& Uout, double dt, void* st) {
    check(gd2_stage(&s.g, &s.sc, dp(U1), dp(Un), dp(Uout), 0.5, 0.5, dt, nullptr, nullptr, st), "gd2_stage");
}

// One SSP-RK2 step Un -> Uout (Uout != Un); U1 (may be null) receives the
// stage-1 state. One pass where the library prefers it, else two stage
// launches (same bits); `tmp` is the stage state of the latter when U1 is null.
void step(const Setup& s, const Tensor& Un, Tensor& Uout, Tensor* U1, Tensor& tmp, double dt, void* st) {
    if (s.one_pass) {
        check(gd2_ssprk2_step(&s.g, &s.sc, dp(Un), dp(Uout), U1 ? dp(*U1) : nullptr, dt, nullptr, nullptr, st),
              "gd2_ssprk2_step");
        return;
    }
    Tensor& S = U1 ? *U1 : tmp;
    stage1(s, Un, S, dt, st);
    stage2(s, S, Un, Uout, dt, st);
}

// One step adjoint in place: lam <- (d U^{n+1} / d U^n)^T lam.
void step_vjp(const Setup& s, const Tensor& Un, const Tensor& U1, Tensor& lam, Tensor& lam1, Tensor& ws,
              double dt, void* st) {
    check(gd2_stage_vjp(&s.g, &s.sc, dp(U1), dp(lam), dp(lam), 0.5, nullptr, 0.0, 0.5 * dt, dp(lam1),
                        ws.data_ptr(), st),
          "gd2_stage_vjp");
    check(gd2_stage_vjp(&s.g, &s.sc, dp(Un), dp(lam1), dp(lam), 0.5, dp(lam1), 1.0, dt, dp(lam), ws.data_ptr(),
                        st),
          "gd2_stage_vjp");
}

std::vector<double> schedule(const Tensor& dts) {
    Tensor d = dts.detach().to(torch::kCPU, torch::kFloat64).contiguous();
    TORCH_CHECK(d.dim() == 1, "dts must be a 1-D tensor of step sizes");
    const double* p = d.data_ptr<double>();
    return std::vector<double>(p, p + d.numel());
}

// Forward rollout; returns the final state and fills the checkpoints (U^n of
// every `every`-th step). The last segment also keeps each step's U^n and
// U^(1) (in last_un / last_u1): the reverse sweep starts there, so its
// recompute would have nothing to overlap with; the forward computes these
// states anyway and only writes them to separate buffers.
Tensor run_forward(const Setup& s, const Tensor& U0, const std::vector<double>& dt, int64_t every, void* st,
                   std::vector<Tensor>* ckpt, std::vector<Tensor>* last_un, std::vector<Tensor>* last_u1) {
    Tensor U = U0.detach().clone();
    Tensor N = torch::empty_like(U);  // next state
    Tensor S = torch::empty_like(U);  // stage state of a two-launch step
    const int64_t n = static_cast<int64_t>(dt.size());
    const int64_t k_last = n > 0 ? ((n - 1) / every) * every : 0;
    const bool keep = ckpt && last_un && last_u1;
    for (int64_t k = 0; k < n; ++k) {
        if (ckpt && k % every == 0) ckpt->push_back(U.clone());
        if (keep && k >= k_last) {
            if (k == k_last) last_un->push_back(ckpt->back());
            const Tensor& un = last_un->back();
            Tensor u1 = torch::empty_like(U);
            Tensor next = torch::empty_like(U);
            step(s, un, next, &u1, S, dt[k], st);
            last_u1->push_back(u1);
            if (k + 1 < n)
                last_un->push_back(next);
            else
                U = next;
            continue;
        }
        step(s, U, N, nullptr, S, dt[k], st);
        std::swap(U, N);
    }
    return U;
}

void hip_check(hipError_t e, const char* what) { TORCH_CHECK(e == hipSuccess, what, ": ", hipGetErrorString(e)); }

// A second stream (default priority) and its events, created once per
// process: the backward sweep recomputes the next segment's states on it
// while the caller's stream runs the current segment's adjoint.
struct SideStream {
    hipStream_t stream = nullptr;
    hipEvent_t start = nullptr, rec[2] = {nullptr, nullptr}, vjp[2] = {nullptr, nullptr};
    SideStream() {
        hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreateWithFlags");
        hip_check(hipEventCreateWithFlags(&start, hipEventDisableTiming), "hipEventCreateWithFlags");
        for (int k = 0; k < 2; ++k) {
            hip_check(hipEventCreateWithFlags(&rec[k], hipEventDisableTiming), "hipEventCreateWithFlags");
            hip_check(hipEventCreateWithFlags(&vjp[k], hipEventDisableTiming), "hipEventCreateWithFlags");
        }
    }
};

SideStream& side_stream() {
    static SideStream s;
    return s;
}

// Reverse sweep from the checkpoints: returns (d U_T / d U_0)^T lamT.
// Segment seg's states U^n, U^(1) are recomputed from its checkpoint into one
// of two buffer sets on the side stream; the adjoint of segment seg (caller's
// stream) overlaps the recompute of segment seg-1 into the other set.
Tensor run_backward(const Setup& s, const std::vector<Tensor>& ckpt, const std::vector<double>& dt, int64_t every,
                    const Tensor& lamT, void* st, const std::vector<Tensor>& last_un,
                    const std::vector<Tensor>& last_u1) {
    const int64_t n = static_cast<int64_t>(dt.size());
    const int64_t nseg = static_cast<int64_t>(ckpt.size());
    Tensor lam = lamT.to(torch::kFloat64).contiguous().clone();
    Tensor lam1 = torch::empty_like(lam);
    const long long wsb = gd2_vjp_workspace_bytes(&s.g, &s.sc);
    TORCH_CHECK(wsb > 0, "gd2_vjp_workspace_bytes failed");
    Tensor ws = torch::empty({(wsb + 7) / 8}, lam.options());
    const int64_t seg_len = std::min<int64_t>(every, n);
    std::vector<Tensor> Un[2], U1[2];
    for (int k = 0; k < 2 && k < nseg; ++k)
        for (int64_t i = 0; i < seg_len; ++i) {
            Un[k].push_back(torch::empty_like(lam));
            U1[k].push_back(torch::empty_like(lam));
        }
    SideStream& ss = side_stream();
    hipStream_t main = reinterpret_cast<hipStream_t>(st);
    // the checkpoints and lamT come from work on the caller's stream
    hip_check(hipEventRecord(ss.start, main), "hipEventRecord");
    hip_check(hipStreamWaitEvent(ss.stream, ss.start, 0), "hipStreamWaitEvent");
    const size_t bytes = static_cast<size_t>(lam.numel()) * sizeof(double);
    auto recompute = [&](int64_t seg) {
        const int set = static_cast<int>(seg & 1);
        const int64_t k0 = seg * every, m = std::min(n, k0 + every) - k0;
        hip_check(hipMemcpyAsync(dp(Un[set][0]), dp(ckpt[seg]), bytes, hipMemcpyDeviceToDevice, ss.stream),
                  "hipMemcpyAsync");
        for (int64_t i = 0; i < m; ++i) {
            if (i + 1 < m)
                step(s, Un[set][i], Un[set][i + 1], &U1[set][i], U1[set][i], dt[k0 + i], ss.stream);
            else
                stage1(s, Un[set][i], U1[set][i], dt[k0 + i], ss.stream);
        }
        hip_check(hipEventRecord(ss.rec[set], ss.stream), "hipEventRecord");
    };
    // the last segment's states come from the forward when it kept them
    const bool have_last = nseg > 0 && static_cast<int64_t>(last_un.size()) == n - (nseg - 1) * every &&
                           last_u1.size() == last_un.size();
    if (nseg > 0 && !have_last) recompute(nseg - 1);
    for (int64_t seg = nseg - 1; seg >= 0; --seg) {
        const int set = static_cast<int>(seg & 1);
        const bool kept = have_last && seg == nseg - 1;
        if (seg >= 1) {
            // segment seg-1 reuses the buffers of segment seg+1: after its adjoint
            const bool seg1_used_set = seg + 1 < nseg && !(have_last && seg + 1 == nseg - 1);
            if (seg1_used_set) hip_check(hipStreamWaitEvent(ss.stream, ss.vjp[set ^ 1], 0), "hipStreamWaitEvent");
            recompute(seg - 1);
        }
        const int64_t k0 = seg * every, m = std::min(n, k0 + every) - k0;
        if (kept) {
            for (int64_t i = m - 1; i >= 0; --i) step_vjp(s, last_un[i], last_u1[i], lam, lam1, ws, dt[k0 + i], st);
        } else {
            hip_check(hipStreamWaitEvent(main, ss.rec[set], 0), "hipStreamWaitEvent");
            for (int64_t i = m - 1; i >= 0; --i) step_vjp(s, Un[set][i], U1[set][i], lam, lam1, ws, dt[k0 + i], st);
        }
        hip_check(hipEventRecord(ss.vjp[set], main), "hipEventRecord");
    }
    // every side-stream launch was waited for by the caller's stream above, so
    // the buffers may be released in the caller's stream order
    return lam;
}

struct Rollout : public torch::autograd::Function<Rollout> {
    static Tensor forward(AutogradContext* ctx, Tensor U0, Tensor dts, std::vector<double> geom,
                          std::vector<int64_t> bc, std::vector<int64_t> scheme, int64_t every, int64_t stream) {
        const Setup s = make_setup(geom, bc, scheme);
        check_state(s, U0, "U0");
        TORCH_CHECK(every >= 1, "checkpoint interval must be >= 1");
        const std::vector<double> dt = schedule(dts);
        std::vector<Tensor> ckpt, last_un, last_u1;
        Tensor UT = run_forward(s, U0, dt, every, reinterpret_cast<void*>(stream), &ckpt, &last_un, &last_u1);
        ctx->saved_data["ckpt"] = ckpt;
        ctx->saved_data["last_un"] = last_un;
        ctx->saved_data["last_u1"] = last_u1;
        ctx->saved_data["dt"] = dt;
        ctx->saved_data["geom"] = geom;
        ctx->saved_data["bc"] = bc;
        ctx->saved_data["scheme"] = scheme;
        ctx->saved_data["every"] = every;
        ctx->saved_data["stream"] = stream;
        return UT;
    }

    static variable_list backward(AutogradContext* ctx, variable_list grads) {
        const Setup s = make_setup(ctx->saved_data["geom"].toDoubleVector(), ctx->saved_data["bc"].toIntVector(),
                                   ctx->saved_data["scheme"].toIntVector());
        const std::vector<Tensor> ckpt = ctx->saved_data["ckpt"].toTensorVector();
        const std::vector<double> dt = ctx->saved_data["dt"].toDoubleVector();
        const int64_t every = ctx->saved_data["every"].toInt();
        void* st = reinterpret_cast<void*>(ctx->saved_data["stream"].toInt());
        Tensor g = grads[0];
        if (!g.defined()) return {Tensor(), Tensor(), Tensor(), Tensor(), Tensor(), Tensor(), Tensor()};
        Tensor lam0 = run_backward(s, ckpt, dt, every, g, st, ctx->saved_data["last_un"].toTensorVector(),
                                   ctx->saved_data["last_u1"].toTensorVector());
        return {lam0, Tensor(), Tensor(), Tensor(), Tensor(), Tensor(), Tensor()};
    }
};

Tensor rollout(Tensor U0, Tensor dts, std::vector<double> geom, std::vector<int64_t> bc,
               std::vector<int64_t> scheme, int64_t every, int64_t stream) {
    return Rollout::apply(U0, dts, geom, bc, scheme, every, stream);
}

// The same two sweeps without the autograd graph: returns (U_T, lam_0).
std::vector<Tensor> rollout_vjp(Tensor U0, Tensor dts, Tensor lamT, std::vector<double> geom,
                                std::vector<int64_t> bc, std::vector<int64_t> scheme, int64_t every,
                                int64_t stream) {
    const Setup s = make_setup(geom, bc, scheme);
    check_state(s, U0, "U0");
    check_state(s, lamT, "lamT");
    TORCH_CHECK(every >= 1, "checkpoint interval must be >= 1");
    const std::vector<double> dt = schedule(dts);
    std::vector<Tensor> ckpt, last_un, last_u1;
    void* st = reinterpret_cast<void*>(stream);
    Tensor UT = run_forward(s, U0, dt, every, st, &ckpt, &last_un, &last_u1);
    Tensor lam0 = run_backward(s, ckpt, dt, every, lamT, st, last_un, last_u1);
    return {UT, lam0};
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
    m.doc() = "Differentiable SSP-RK2 rollouts of the Euler and MHD MUSCL schemes on libgd2";
    m.def("rollout", &rollout,
          "U_T = rollout(U0, dts, geom, bc, (physics, flux, order), every, stream) (autograd-aware)");
    m.def("rollout_vjp", &rollout_vjp,
          "(U_T, lam_0) = rollout_vjp(U0, dts, lamT, geom, bc, (physics, flux, order), every, stream)");
}
