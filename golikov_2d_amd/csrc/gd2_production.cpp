// gd2_production -- the reference's production run (its main, kernel.cu:3810-4469)
// over the libgd2 C ABI: the stellar-wind initial state (kernel.cu:3901-3941),
// a first-order phase of add2 steps and a MUSCL phase of Kernel_TVD steps, each
// step followed by funk_time (kernel.cu:4029-4174), Tecplot snapshots of either
// phase every --snap1 / --snap2 steps (kernel.cu:4085-4098, 4160-4173; the
// reference's schedule, 50000 and 100000, is the default; 0 turns one off) and
// the end-of-run files (kernel.cu:4213-4460) into --out. Host C++; the device
// work is the library's, one stream, no host synchronisation between steps.
//
//   gd2_production [--nx 3226] [--ny 1792] [--x-max 9] [--y-max 5] [--steps1 200000]
//                  [--steps2 400000] [--method 2] [--snap1 50000] [--snap2 100000]
//                  [--progress 10000] [--restart FILE] [--out .]
//
// Every --progress steps (0: never) the host waits for the device and prints
// the step counter and the model time TT, as the reference's loop prints its
// counter (kernel.cu:4033-4036, 4105-4108).
//
// Step counts are rounded down to even numbers (the reference advances two
// steps per loop pass, ping-ponging its two state buffers).
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gd2.h"

namespace {

// Header.h:25-37 and kernel.cu:12-20
constexpr double kPhi0 = 4.0, kM0 = 10.0, kMalf = 8.0, kRr0 = 0.07, kDdist = 0.3;
constexpr double kPi = 3.14159265358979323846;

struct Opts {
    int nx = 3226, ny = 1792;
    double x_max = 9.0, y_max = 5.0;
    long steps1 = 200000, steps2 = 400000, snap1 = 50000, snap2 = 100000, progress = 10000;
    int method = 2;
    std::string out = ".", restart;
};

bool parse(int argc, char** argv, Opts& o) {
    for (int a = 1; a < argc; ++a) {
        const std::string k = argv[a];
        if (a + 1 >= argc) return false;
        const char* v = argv[++a];
        if (k == "--nx") o.nx = std::atoi(v);
        else if (k == "--ny") o.ny = std::atoi(v);
        else if (k == "--x-max") o.x_max = std::atof(v);
        else if (k == "--y-max") o.y_max = std::atof(v);
        else if (k == "--steps1") o.steps1 = std::atol(v);
        else if (k == "--steps2") o.steps2 = std::atol(v);
        else if (k == "--snap1") o.snap1 = std::atol(v);
        else if (k == "--snap2") o.snap2 = std::atol(v);
        else if (k == "--progress") o.progress = std::atol(v);
        else if (k == "--method") o.method = std::atoi(v);
        else if (k == "--out") o.out = v;
        else if (k == "--restart") o.restart = v;
        else return false;
    }
    return o.nx >= 2 && o.ny >= 2 && o.x_max > 0.0 && o.y_max > 0.0 && o.steps1 >= 0 && o.steps2 >= 0;
}

#define HIPCHK(x)                                                                  \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                    \
            std::fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));    \
            std::exit(1);                                                          \
        }                                                                          \
    } while (0)
#define GD2CHK(x)                                                                          \
    do {                                                                                   \
        int r_ = (x);                                                                      \
        if (r_ != GD2_OK) {                                                                \
            std::fprintf(stderr, "%s returned %d (hip error %d)\n", #x, r_, gd2_last_hip_error()); \
            std::exit(1);                                                                  \
        }                                                                                  \
    } while (0)

}  // namespace

int main(int argc, char** argv) {
    Opts o;
    if (!parse(argc, argv, o)) {
        std::fprintf(stderr, "usage: %s [--nx N] [--ny M] [--x-max X] [--y-max Y] [--steps1 S] [--steps2 S] "
                             "[--method 2|3] [--snap1 S] [--snap2 S] [--progress S] [--restart FILE] [--out DIR]\n", argv[0]);
        return 2;
    }
    const int N = o.nx, M = o.ny;
    const long long K = (long long)N * M;
    gd2_grid g;
    std::memset(&g, 0, sizeof g);
    g.nx = N;
    g.ny = M;
    g.dx = o.x_max / N;            // kernel.cu:19-20
    g.dy = o.y_max / M;
    g.x0 = o.x_max / (2.0 * N);    // kernel.cu:16, 18
    g.y0 = o.y_max / (2.0 * M);
    g.gamma = 5.0 / 3.0;           // Header.h:4
    g.krit = 0.25;                 // Header.h:3
    g.bc[0] = GD2_BC_SYMMETRY;
    g.bc[1] = GD2_BC_OUTFLOW;
    g.bc[2] = GD2_BC_AXIS;
    g.bc[3] = GD2_BC_INFLOW;
    g.r_frozen = kDdist;
    g.axisym = 1;

    // initial state: the stellar wind inside 1.01 ddist and the quiescent
    // medium outside (kernel.cu:3901-3941); or, with --restart, a text state
    // with the wind imposed again inside 1.001 ddist (the reference's restart
    // branch, kernel.cu:3944-3996)
    std::vector<gd2_double2> s(K), u(K);
    std::vector<gd2_double3> b(K);
    const double gam = g.gamma;
    const double V_E = kPhi0, ro_E = 1.0 / (kPhi0 * kPhi0 * kRr0 * kRr0);
    const double P_E = ro_E * V_E * V_E / (gam * kM0 * kM0), B_E = std::sqrt(4.0 * kPi) / (kMalf * kRr0);
    if (!o.restart.empty()) GD2CHK(gd2_read_state_text(&g, o.restart.c_str(), s.data(), u.data(), b.data()));
    const double r_wind = kDdist * (o.restart.empty() ? 1.01 : 1.001);
    for (long long k = 0; k < K; ++k) {
        const int n = (int)(k % N), m = (int)(k / N);
        const double y = g.y0 + m * g.dy, x = g.x0 + n * g.dx;
        const double dist = std::sqrt(x * x + y * y);
        if (dist < r_wind) {
            s[k] = {ro_E / std::pow(dist / kRr0, 2.0), P_E * std::pow(kRr0 / dist, 2.0 * gam)};
            u[k] = {V_E * x / dist, V_E * y / dist};
            // B_E (rr_0 / r) sin(theta) in the azimuthal direction, whose
            // Cartesian image at phi = 0 is the third component (0 on the axis)
            const double BE = B_E / (dist / kRr0);
            const double bphi = std::fabs(y) < 0.000001 ? 0.0 : BE * std::sin(std::acos(x / dist));
            b[k] = {0.0, 0.0, bphi};
        } else if (o.restart.empty()) {
            s[k] = {1.0, 1.0 / gam};
            u[k] = {0.0, 0.0};
            b[k] = {0.0, 0.0, 0.0};
        }
    }

    gd2_double2 *ds, *du, *ds2, *du2;
    gd2_double3 *db, *db2;
    double *T, *T_do, *TT;
    int* di;
    const size_t b2 = sizeof(gd2_double2) * K, b3 = sizeof(gd2_double3) * K;
    HIPCHK(hipMalloc(&ds, b2));
    HIPCHK(hipMalloc(&du, b2));
    HIPCHK(hipMalloc(&ds2, b2));
    HIPCHK(hipMalloc(&du2, b2));
    HIPCHK(hipMalloc(&db, b3));
    HIPCHK(hipMalloc(&db2, b3));
    HIPCHK(hipMalloc(&T, sizeof(double)));
    HIPCHK(hipMalloc(&T_do, sizeof(double)));
    HIPCHK(hipMalloc(&TT, sizeof(double)));
    HIPCHK(hipMalloc(&di, sizeof(int)));
    HIPCHK(hipMemcpy(ds, s.data(), b2, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(du, u.data(), b2, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(db, b.data(), b3, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ds2, s.data(), b2, hipMemcpyHostToDevice));  // both buffer sets start equal (kernel.cu:4005-4014)
    HIPCHK(hipMemcpy(du2, u.data(), b2, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(db2, b.data(), b3, hipMemcpyHostToDevice));
    const double T0 = 10000000.0, Tdo0 = 0.00000001, TT0 = 0.0;  // kernel.cu:3860-3863
    const int i0 = 0;
    HIPCHK(hipMemcpy(T, &T0, sizeof T0, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(T_do, &Tdo0, sizeof Tdo0, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(TT, &TT0, sizeof TT0, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(di, &i0, sizeof i0, hipMemcpyHostToDevice));
    hipStream_t st;
    HIPCHK(hipStreamCreate(&st));

    auto to_host = [&]() {
        HIPCHK(hipStreamSynchronize(st));
        HIPCHK(hipMemcpy(s.data(), ds, b2, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(u.data(), du, b2, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(b.data(), db, b3, hipMemcpyDeviceToHost));
        double tt = 0.0;
        HIPCHK(hipMemcpy(&tt, TT, sizeof tt, hipMemcpyDeviceToHost));
        return tt;
    };
    double time_null = -1.0;
    auto phase = [&](int order, long steps, long snap, const char* tag, const char* prefix) {
        const auto t0 = std::chrono::steady_clock::now();
        for (long i = 0; i + 2 <= steps; i += 2) {
            if (order == 1) {
                GD2CHK(gd2_add2(&g, ds, du, db, ds2, du2, db2, T, T_do, (int)i, o.method, st));
                GD2CHK(gd2_funk_time(T, T_do, TT, di, st));
                GD2CHK(gd2_add2(&g, ds2, du2, db2, ds, du, db, T, T_do, (int)i, o.method, st));
            } else {
                GD2CHK(gd2_kernel_tvd(&g, ds, du, db, ds2, du2, db2, T, T_do, (int)i, o.method, st));
                GD2CHK(gd2_funk_time(T, T_do, TT, di, st));
                GD2CHK(gd2_kernel_tvd(&g, ds2, du2, db2, ds, du, db, T, T_do, (int)i, o.method, st));
            }
            GD2CHK(gd2_funk_time(T, T_do, TT, di, st));
            if (o.progress > 0 && i > 0 && i % o.progress == 0) {
                double tt = 0.0;
                HIPCHK(hipStreamSynchronize(st));
                HIPCHK(hipMemcpy(&tt, TT, sizeof tt, hipMemcpyDeviceToHost));
                std::printf("%s %ld TT = %g\n", tag, i, tt);
                std::fflush(stdout);
            }
            // snapshot after the pass whose counter i is a positive multiple of
            // snap (kernel.cu:4085-4098, 4160-4173): the first phase without a
            // solution time, the second with TT minus its first snapshot's TT
            if (snap > 0 && i > 0 && i % snap == 0) {
                const double tt = to_host();
                const std::string name = o.out + "/" + prefix + std::to_string(i) + ".txt";
                if (order == 1) {
                    GD2CHK(gd2_write_tecplot(&g, s.data(), u.data(), b.data(), 3, nullptr, name.c_str()));
                } else {
                    if (time_null < 0.0) time_null = tt;
                    const double ts = tt - time_null;
                    GD2CHK(gd2_write_tecplot(&g, s.data(), u.data(), b.data(), 3, &ts, name.c_str()));
                }
            }
        }
        HIPCHK(hipStreamSynchronize(st));
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        const long done = steps - steps % 2;
        std::printf("%s: %ld steps in %.3f s (%.3f Gcell-updates/s incl. snapshots)\n", tag, done, sec,
                    sec > 0.0 ? (double)K * done / sec / 1e9 : 0.0);
    };
    phase(1, o.steps1, o.snap1, "add2", "02_12_");
    phase(2, o.steps2, o.snap2, "kernel_tvd", "11_01_");

    const double tt = to_host();
    std::printf("TT = %g\n", tt);
    const std::string d = o.out + "/";
    GD2CHK(gd2_write_state_text(&g, s.data(), u.data(), b.data(), (d + "instable_for_inner_b3.txt").c_str()));
    GD2CHK(gd2_write_state_b3(&g, s.data(), u.data(), b.data(), (d + "save_b3.dat").c_str()));
    GD2CHK(gd2_write_tecplot_derived(&g, s.data(), u.data(), b.data(), 3, (d + "param_for_texplot_mini.txt").c_str()));
    GD2CHK(gd2_write_line_cuts(&g, s.data(), u.data(), b.data(), (d + "param_y=0.txt").c_str(),
                               (d + "param_x=0.txt").c_str()));
    gd2_diag dg;
    GD2CHK(gd2_get_diagnostics(&dg, 0));
    std::printf("clamps: rho %lld, p %lld\n", dg.clamp_rho, dg.clamp_p);
    HIPCHK(hipStreamDestroy(st));
    for (void* p : {(void*)ds, (void*)du, (void*)ds2, (void*)du2, (void*)db, (void*)db2, (void*)T, (void*)T_do,
                    (void*)TT, (void*)di})
        HIPCHK(hipFree(p));
    return 0;
}
