// On-disk formats of the reference's host driver, host side of libgd2:
//   text state / restart  "x y rho p vx vy Bx By Bz" per cell   (kernel.cu:4257-4260 writer,
//                                                                kernel.cu:3944-3969 reader)
//   binary save_b3.dat    x, y, rho, p, vx, vy, Bz doubles     (kernel.cu:4261-4267)
//   Tecplot FEPOINT zone  every nmin-th cell, quadrilaterals   (kernel.cu:3682-3808)
//   end-of-run derived    Tecplot with PP, Fx, Fy, zav; line  (kernel.cu:4231-4460)
//                         cuts along the first row / column
//
// Numbers are written as the reference's std::ofstream writes them: default
// float format at precision 6, i.e. printf "%g" (std::to_chars with
// chars_format::general and precision 6 is specified as that conversion).
// Rows are formatted in parallel into per-thread buffers and written in order.
#include <algorithm>
#include <atomic>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "gd2.h"

#pragma STDC FP_CONTRACT OFF

namespace {

constexpr int kMaxLine = 9 * 32 + 2;  // 9 numbers of at most 31 characters each

int io_threads() {
    const char* e = std::getenv("GD2_IO_THREADS");
    int n = e ? std::atoi(e) : (int)std::thread::hardware_concurrency();
    return std::max(1, std::min(n, 16));
}

inline char* put_g(char* p, double v) {
    auto r = std::to_chars(p, p + 32, v, std::chars_format::general, 6);
    return r.ptr;
}

inline char* put_i(char* p, long long v) {
    auto r = std::to_chars(p, p + 24, v);
    return r.ptr;
}

// Runs body(lo, hi, out) over [0, n) split into contiguous chunks, one per
// thread, then writes the chunks' text in order. Returns 0 or GD2_ERR_ARG on
// a write failure.
template <class Body>
int parallel_write(std::FILE* f, long long n, long long bytes_per_item, Body body) {
    const int T = (int)std::min<long long>(io_threads(), std::max<long long>(1, n / 4096));
    std::vector<std::string> buf(T);
    std::vector<std::thread> pool;
    auto run = [&](int t) {
        const long long lo = n * t / T, hi = n * (t + 1) / T;
        std::string& s = buf[t];
        s.resize((size_t)((hi - lo) * bytes_per_item));
        char* p = body(lo, hi, s.data());
        s.resize((size_t)(p - s.data()));
    };
    for (int t = 1; t < T; ++t) pool.emplace_back(run, t);
    run(0);
    for (auto& th : pool) th.join();
    for (int t = 0; t < T; ++t)
        if (!buf[t].empty() && std::fwrite(buf[t].data(), 1, buf[t].size(), f) != buf[t].size()) return GD2_ERR_ARG;
    return GD2_OK;
}

bool grid_shape_ok(const gd2_grid* g) { return g && g->nx >= 1 && g->ny >= 1; }

}  // namespace

// Derived fields of the reference's end-of-run output (kernel.cu:4283-4446):
// PP = p + Bz^2/8pi, the forces of the azimuthal field Fx = Bz dBz/dx / 4pi,
// Fy = (-Bz^2/y - Bz dBz/dy) / 4pi (central differences inside, forward at the
// first column / row, 0 on the last), the vorticity zav.
namespace {
constexpr double kPiRef = 3.14159265358979323846;  // Header.h:18

struct Derived {
    double Fx, Fy, zav;
};

// kernel.cu:4400-4436 (cdiff_y: central y difference where m > 0, as the
// Tecplot file; the x = 0 / y = 0 line cuts take the forward one)
inline Derived derived_at(const gd2_grid* g, const gd2_double2* u, const gd2_double3* b, int n, int m,
                          bool cdiff_y) {
    const int N = g->nx, M = g->ny;
    const long long k = (long long)m * N + n;
    Derived d = {0.0, 0.0, 0.0};
    if (m < M - 1 && n < N - 1) {
        const double y = g->y0 + m * g->dy;
        const long long kk = k + N, kk2 = k - N, e = k + 1, w = k - 1;
        const double dAz = n > 0 ? (b[e].z - b[w].z) / (2.0 * g->dx) : (b[e].z - b[k].z) / (g->dx);
        const double dAr = (cdiff_y && m > 0) ? (b[kk].z - b[kk2].z) / (2.0 * g->dy) : (b[kk].z - b[k].z) / (g->dy);
        d.Fx = b[k].z * dAz / (4.0 * kPiRef);
        d.Fy = (-b[k].z * b[k].z / y - b[k].z * dAr) / (4.0 * kPiRef);
        if (n > 0 && m > 0) d.zav = (u[e].y - u[w].y) / (2.0 * g->dx) - (u[kk].x - u[kk2].x) / (2.0 * g->dy);
    }
    return d;
}

inline void mach_temp(const gd2_grid* g, const gd2_double2& s, const gd2_double2& u, double& Max, double& Temp) {
    Max = 0.0;
    Temp = 0.0;
    if (s.x > 0.0) {
        Max = std::sqrt((u.x * u.x + u.y * u.y) / (g->gamma * s.y / s.x));
        Temp = s.y / s.x;
    }
}
}  // namespace

extern "C" {

int gd2_write_state_text(const gd2_grid* g, const gd2_double2* s, const gd2_double2* u, const gd2_double3* b,
                         const char* path) {
    if (!grid_shape_ok(g) || !s || !u || !path) return GD2_ERR_ARG;
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return GD2_ERR_ARG;
    const int nx = g->nx;
    const double x0 = g->x0, y0 = g->y0, dx = g->dx, dy = g->dy;
    const long long K = (long long)nx * g->ny;
    int rc = parallel_write(f, K, kMaxLine, [&](long long lo, long long hi, char* p) {
        for (long long k = lo; k < hi; ++k) {
            const int n = (int)(k % nx), m = (int)(k / nx);
            const double y = y0 + m * dy, x = x0 + n * dx;
            p = put_g(p, x); *p++ = ' ';
            p = put_g(p, y); *p++ = ' ';
            p = put_g(p, s[k].x); *p++ = ' ';
            p = put_g(p, s[k].y); *p++ = ' ';
            p = put_g(p, u[k].x); *p++ = ' ';
            p = put_g(p, u[k].y); *p++ = ' ';
            p = put_g(p, b ? b[k].x : 0.0); *p++ = ' ';
            p = put_g(p, b ? b[k].y : 0.0); *p++ = ' ';
            p = put_g(p, b ? b[k].z : 0.0); *p++ = '\n';
        }
        return p;
    });
    if (std::fclose(f) != 0) rc = GD2_ERR_ARG;
    return rc;
}

int gd2_read_state_text(const gd2_grid* g, const char* path, gd2_double2* s, gd2_double2* u, gd2_double3* b) {
    if (!grid_shape_ok(g) || !s || !u || !path) return GD2_ERR_ARG;
    std::FILE* f = std::fopen(path, "rb");
    if (!f) return GD2_ERR_ARG;
    std::string text;
    if (std::fseek(f, 0, SEEK_END) == 0) {
        const long sz = std::ftell(f);
        if (sz > 0) {
            text.resize((size_t)sz);
            std::rewind(f);
            if (std::fread(text.data(), 1, text.size(), f) != text.size()) text.clear();
        }
    }
    std::fclose(f);
    const long long K = (long long)g->nx * g->ny;
    const long long need = 9 * K;
    const char* base = text.data();
    const size_t len = text.size();
    auto is_ws = [](char c) { return c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; };
    // split at whitespace into T chunks; count each chunk's tokens, then parse
    const int T = (int)std::min<long long>(io_threads(), std::max<long long>(1, (long long)len / (1 << 20)));
    std::vector<size_t> cut(T + 1, len);
    cut[0] = 0;
    for (int t = 1; t < T; ++t) {
        size_t c = len * t / T;
        while (c < len && !is_ws(base[c])) ++c;
        cut[t] = std::max(c, cut[t - 1]);
    }
    std::vector<long long> count(T, 0);
    auto count_tokens = [&](int t) {
        long long c = 0;
        bool in = false;
        for (size_t i = cut[t]; i < cut[t + 1]; ++i) {
            const bool w = is_ws(base[i]);
            if (!w && !in) ++c;
            in = !w;
        }
        count[t] = c;
    };
    {
        std::vector<std::thread> pool;
        for (int t = 1; t < T; ++t) pool.emplace_back(count_tokens, t);
        count_tokens(0);
        for (auto& th : pool) th.join();
    }
    std::vector<long long> first(T + 1, 0);
    for (int t = 0; t < T; ++t) first[t + 1] = first[t] + count[t];
    if (first[T] < need) return GD2_ERR_ARG;
    std::atomic<int> bad{0};
    auto parse = [&](int t) {
        long long idx = first[t];
        size_t i = cut[t];
        const size_t end = cut[t + 1];
        while (i < end && idx < need) {
            while (i < end && is_ws(base[i])) ++i;
            if (i >= end) break;
            size_t j = i;
            while (j < end && !is_ws(base[j])) ++j;
            double v = 0.0;
            const char* tb = base + i;
            if (*tb == '+') ++tb;  // operator>> accepts a leading '+', from_chars does not
            auto r = std::from_chars(tb, base + j, v);
            if (r.ec != std::errc() || r.ptr != base + j) {
                bad.store(1);
                return;
            }
            const long long k = idx / 9;
            switch ((int)(idx % 9)) {
                case 2: s[k].x = v; break;
                case 3: s[k].y = v; break;
                case 4: u[k].x = v; break;
                case 5: u[k].y = v; break;
                case 6: if (b) b[k].x = v; break;
                case 7: if (b) b[k].y = v; break;
                case 8: if (b) b[k].z = v; break;
                default: break;  // columns 0, 1: the cell's coordinates
            }
            ++idx;
            i = j;
        }
    };
    {
        std::vector<std::thread> pool;
        for (int t = 1; t < T; ++t) pool.emplace_back(parse, t);
        parse(0);
        for (auto& th : pool) th.join();
    }
    return bad.load() ? GD2_ERR_ARG : GD2_OK;
}

int gd2_write_state_b3(const gd2_grid* g, const gd2_double2* s, const gd2_double2* u, const gd2_double3* b,
                       const char* path) {
    if (!grid_shape_ok(g) || !s || !u || !b || !path) return GD2_ERR_ARG;
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return GD2_ERR_ARG;
    const int nx = g->nx;
    const double x0 = g->x0, y0 = g->y0, dx = g->dx, dy = g->dy;
    const long long K = (long long)nx * g->ny;
    int rc = parallel_write(f, K, 7 * sizeof(double), [&](long long lo, long long hi, char* p) {
        for (long long k = lo; k < hi; ++k) {
            const int n = (int)(k % nx), m = (int)(k / nx);
            const double v[7] = {x0 + n * dx, y0 + m * dy, s[k].x, s[k].y, u[k].x, u[k].y, b[k].z};
            std::memcpy(p, v, sizeof v);
            p += sizeof v;
        }
        return p;
    });
    if (std::fclose(f) != 0) rc = GD2_ERR_ARG;
    return rc;
}

int gd2_write_tecplot(const gd2_grid* g, const gd2_double2* s, const gd2_double2* u, const gd2_double3* b, int nmin,
                      const double* solution_time, const char* path) {
    if (!grid_shape_ok(g) || !s || !u || !path || nmin < 1) return GD2_ERR_ARG;
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return GD2_ERR_ARG;
    const int nx = g->nx, ny = g->ny;
    const long long nn = (nx + nmin - 1) / nmin, mm = (ny + nmin - 1) / nmin;
    // the writer's own coordinates: cell centres stretched over [x_min, x_max]
    // with an (N - 1) spacing (kernel.cu:3741-3742), x_max the domain's edge
    const double x_min = g->x0, y_min = g->y0;
    const double x_max = g->x0 + (nx - 0.5) * g->dx, y_max = g->y0 + (ny - 0.5) * g->dy;
    const double gam = g->gamma;
    std::string head = b ? "TITLE = \"HP\"  VARIABLES = \"X\", \"Y\", \"Ro\", \"P\", \"Vx\", \"Vy\", \"Bx\", \"By\", \"Bz\", "
                           "\"Max\", \"T\", ZONE T = \"HP\", N = "
                         : "TITLE = \"HP\"  VARIABLES = \"X\", \"Y\", \"Ro\", \"P\", \"Vx\", \"Vy\", \"Max\", \"T\", "
                           "ZONE T = \"HP\", N = ";
    char num[64];
    head.append(num, put_i(num, nn * mm));
    head += " , E = ";
    head.append(num, put_i(num, (nn - 1) * (mm - 1)));
    head += ", F = FEPOINT, ET = quadrilateral";
    if (solution_time) {
        head += ", SOLUTIONTIME = ";
        head.append(num, put_g(num, *solution_time));
    }
    head += '\n';
    int rc = std::fwrite(head.data(), 1, head.size(), f) == head.size() ? GD2_OK : GD2_ERR_ARG;
    if (rc == GD2_OK)
        rc = parallel_write(f, nn * mm, 11 * 32 + 2, [&](long long lo, long long hi, char* p) {
            for (long long q = lo; q < hi; ++q) {
                const int n = (int)(q % nn) * nmin, m = (int)(q / nn) * nmin;
                const long long k = (long long)m * nx + n;
                const double y = ny > 1 ? y_min + m * (y_max - y_min) / (ny - 1) : y_min;
                const double x = nx > 1 ? x_min + n * (x_max - x_min) / (nx - 1) : x_min;
                double Max = 0.0, Temp = 0.0;
                if (s[k].x > 0.0) {
                    Max = std::sqrt((u[k].x * u[k].x + u[k].y * u[k].y) / (gam * s[k].y / s[k].x));
                    Temp = s[k].y / s[k].x;
                }
                p = put_g(p, x); *p++ = ' ';
                p = put_g(p, y); *p++ = ' ';
                p = put_g(p, s[k].x); *p++ = ' ';
                p = put_g(p, s[k].y); *p++ = ' ';
                p = put_g(p, u[k].x); *p++ = ' ';
                p = put_g(p, u[k].y); *p++ = ' ';
                if (b) {
                    p = put_g(p, b[k].x); *p++ = ' ';
                    p = put_g(p, b[k].y); *p++ = ' ';
                    p = put_g(p, b[k].z); *p++ = ' ';
                }
                p = put_g(p, Max); *p++ = ' ';
                p = put_g(p, Temp); *p++ = '\n';
            }
            return p;
        });
    // quadrilaterals, 1-based node numbers (kernel.cu:3756-3764)
    if (rc == GD2_OK && nn > 1 && mm > 1)
        rc = parallel_write(f, (nn - 1) * (mm - 1), 4 * 21 + 4, [&](long long lo, long long hi, char* p) {
            for (long long q = lo; q < hi; ++q) {
                const long long n = q % (nn - 1), m = q / (nn - 1);
                p = put_i(p, m * nn + n + 1); *p++ = ' ';
                p = put_i(p, m * nn + n + 2); *p++ = ' ';
                p = put_i(p, (m + 1) * nn + n + 2); *p++ = ' ';
                p = put_i(p, (m + 1) * nn + n + 1); *p++ = '\n';
            }
            return p;
        });
    if (std::fclose(f) != 0) rc = GD2_ERR_ARG;
    return rc;
}


int gd2_write_tecplot_derived(const gd2_grid* g, const gd2_double2* s, const gd2_double2* u, const gd2_double3* b,
                              int nmin, const char* path) {
    if (!grid_shape_ok(g) || !s || !u || !b || !path || nmin < 1) return GD2_ERR_ARG;
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return GD2_ERR_ARG;
    const int nx = g->nx, ny = g->ny;
    const long long nn = (nx + nmin - 1) / nmin, mm = (ny + nmin - 1) / nmin;
    std::string head = "TITLE = \"HP\"  VARIABLES = \"X\", \"Y\", \"Ro\", \"P\", \"Vx\", \"Vy\", \"Bx\", \"By\", \"Bz\", "
                       "\"Max\", \"T\", \"PP\", \"Fx\", \"Fy\", \"zav\",  ZONE T = \"HP\", N = ";
    char num[64];
    head.append(num, put_i(num, nn * mm));
    head += " , E = ";
    head.append(num, put_i(num, (nn - 1) * (mm - 1)));
    head += ", F = FEPOINT, ET = quadrilateral\n";
    int rc = std::fwrite(head.data(), 1, head.size(), f) == head.size() ? GD2_OK : GD2_ERR_ARG;
    if (rc == GD2_OK)
        rc = parallel_write(f, nn * mm, 15 * 32 + 2, [&](long long lo, long long hi, char* p) {
            for (long long q = lo; q < hi; ++q) {
                const int n = (int)(q % nn) * nmin, m = (int)(q / nn) * nmin;
                const long long k = (long long)m * nx + n;
                const double y = g->y0 + m * g->dy, x = g->x0 + n * g->dx;
                const Derived d = derived_at(g, u, b, n, m, true);
                double Max, Temp;
                mach_temp(g, s[k], u[k], Max, Temp);
                const double v[15] = {x, y, s[k].x, s[k].y, u[k].x, u[k].y, b[k].x, b[k].y, b[k].z, Max, Temp,
                                      s[k].y + b[k].z * b[k].z / (8.0 * kPiRef), d.Fx, d.Fy, d.zav};
                for (int c = 0; c < 15; ++c) {
                    p = put_g(p, v[c]);
                    *p++ = c == 14 ? '\n' : ' ';
                }
            }
            return p;
        });
    if (rc == GD2_OK && nn > 1 && mm > 1)
        rc = parallel_write(f, (nn - 1) * (mm - 1), 4 * 21 + 4, [&](long long lo, long long hi, char* p) {
            for (long long q = lo; q < hi; ++q) {
                const long long n = q % (nn - 1), m = q / (nn - 1);
                p = put_i(p, m * nn + n + 1); *p++ = ' ';
                p = put_i(p, m * nn + n + 2); *p++ = ' ';
                p = put_i(p, (m + 1) * nn + n + 2); *p++ = ' ';
                p = put_i(p, (m + 1) * nn + n + 1); *p++ = '\n';
            }
            return p;
        });
    if (std::fclose(f) != 0) rc = GD2_ERR_ARG;
    return rc;
}

int gd2_write_line_cuts(const gd2_grid* g, const gd2_double2* s, const gd2_double2* u, const gd2_double3* b,
                        const char* path_y0, const char* path_x0) {
    if (!grid_shape_ok(g) || !s || !u || !b || (!path_y0 && !path_x0)) return GD2_ERR_ARG;
    const int N = g->nx, M = g->ny;
    auto pp = [&](long long k) { return s[k].y + b[k].z * b[k].z / (8.0 * kPiRef); };
    int rc = GD2_OK;
    char line[16 * 32];
    // the first row (kernel.cu:4283-4327): cells 5 .. N-2
    if (path_y0) {
        std::FILE* f = std::fopen(path_y0, "wb");
        if (!f) return GD2_ERR_ARG;
        const char* head = "TITLE = \"HP\"  VARIABLES = \"X\", \"Ro\", \"P\", \"Vx\", \"Bz\", \"Max\", \"T\", \"PP\", "
                           "\"Fx\", \"Fy\", \"dP\", ZONE T = \"HP\"\n";
        if (std::fputs(head, f) < 0) rc = GD2_ERR_ARG;
        for (int n = 5; n < N - 1 && rc == GD2_OK; ++n) {
            const long long k = n;
            const Derived d = derived_at(g, u, b, n, 0, false);
            double Max, Temp;
            mach_temp(g, s[k], u[k], Max, Temp);
            const double v[11] = {g->x0 + n * g->dx, s[k].x, s[k].y, u[k].x, b[k].z, Max, Temp, pp(k), d.Fx, d.Fy,
                                  (pp(k + 1) - pp(k)) / g->dx};
            char* p = line;
            for (int c = 0; c < 11; ++c) {
                p = put_g(p, v[c]);
                *p++ = c == 10 ? '\n' : ' ';
            }
            if (std::fwrite(line, 1, (size_t)(p - line), f) != (size_t)(p - line)) rc = GD2_ERR_ARG;
        }
        if (std::fclose(f) != 0) rc = GD2_ERR_ARG;
    }
    // the first column (kernel.cu:4330-4374): cells 5 .. M-2
    if (path_x0 && rc == GD2_OK) {
        std::FILE* f = std::fopen(path_x0, "wb");
        if (!f) return GD2_ERR_ARG;
        const char* head = "TITLE = \"HP\"  VARIABLES = \"Y\", \"Ro\", \"P\", \"Vy\", \"Bz\", \"Max\", \"T\", \"PP\", "
                           "\"Fx\", \"Fy\", \"dP\", ZONE T = \"HP\"\n";
        if (std::fputs(head, f) < 0) rc = GD2_ERR_ARG;
        for (int m = 5; m < M - 1 && rc == GD2_OK; ++m) {
            const long long k = (long long)m * N;
            const Derived d = derived_at(g, u, b, 0, m, false);
            double Max, Temp;
            mach_temp(g, s[k], u[k], Max, Temp);
            const double v[11] = {g->y0 + m * g->dy, s[k].x, s[k].y, u[k].y, b[k].z, Max, Temp, pp(k), d.Fx, d.Fy,
                                  (pp(k + N) - pp(k)) / g->dy};
            char* p = line;
            for (int c = 0; c < 11; ++c) {
                p = put_g(p, v[c]);
                *p++ = c == 10 ? '\n' : ' ';
            }
            if (std::fwrite(line, 1, (size_t)(p - line), f) != (size_t)(p - line)) rc = GD2_ERR_ARG;
        }
        if (std::fclose(f) != 0) rc = GD2_ERR_ARG;
    }
    return rc;
}

}  // extern "C"
