"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
the cpu_baseline leg of bench.py, always as the checker. The product
(golikov_2d_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

BC_TRANSMISSIVE, BC_PERIODIC, BC_REFLECTIVE, BC_FROZEN = 0, 1, 2, 3
BC_SYMMETRY, BC_AXIS, BC_OUTFLOW, BC_INFLOW = 4, 5, 6, 7
BC_WEST_INFLOW = 8
PHYS_EULER, PHYS_MHD = 0, 1
FLUX_HLLC, FLUX_EXACT, FLUX_HLLC_MHD, FLUX_HLLD, FLUX_HLL = 1, 2, 3, 4, 5
FLUX_HLLC_KOROLKOV, FLUX_HLLC_2D = 6, 7


class D2(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double)]


class D3(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("dx", C.c_double), ("dy", C.c_double),
                ("x0", C.c_double), ("y0", C.c_double), ("gamma", C.c_double),
                ("krit", C.c_double), ("bc", C.c_int * 4), ("r_frozen", C.c_double), ("axisym", C.c_int),
                ("m_inf", C.c_double)]


class Scheme(C.Structure):
    _fields_ = [("physics", C.c_int), ("flux", C.c_int), ("order", C.c_int), ("cfl", C.c_double)]


class Diag(C.Structure):
    _fields_ = [("newton_cap_hits", C.c_long), ("bisect_cap_hits", C.c_long),
                ("clamp_rho", C.c_long), ("clamp_p", C.c_long), ("bad_branch", C.c_long)]


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    so = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(so):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return so


def lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(so):
            build()
        L = C.CDLL(so)
        dp = C.POINTER(C.c_double)
        L.o_minmod.restype = C.c_double
        L.o_minmod.argtypes = [C.c_double, C.c_double]
        L.o_linear.restype = C.c_double
        L.o_linear.argtypes = [C.c_double] * 7
        L.o_linear2.restype = None
        L.o_linear2.argtypes = [C.c_double] * 8 + [dp, dp]
        for fn in (L.o_hllc, L.o_godunov, L.o_hll, L.o_hllc_aleksashov2, L.o_hllc_korolkov):
            fn.restype = C.c_double
            fn.argtypes = [C.c_double, C.c_double, D2, D2, D2, D2, C.c_double, C.c_double,
                           C.POINTER(D2), C.POINTER(D2), C.c_double]
        for fn in (L.o_hllcq_korolkov_2d, L.o_hllcq_aleksashov):
            fn.restype = C.c_double
            fn.argtypes = [C.c_double, C.c_double, D2, D2, D2, D2, C.c_double, C.c_double, C.c_double,
                           C.c_double, C.POINTER(D2), C.POINTER(D2), dp, C.c_double]
        L.o_godunov_star.restype = None
        L.o_godunov_star.argtypes = [C.c_double, D2, D2, D2, D2, C.c_double, C.c_double, dp, dp,
                                     C.POINTER(C.c_int)]
        L.o_hlld.restype = C.c_double
        L.o_hlld.argtypes = [C.c_double, C.c_double] + [C.c_double] * 16 + [dp] + [C.c_double] * 4 + [C.c_int]
        vp = C.c_void_p
        gp = C.POINTER(Grid)
        for fn in (L.o_ker_dekard, L.o_kernel_tvd_old):
            fn.restype = C.c_int
            fn.argtypes = [gp, vp, vp, vp, vp, dp, dp, C.c_int]
        for fn in (L.o_add2, L.o_kernel_tvd):
            fn.restype = C.c_int
            fn.argtypes = [gp, vp, vp, vp, vp, vp, vp, dp, dp, C.c_int, C.c_int]
        L.o_kernel_gkr.restype = C.c_int
        L.o_kernel_gkr.argtypes = [gp, vp, vp, vp, vp, vp, vp, dp, dp, C.c_int]
        L.o_potok.restype = None
        L.o_potok.argtypes = [C.c_double] * 9 + [dp] + [C.c_double] * 3
        L.o_gkr_predictor.restype = None
        L.o_gkr_predictor.argtypes = [gp, vp, vp, vp, vp, C.c_double, C.c_double]
        L.o_funk_time.restype = None
        L.o_funk_time.argtypes = [dp, dp, dp, C.POINTER(C.c_int)]
        sp = C.POINTER(Scheme)
        L.o_stage.restype = C.c_int
        L.o_stage.argtypes = [gp, sp, vp, vp, vp, C.c_double, C.c_double, C.c_double, dp]
        L.o_wavespeed_max.restype = C.c_double
        L.o_wavespeed_max.argtypes = [gp, sp, vp]
        L.o_run.restype = C.c_long
        L.o_run.argtypes = [gp, sp, C.c_int, vp, vp, C.c_long, C.c_double, dp]
        L.o_prim_to_cons.restype = None
        L.o_prim_to_cons.argtypes = [gp, C.c_int, vp, vp, vp, vp]
        L.o_cons_to_prim.restype = None
        L.o_cons_to_prim.argtypes = [gp, C.c_int, vp, vp, vp, vp]
        L.o_diag_reset.restype = None
        L.o_diag_get.restype = None
        L.o_diag_get.argtypes = [C.POINTER(Diag)]
        for fn in (L.o_write_state_text, L.o_write_state_b3):
            fn.restype = C.c_int
            fn.argtypes = [gp, vp, vp, vp, C.c_char_p]
        L.o_read_state_text.restype = C.c_int
        L.o_read_state_text.argtypes = [gp, C.c_char_p, vp, vp, vp]
        L.o_write_tecplot_derived.restype = C.c_int
        L.o_write_tecplot_derived.argtypes = [gp, vp, vp, vp, C.c_int, C.c_char_p]
        L.o_write_line_cuts.restype = C.c_int
        L.o_write_line_cuts.argtypes = [gp, vp, vp, vp, C.c_char_p, C.c_char_p]
        L.o_write_tecplot.restype = C.c_int
        L.o_write_tecplot.argtypes = [gp, vp, vp, vp, C.c_int, dp, C.c_char_p]
        L.o_num_threads.restype = C.c_int
        L.o_set_num_threads.argtypes = [C.c_int]
        _LIB = L
    return _LIB


# ---------------------------------------------------------------- helpers

def make_grid(nx, ny, *, dx=None, dy=None, x0=None, y0=None, gamma=1.4, krit=0.25,
              bc=(BC_TRANSMISSIVE,) * 4, r_frozen=0.0, axisym=False, m_inf=0.0) -> Grid:
    dx = 1.0 / nx if dx is None else dx
    dy = 1.0 / ny if dy is None else dy
    x0 = dx / 2.0 if x0 is None else x0
    y0 = dy / 2.0 if y0 is None else y0
    g = Grid(nx, ny, dx, dy, x0, y0, gamma, krit)
    for i in range(4):
        g.bc[i] = bc[i]
    g.r_frozen = float(r_frozen)
    g.axisym = 1 if axisym else 0
    g.m_inf = float(m_inf)
    return g


def _ptr(a: np.ndarray):
    assert a.flags.c_contiguous and a.dtype == np.float64
    return a.ctypes.data_as(C.c_void_p)


def minmod(x, y):
    return lib().o_minmod(x, y)


def linear(*a):
    return lib().o_linear(*a)


def linear2(x1, t1, x2, t2, x3, t3, y1, y2):
    A, B = C.c_double(), C.c_double()
    lib().o_linear2(x1, t1, x2, t2, x3, t3, y1, y2, C.byref(A), C.byref(B))
    return A.value, B.value


EULER_FLUX_KINDS = ("hllc", "hll", "godunov", "exact", "hllc_aleksashov2", "hllc_korolkov", "hllc_korolkov_2d",
                    "hllcq_korolkov_2d", "hllcq_aleksashov")


def euler_flux(kind, gamma, krit, L, R, n, rad=1.0):
    """L, R = (rho, p, u, v). Returns (F_rho, F_E, F_mx, F_my), dt_face."""
    if kind in ("hllc_korolkov_2d", "hllcq_korolkov_2d", "hllcq_aleksashov"):
        F, _, t = euler_flux_q(kind, gamma, krit, L, R, 1.0, 1.0, n, rad)
        return F, t
    Ps, Pu = D2(), D2()
    fn = {"hllc": lib().o_hllc, "hll": lib().o_hll, "godunov": lib().o_godunov, "exact": lib().o_godunov,
          "hllc_aleksashov2": lib().o_hllc_aleksashov2, "hllc_korolkov": lib().o_hllc_korolkov}[kind]
    t = fn(gamma, krit, D2(L[0], L[1]), D2(L[2], L[3]), D2(R[0], R[1]), D2(R[2], R[3]),
           float(n[0]), float(n[1]), C.byref(Ps), C.byref(Pu), rad)
    return (Ps.x, Ps.y, Pu.x, Pu.y), t


def euler_flux_q(kind, gamma, krit, L, R, QL, QR, n, rad=1.0):
    """Passive-scalar variants (HLLCQ_Korolkov_2D, HLLCQ_Aleksashov; kind
    hllc_korolkov_2d is the former without the scalar). Returns
    ((F_rho, F_E, F_mx, F_my), F_Q, dt_face)."""
    Ps, Pu, PQ = D2(), D2(), C.c_double()
    fn = {"hllc_korolkov_2d": lib().o_hllcq_korolkov_2d, "hllcq_korolkov_2d": lib().o_hllcq_korolkov_2d,
          "hllcq_aleksashov": lib().o_hllcq_aleksashov}[kind]
    t = fn(gamma, krit, D2(L[0], L[1]), D2(L[2], L[3]), D2(R[0], R[1]), D2(R[2], R[3]), float(QL), float(QR),
           float(n[0]), float(n[1]), C.byref(Ps), C.byref(Pu), C.byref(PQ), rad)
    return (Ps.x, Ps.y, Pu.x, Pu.y), PQ.value, t


def godunov_star(gamma, L, R, n=(1.0, 0.0)):
    p, u, pat = C.c_double(), C.c_double(), C.c_int()
    lib().o_godunov_star(gamma, D2(L[0], L[1]), D2(L[2], L[3]), D2(R[0], R[1]), D2(R[2], R[3]),
                         float(n[0]), float(n[1]), C.byref(p), C.byref(u), C.byref(pat))
    return p.value, u.value, pat.value


def mhd_flux(gamma, krit, L, R, n, metod, rad=1.0):
    """L, R = (rho, p, vx, vy, vz, Bx, By, Bz) Gaussian B. Returns P[8], dt_face."""
    P = (C.c_double * 8)()
    t = lib().o_hlld(gamma, krit, *[float(v) for v in L], *[float(v) for v in R],
                     P, float(n[0]), float(n[1]), float(n[2]), rad, int(metod))
    return np.array(list(P)), t


def diag_reset():
    lib().o_diag_reset()


def diag():
    d = Diag()
    lib().o_diag_get(C.byref(d))
    return {f: getattr(d, f) for f, _ in Diag._fields_}


def nfields(physics):
    return 7 if physics == PHYS_MHD else 4


def stage(g: Grid, sc: Scheme, Uin, Ubase, a, b, dt, want_w=False):
    Uout = np.empty_like(Uin)
    w = C.c_double(0.0)
    rc = lib().o_stage(C.byref(g), C.byref(sc), _ptr(Uin), _ptr(Ubase) if Ubase is not None else None,
                       _ptr(Uout), a, b, dt, C.byref(w) if want_w else None)
    if rc != 0:
        raise ValueError(f"o_stage rejected the scheme/grid (rc={rc})")
    return (Uout, w.value) if want_w else Uout


def wavespeed_max(g, sc, U):
    return lib().o_wavespeed_max(C.byref(g), C.byref(sc), _ptr(U))


def run(g, sc, integrator, U, nsteps=0, t_end=0.0, t0=0.0):
    """Advance U (modified in place). Returns (steps, t)."""
    tmp = np.empty((2,) + U.shape, dtype=np.float64)
    t = C.c_double(t0)
    steps = lib().o_run(C.byref(g), C.byref(sc), integrator, _ptr(U), _ptr(tmp), nsteps, t_end, C.byref(t))
    if steps < 0:
        raise ValueError("o_run rejected the scheme/grid")
    return steps, t.value


def prim_to_cons(g, physics, s, u, b=None):
    U = np.empty((nfields(physics), g.ny, g.nx))
    lib().o_prim_to_cons(C.byref(g), physics, _ptr(s), _ptr(u), _ptr(b) if b is not None else None, _ptr(U))
    return U


def cons_to_prim(g, physics, U):
    s = np.empty((g.ny, g.nx, 2))
    u = np.empty((g.ny, g.nx, 2))
    b = np.empty((g.ny, g.nx, 3)) if physics == PHYS_MHD else None
    lib().o_cons_to_prim(C.byref(g), physics, _ptr(U), _ptr(s), _ptr(u), _ptr(b) if b is not None else None)
    return s, u, b


def ref_euler_step(kind, g, s, u, T, T_do, method):
    """kind: 'ker_dekard' | 'kernel_tvd_old'. s, u: (ny, nx, 2). Returns s2, u2, T."""
    s2 = np.empty_like(s)
    u2 = np.empty_like(u)
    Tc, Td = C.c_double(T), C.c_double(T_do)
    fn = lib().o_ker_dekard if kind == "ker_dekard" else lib().o_kernel_tvd_old
    rc = fn(C.byref(g), _ptr(s), _ptr(u), _ptr(s2), _ptr(u2), C.byref(Tc), C.byref(Td), method)
    if rc != 0:
        raise ValueError(f"{kind} rejected method {method}")
    return s2, u2, Tc.value


def ref_mhd_step(kind, g, s, u, b, T, T_do, i, method):
    """kind: 'add2' | 'kernel_tvd'. Returns s2, u2, b2, T."""
    s2, u2, b2 = np.empty_like(s), np.empty_like(u), np.empty_like(b)
    Tc, Td = C.c_double(T), C.c_double(T_do)
    fn = lib().o_add2 if kind == "add2" else lib().o_kernel_tvd
    rc = fn(C.byref(g), _ptr(s), _ptr(u), _ptr(b), _ptr(s2), _ptr(u2), _ptr(b2), C.byref(Tc), C.byref(Td),
            i, method)
    if rc != 0:
        raise ValueError(f"{kind} rejected method {method}")
    return s2, u2, b2, Tc.value


def ref_gkr_step(g, s, u, b, T, T_do, method=3):
    """Kernel_GKR (kernel.cu:2663): one predictor-corrector step. Returns s2, u2, b2, T."""
    s2, u2, b2 = np.empty_like(s), np.empty_like(u), np.empty_like(b)
    Tc, Td = C.c_double(T), C.c_double(T_do)
    rc = lib().o_kernel_gkr(C.byref(g), _ptr(s), _ptr(u), _ptr(b), _ptr(s2), _ptr(u2), _ptr(b2), C.byref(Tc),
                            C.byref(Td), method)
    if rc != 0:
        raise ValueError(f"o_kernel_gkr rejected method {method} or the grid")
    return s2, u2, b2, Tc.value


def potok(gamma, q, n):
    """POTOK_Korolkov: physical MHD flux of q = (rho, p, vx, vy, vz, Bx, By, Bz)
    (Gaussian B) along an axis normal n; P[8] with the energy flux last."""
    P = (C.c_double * 8)()
    lib().o_potok(gamma, *[float(v) for v in q], P, float(n[0]), float(n[1]), float(n[2]))
    return np.array(list(P))


def gkr_predictor(g, Q, DX, DY, step, y):
    """predictor (kernel.cu:1611): Q, DX, DY of 7 entries (rho, p, vx, vy, Bx, By, Bz)."""
    Q, DX, DY = (np.ascontiguousarray(a, dtype=np.float64) for a in (Q, DX, DY))
    QQ = np.empty(7)
    lib().o_gkr_predictor(C.byref(g), _ptr(Q), _ptr(DX), _ptr(DY), _ptr(QQ), float(step), float(y))
    return QQ


def funk_time(T, T_do, TT, i):
    a, b, c, d = C.c_double(T), C.c_double(T_do), C.c_double(TT), C.c_int(i)
    lib().o_funk_time(C.byref(a), C.byref(b), C.byref(c), C.byref(d))
    return a.value, b.value, c.value, d.value


def num_threads():
    return lib().o_num_threads()


def set_num_threads(n):
    lib().o_set_num_threads(n)


# ---------------------------------------------------------------- on-disk formats

def write_state_text(g, s, u, b, path):
    if lib().o_write_state_text(C.byref(g), _ptr(s), _ptr(u), _ptr(b) if b is not None else None,
                                os.fsencode(path)) != 0:
        raise OSError(path)


def read_state_text(g, path, mhd=True):
    s = np.empty((g.ny, g.nx, 2)); u = np.empty((g.ny, g.nx, 2))
    b = np.empty((g.ny, g.nx, 3)) if mhd else None
    if lib().o_read_state_text(C.byref(g), os.fsencode(path), _ptr(s), _ptr(u),
                               _ptr(b) if b is not None else None) != 0:
        raise OSError(path)
    return s, u, b


def write_state_b3(g, s, u, b, path):
    if lib().o_write_state_b3(C.byref(g), _ptr(s), _ptr(u), _ptr(b), os.fsencode(path)) != 0:
        raise OSError(path)


def write_tecplot(g, s, u, b, path, nmin=3, solution_time=None):
    t = None if solution_time is None else C.byref(C.c_double(solution_time))
    if lib().o_write_tecplot(C.byref(g), _ptr(s), _ptr(u), _ptr(b) if b is not None else None, nmin, t,
                             os.fsencode(path)) != 0:
        raise OSError(path)


def write_tecplot_derived(g, s, u, b, path, nmin=3):
    if lib().o_write_tecplot_derived(C.byref(g), _ptr(s), _ptr(u), _ptr(b), nmin, os.fsencode(path)) != 0:
        raise OSError(path)


def write_line_cuts(g, s, u, b, path_y0=None, path_x0=None):
    enc = lambda p: None if p is None else os.fsencode(p)
    if lib().o_write_line_cuts(C.byref(g), _ptr(s), _ptr(u), _ptr(b), enc(path_y0), enc(path_x0)) != 0:
        raise OSError(str(path_y0))
