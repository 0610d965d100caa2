"""Host-side Python mirror of the libgd2.so C ABI (include/gd2.h).

Device buffers are torch tensors on a ROCm device (torch is the allocator and
stream provider only); every compute call goes through libgd2.so. If the
library is not built the first call raises: there is no CPU fallback.

Reference-layout functions keep the reference kernels' names and argument
order (kernel.cu:1736 Ker_Dekard, 3286 Kernel_TVD_old, 1877 add2, 2198
Kernel_TVD, 1722 funk_time); tensors s, u have shape (ny, nx, 2) (double2 per
cell) and b has shape (ny, nx, 3).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
SO_PATH = os.path.join(HERE, "libgd2.so")

BC_TRANSMISSIVE, BC_PERIODIC, BC_REFLECTIVE, BC_FROZEN = 0, 1, 2, 3
BC_SYMMETRY, BC_AXIS, BC_OUTFLOW, BC_INFLOW = 4, 5, 6, 7
BC_WEST_INFLOW = 8
EULER, MHD = 0, 1
FLUX_HLLC, FLUX_EXACT, FLUX_HLLC_MHD, FLUX_HLLD, FLUX_HLL = 1, 2, 3, 4, 5
FLUX_HLLC_KOROLKOV, FLUX_HLLC_2D = 6, 7  # HLLC_Korolkov (kernel.cu:1349), HLLC_Korolkov_2D (Gas_Dinamic.cu:265)
# pointwise flux kinds of gd2_euler_face_fluxes
(FACE_HLL, FACE_HLLC_ALEKSASHOV, FACE_GODUNOV, FACE_HLLC_ALEKSASHOV2, FACE_HLLC_KOROLKOV, FACE_HLLC_KOROLKOV_2D,
 FACE_HLLCQ_KOROLKOV_2D, FACE_HLLCQ_ALEKSASHOV) = range(8)
FORWARD_EULER, SSPRK2 = 1, 2

_lib = None


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("dx", C.c_double), ("dy", C.c_double),
                ("x0", C.c_double), ("y0", C.c_double), ("gamma", C.c_double),
                ("krit", C.c_double), ("bc", C.c_int * 4), ("r_frozen", C.c_double), ("axisym", C.c_int),
                ("m_inf", C.c_double)]


class Scheme(C.Structure):
    _fields_ = [("physics", C.c_int), ("flux", C.c_int), ("order", C.c_int), ("integrator", C.c_int),
                ("cfl", C.c_double)]


class Diag(C.Structure):
    _fields_ = [("clamp_rho", C.c_longlong), ("clamp_p", C.c_longlong),
                ("newton_cap_hits", C.c_longlong), ("bisect_cap_hits", C.c_longlong)]


class GD2Error(RuntimeError):
    pass


def make_grid(nx, ny, *, dx=None, dy=None, x0=None, y0=None, gamma=1.4, krit=0.25,
              bc=(BC_TRANSMISSIVE,) * 4, r_frozen=0.0, axisym=False, m_inf=0.0) -> Grid:
    """Uniform grid on [0, nx*dx] x [0, ny*dy]; defaults to the unit square."""
    dx = 1.0 / nx if dx is None else dx
    dy = 1.0 / ny if dy is None else dy
    g = Grid(nx, ny, dx, dy, dx / 2.0 if x0 is None else x0, dy / 2.0 if y0 is None else y0, gamma, krit)
    for i in range(4):
        g.bc[i] = bc[i]
    g.r_frozen = float(r_frozen)
    g.axisym = 1 if axisym else 0
    g.m_inf = float(m_inf)
    return g


def make_scheme(physics=EULER, flux=FLUX_HLLC, order=2, integrator=SSPRK2, cfl=0.4) -> Scheme:
    return Scheme(physics, flux, order, integrator, cfl)


def lib():
    """Load libgd2.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(SO_PATH):
            raise GD2Error(f"{SO_PATH} is missing: build it with golikov_2d_amd/build.py "
                           "(or __graft_entry__.build()); there is no CPU fallback")
        L = C.CDLL(SO_PATH)
        vp, ip, dp = C.c_void_p, C.c_int, C.c_double
        gp, sp = C.POINTER(Grid), C.POINTER(Scheme)
        sig = {
            "gd2_nfields": (ip, [ip]),
            "gd2_last_hip_error": (ip, []),
            "gd2_version": (C.c_char_p, []),
            "gd2_stage": (ip, [gp, sp, vp, vp, vp, dp, dp, dp, vp, vp, vp]),
            "gd2_ssprk2_step": (ip, [gp, sp, vp, vp, vp, dp, vp, vp, vp]),
            "gd2_wavespeed_max": (ip, [gp, sp, vp, vp, vp]),
            "gd2_ker_dekard": (ip, [gp, vp, vp, vp, vp, vp, vp, ip, vp]),
            "gd2_kernel_tvd_old": (ip, [gp, vp, vp, vp, vp, vp, vp, ip, vp]),
            "gd2_add2": (ip, [gp, vp, vp, vp, vp, vp, vp, vp, vp, ip, ip, vp]),
            "gd2_kernel_tvd": (ip, [gp, vp, vp, vp, vp, vp, vp, vp, vp, ip, ip, vp]),
            "gd2_kernel_gkr": (ip, [gp, vp, vp, vp, vp, vp, vp, vp, vp, ip, ip, vp]),
            "gd2_funk_time": (ip, [vp, vp, vp, vp, vp]),
            "gd2_prim_to_cons": (ip, [gp, ip, vp, vp, vp, vp, vp]),
            "gd2_cons_to_prim": (ip, [gp, ip, vp, vp, vp, vp, vp]),
            "gd2_ctx_create": (ip, [gp, sp, C.POINTER(vp)]),
            "gd2_ctx_destroy": (ip, [vp]),
            "gd2_ctx_start": (ip, [vp, vp, dp, vp]),
            "gd2_ctx_advance": (ip, [vp, vp, ip, dp, vp]),
            "gd2_ctx_time": (ip, [vp, C.POINTER(dp), C.POINTER(C.c_longlong), C.POINTER(dp), vp]),
            "gd2_get_diagnostics": (ip, [C.POINTER(Diag), ip]),
            "gd2_ensemble_create": (ip, [ip, C.POINTER(Grid), sp, C.POINTER(vp), C.POINTER(vp)]),
            "gd2_ensemble_destroy": (ip, [vp]),
            "gd2_ensemble_set_tiling": (ip, [vp, ip, ip]),
            "gd2_ensemble_get_tiling": (ip, [vp, C.POINTER(ip), C.POINTER(ip), C.POINTER(ip)]),
            "gd2_ensemble_start": (ip, [vp, vp]),
            "gd2_ensemble_advance": (ip, [vp, ip, ip, vp]),
            "gd2_ensemble_time": (ip, [vp, ip, C.POINTER(dp), C.POINTER(C.c_longlong), C.POINTER(dp), vp]),
            "gd2_ensemble_tune": (ip, [vp, vp, C.POINTER(ip), C.POINTER(ip), C.POINTER(dp)]),
            "gd2_ctx_log_dt": (ip, [vp, vp, C.c_longlong]),
            "gd2_ctx_time_kernels": (ip, [vp, ip, ip]),
            "gd2_ctx_kernel_ms": (ip, [vp, C.POINTER(dp), C.POINTER(ip)]),
            "gd2_vjp_workspace_bytes": (C.c_longlong, [gp, sp]),
            "gd2_stage_vjp": (ip, [gp, sp, vp, vp, vp, dp, vp, dp, dp, vp, vp, vp]),
            "gd2_stage_vjp_noalloc": (ip, [gp, sp, vp, vp, vp, dp, vp, dp, dp, vp, vp]),
            "gd2_euler_face_fluxes": (ip, [ip, dp, dp, dp, C.c_longlong] + [vp] * 11),
            "gd2_mhd_face_fluxes": (ip, [ip, dp, dp, dp, C.c_longlong] + [vp] * 6),
            "gd2_set_step_path": (ip, [ip]),
            "gd2_one_pass_preferred": (ip, [gp, sp]),
            "gd2_write_state_text": (ip, [gp, vp, vp, vp, C.c_char_p]),
            "gd2_read_state_text": (ip, [gp, C.c_char_p, vp, vp, vp]),
            "gd2_write_state_b3": (ip, [gp, vp, vp, vp, C.c_char_p]),
            "gd2_write_tecplot": (ip, [gp, vp, vp, vp, ip, C.POINTER(dp), C.c_char_p]),
            "gd2_write_tecplot_derived": (ip, [gp, vp, vp, vp, ip, C.c_char_p]),
            "gd2_write_line_cuts": (ip, [gp, vp, vp, vp, C.c_char_p, C.c_char_p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


EXPORTED = ["gd2_nfields", "gd2_last_hip_error", "gd2_version", "gd2_stage", "gd2_wavespeed_max",
            "gd2_ker_dekard", "gd2_kernel_tvd_old", "gd2_add2", "gd2_kernel_tvd", "gd2_kernel_gkr", "gd2_funk_time",
            "gd2_prim_to_cons", "gd2_cons_to_prim", "gd2_ctx_create", "gd2_ctx_destroy", "gd2_ctx_start",
            "gd2_ctx_advance", "gd2_ctx_time", "gd2_get_diagnostics", "gd2_ensemble_create",
            "gd2_ensemble_destroy", "gd2_ensemble_set_tiling", "gd2_ensemble_get_tiling", "gd2_ensemble_start",
            "gd2_ensemble_advance", "gd2_ensemble_time", "gd2_ensemble_tune", "gd2_ctx_log_dt",
            "gd2_ctx_time_kernels", "gd2_ctx_kernel_ms", "gd2_vjp_workspace_bytes", "gd2_stage_vjp", "gd2_stage_vjp_noalloc", "gd2_euler_face_fluxes", "gd2_mhd_face_fluxes", "gd2_ssprk2_step", "gd2_set_step_path",
            "gd2_one_pass_preferred", "gd2_write_state_text", "gd2_read_state_text", "gd2_write_state_b3",
            "gd2_write_tecplot", "gd2_write_tecplot_derived", "gd2_write_line_cuts"]


def _check(rc, name):
    if rc != 0:
        raise GD2Error(f"{name} returned {rc} (hip error {lib().gd2_last_hip_error()})")


def _p(t):
    if t is None:
        return None
    if not t.is_cuda:
        raise GD2Error("device tensor expected")
    if not t.is_contiguous():
        raise GD2Error("contiguous tensor expected")
    return C.c_void_p(t.data_ptr())


def _f64(t):
    if t is not None and t.dtype != torch.float64:
        raise GD2Error("float64 tensor expected")
    return _p(t)


def _stream(stream=None):
    s = torch.cuda.current_stream() if stream is None else stream
    return C.c_void_p(s.cuda_stream)


def nfields(physics):
    return lib().gd2_nfields(physics)


# ------------------------------------------------ reference-layout entry points

def _check_aos(g, *ts):
    for t in ts:
        if t is not None and (t.shape[0] != g.ny or t.shape[1] != g.nx):
            raise GD2Error(f"state tensor shape {tuple(t.shape)} does not match grid {g.ny}x{g.nx}")


def ker_dekard(g, s, u, s2, u2, T, T_do, method, stream=None):
    _check_aos(g, s, u, s2, u2)
    _check(lib().gd2_ker_dekard(C.byref(g), _f64(s), _f64(u), _f64(s2), _f64(u2), _f64(T), _f64(T_do),
                                method, _stream(stream)), "gd2_ker_dekard")


def kernel_tvd_old(g, s, u, s2, u2, T, T_do, method, stream=None):
    _check_aos(g, s, u, s2, u2)
    _check(lib().gd2_kernel_tvd_old(C.byref(g), _f64(s), _f64(u), _f64(s2), _f64(u2), _f64(T), _f64(T_do),
                                    method, _stream(stream)), "gd2_kernel_tvd_old")


def add2(g, s, u, b, s2, u2, b2, T, T_do, i, method, stream=None):
    _check_aos(g, s, u, b, s2, u2, b2)
    _check(lib().gd2_add2(C.byref(g), _f64(s), _f64(u), _f64(b), _f64(s2), _f64(u2), _f64(b2), _f64(T),
                          _f64(T_do), i, method, _stream(stream)), "gd2_add2")


def kernel_tvd(g, s, u, b, s2, u2, b2, T, T_do, i, method, stream=None):
    _check_aos(g, s, u, b, s2, u2, b2)
    _check(lib().gd2_kernel_tvd(C.byref(g), _f64(s), _f64(u), _f64(b), _f64(s2), _f64(u2), _f64(b2),
                                _f64(T), _f64(T_do), i, method, _stream(stream)), "gd2_kernel_tvd")


def kernel_gkr(g, s, u, b, s2, u2, b2, T, T_do, i, method=3, stream=None):
    """Kernel_GKR (kernel.cu:2663): the predictor-corrector MHD step."""
    _check_aos(g, s, u, b, s2, u2, b2)
    _check(lib().gd2_kernel_gkr(C.byref(g), _f64(s), _f64(u), _f64(b), _f64(s2), _f64(u2), _f64(b2), _f64(T),
                                _f64(T_do), i, method, _stream(stream)), "gd2_kernel_gkr")


def funk_time(T, T_do, TT, i, stream=None):
    if i.dtype != torch.int32:
        raise GD2Error("step counter must be int32")
    _check(lib().gd2_funk_time(_f64(T), _f64(T_do), _f64(TT), _p(i), _stream(stream)), "gd2_funk_time")


# ------------------------------------------------ native path

def stage(g, sc, Uin, Ubase, Uout, a=0.0, b=1.0, dt=0.0, dt_dev=None, wmax_bits=None, stream=None):
    _check(lib().gd2_stage(C.byref(g), C.byref(sc), _f64(Uin), _f64(Ubase), _f64(Uout), float(a), float(b),
                           float(dt), _f64(dt_dev), _p(wmax_bits), _stream(stream)), "gd2_stage")


def ssprk2_step(g, sc, Uin, Uout, dt=0.0, dt_dev=None, wmax_bits=None, stream=None, ustage=None):
    """One SSP-RK2 step Uin -> Uout in a single pass (Euler); bitwise equal to
    stage(Uin -> S) followed by stage(S, Ubase=Uin -> Uout, a=b=1/2). When
    `ustage` is given it receives the stage-1 state S."""
    _check(lib().gd2_ssprk2_step(C.byref(g), C.byref(sc), _f64(Uin), _f64(Uout), _f64(ustage), float(dt),
                                 _f64(dt_dev), _p(wmax_bits), _stream(stream)), "gd2_ssprk2_step")


STEP_AUTO, STEP_TWO_PASS, STEP_ONE_PASS = 0, 1, 2


def set_step_path(mode) -> int:
    """Process-wide choice between the one-pass SSP-RK2 kernel and two stage
    launches where the library steps by itself (the solver context); returns
    the previous mode. Both paths give the same bits."""
    rc = lib().gd2_set_step_path(int(mode))
    if rc < 0:
        raise GD2Error(f"gd2_set_step_path({mode}) failed ({rc})")
    return rc


def one_pass_preferred(g, sc) -> bool:
    rc = lib().gd2_one_pass_preferred(C.byref(g), C.byref(sc))
    if rc < 0:
        raise GD2Error(f"gd2_one_pass_preferred failed ({rc})")
    return rc == 1


def vjp_workspace_bytes(g, sc=None) -> int:
    sc = make_scheme() if sc is None else sc
    n = lib().gd2_vjp_workspace_bytes(C.byref(g), C.byref(sc))
    if n < 0:
        raise GD2Error(f"gd2_vjp_workspace_bytes returned {n}")
    return n


def vjp_workspace(g, device, sc=None):
    return torch.empty(vjp_workspace_bytes(g, sc) // 8, dtype=torch.float64, device=device)


def stage_vjp(g, sc, U, V, W, c=1.0, A=None, a=0.0, B=None, b=0.0, workspace=None, stream=None):
    """W = a*A + b*B + c * J_L(U)^T V for the stage operator L (Euler HLLC, or
    MHD HLLC-MHD / HLLD with the Powell term)."""
    nf = 7 if sc.physics == MHD else 4
    for t in (U, V, W, A, B):
        if t is not None and tuple(t.shape) != (nf, g.ny, g.nx):
            raise GD2Error(f"state shape {tuple(t.shape)} is not ({nf}, {g.ny}, {g.nx})")
    if workspace is None:
        workspace = vjp_workspace(g, U.device, sc)
    elif workspace.numel() * workspace.element_size() < vjp_workspace_bytes(g, sc):
        raise GD2Error("workspace smaller than gd2_vjp_workspace_bytes")
    _check(lib().gd2_stage_vjp(C.byref(g), C.byref(sc), _f64(U), _f64(V), _f64(A), float(a), _f64(B), float(b),
                               float(c), _f64(W), _p(workspace), _stream(stream)), "gd2_stage_vjp")


ERR_TOO_LARGE = -5
NOALLOC_SCRATCH_BYTES = 32 << 20  # GD2_NOALLOC_SCRATCH_BYTES


def stage_vjp_noalloc(g, sc, U, V, W, c=1.0, A=None, a=0.0, B=None, b=0.0, stream=None):
    """stage_vjp through gd2_stage_vjp_noalloc: no workspace, no device
    allocation (the library's fixed per-device scratch array is used)."""
    nf = 7 if sc.physics == MHD else 4
    for t in (U, V, W, A, B):
        if t is not None and tuple(t.shape) != (nf, g.ny, g.nx):
            raise GD2Error(f"state shape {tuple(t.shape)} is not ({nf}, {g.ny}, {g.nx})")
    _check(lib().gd2_stage_vjp_noalloc(C.byref(g), C.byref(sc), _f64(U), _f64(V), _f64(A), float(a), _f64(B),
                                       float(b), float(c), _f64(W), _stream(stream)), "gd2_stage_vjp_noalloc")


def euler_face_fluxes(kind, gamma, krit, Ls, Lu, Rs, Ru, normal, rad=1.0, QL=None, QR=None, stream=None):
    """The reference's Euler flux function `kind` (FACE_*) on n faces: Ls, Lu,
    Rs, Ru, normal of shape (n, 2) (rho, p) / (vx, vy) / (n1, n2); QL, QR of
    shape (n,) for the passive-scalar kinds. Returns Ps (n, 2) = (F_rho, F_E),
    Pu (n, 2) = (F_mx, F_my), PQ (n,) or None, dt (n,)."""
    n = Ls.shape[0]
    for t in (Ls, Lu, Rs, Ru, normal):
        if tuple(t.shape) != (n, 2):
            raise GD2Error("face arrays of shape (n, 2) expected")
    q = kind in (FACE_HLLCQ_KOROLKOV_2D, FACE_HLLCQ_ALEKSASHOV)
    Ps, Pu = torch.empty_like(Ls), torch.empty_like(Ls)
    dt = torch.empty(n, dtype=torch.float64, device=Ls.device)
    PQ = torch.empty(n, dtype=torch.float64, device=Ls.device) if q else None
    _check(lib().gd2_euler_face_fluxes(int(kind), float(gamma), float(krit), float(rad), n, _f64(Ls), _f64(Lu),
                                       _f64(Rs), _f64(Ru), _f64(QL) if q else None, _f64(QR) if q else None,
                                       _f64(normal), _f64(Ps), _f64(Pu), _f64(PQ), _f64(dt), _stream(stream)),
           "gd2_euler_face_fluxes")
    return Ps, Pu, PQ, dt


def mhd_face_fluxes(metod, gamma, krit, L, R, normal, rad=1.0, stream=None):
    """HLLDQ_Korolkov metod 2 / 3 on n faces: L, R of shape (n, 8) (rho, p,
    vx, vy, vz, Bx, By, Bz; Gaussian B), normal (n, 3). Returns P (n, 8) =
    (F_rho, F_mx, F_my, F_mz, F_Bx, F_By, F_Bz, F_E) and dt (n,)."""
    n = L.shape[0]
    if tuple(L.shape) != (n, 8) or tuple(R.shape) != (n, 8) or tuple(normal.shape) != (n, 3):
        raise GD2Error("face arrays of shape (n, 8) and (n, 3) expected")
    P = torch.empty_like(L)
    dt = torch.empty(n, dtype=torch.float64, device=L.device)
    _check(lib().gd2_mhd_face_fluxes(int(metod), float(gamma), float(krit), float(rad), n, _f64(L), _f64(R),
                                     _f64(normal), _f64(P), _f64(dt), _stream(stream)), "gd2_mhd_face_fluxes")
    return P, dt


def bits_to_double(t):
    """View an int64 device scalar holding double bits as a Python float."""
    return t.view(torch.float64).item()


def wavespeed_max(g, sc, U, stream=None) -> float:
    w = torch.zeros(1, dtype=torch.int64, device=U.device)
    _check(lib().gd2_wavespeed_max(C.byref(g), C.byref(sc), _f64(U), _p(w), _stream(stream)), "gd2_wavespeed_max")
    return bits_to_double(w)


def prim_to_cons(g, physics, s, u, b=None, stream=None):
    U = torch.empty((nfields(physics), g.ny, g.nx), dtype=torch.float64, device=s.device)
    _check(lib().gd2_prim_to_cons(C.byref(g), physics, _f64(s), _f64(u), _f64(b), _f64(U), _stream(stream)),
           "gd2_prim_to_cons")
    return U


def cons_to_prim(g, physics, U, stream=None):
    dev = U.device
    s = torch.empty((g.ny, g.nx, 2), dtype=torch.float64, device=dev)
    u = torch.empty((g.ny, g.nx, 2), dtype=torch.float64, device=dev)
    b = torch.empty((g.ny, g.nx, 3), dtype=torch.float64, device=dev) if physics == MHD else None
    _check(lib().gd2_cons_to_prim(C.byref(g), physics, _f64(U), _f64(s), _f64(u), _f64(b), _stream(stream)),
           "gd2_cons_to_prim")
    return s, u, b


def diagnostics(reset=False):
    d = Diag()
    _check(lib().gd2_get_diagnostics(C.byref(d), 1 if reset else 0), "gd2_get_diagnostics")
    return {f: getattr(d, f) for f, _ in Diag._fields_}


class Solver:
    """Native solver context (gd2_ctx): scratch state + device clock."""

    def __init__(self, g: Grid, sc: Scheme):
        self.g, self.sc = g, sc
        h = C.c_void_p()
        _check(lib().gd2_ctx_create(C.byref(g), C.byref(sc), C.byref(h)), "gd2_ctx_create")
        self._h = h

    def start(self, U, t0=0.0, stream=None):
        _check(lib().gd2_ctx_start(self._h, _f64(U), float(t0), _stream(stream)), "gd2_ctx_start")

    def advance(self, U, nsteps, t_end=0.0, stream=None):
        _check(lib().gd2_ctx_advance(self._h, _f64(U), int(nsteps), float(t_end), _stream(stream)),
               "gd2_ctx_advance")

    def time(self, stream=None):
        t, dt, n = C.c_double(), C.c_double(), C.c_longlong()
        _check(lib().gd2_ctx_time(self._h, C.byref(t), C.byref(n), C.byref(dt), _stream(stream)), "gd2_ctx_time")
        return t.value, n.value, dt.value

    def log_dt(self, dt_log):
        """Record the dt of every following step into dt_log[step] (float64 device tensor; None: off)."""
        if dt_log is None:
            _check(lib().gd2_ctx_log_dt(self._h, None, 0), "gd2_ctx_log_dt")
        else:
            _check(lib().gd2_ctx_log_dt(self._h, _f64(dt_log), dt_log.numel()), "gd2_ctx_log_dt")
        self._dt_log = dt_log

    def time_kernels(self, capacity, stride=1):
        """HIP events around the kernel launches of every `stride`-th following step until
        `capacity` steps are timed (0: off)."""
        _check(lib().gd2_ctx_time_kernels(self._h, int(capacity), int(stride)), "gd2_ctx_time_kernels")

    def kernel_ms(self):
        """(mean ms per timed step's kernels, number of steps timed) since the last call; waits for them."""
        ms, n = C.c_double(), C.c_int()
        _check(lib().gd2_ctx_kernel_ms(self._h, C.byref(ms), C.byref(n)), "gd2_ctx_kernel_ms")
        return ms.value, n.value

    def run_to(self, U, t_end, chunk=32, max_steps=10 ** 7):
        """Advance until t_end, checking the device clock every `chunk` steps."""
        while True:
            t, n, _ = self.time()
            if not (t < t_end) or n >= max_steps:
                return t, n
            self.advance(U, chunk, t_end)

    def close(self):
        if self._h:
            lib().gd2_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Ensemble:
    """Batched ensemble of independent Euler problems (gd2_ensemble)."""

    def __init__(self, grids, sc: Scheme, states):
        if len(grids) != len(states) or not grids:
            raise GD2Error("one state per grid expected")
        n = len(grids)
        for g, U in zip(grids, states):
            if tuple(U.shape) != (4, g.ny, g.nx):
                raise GD2Error("state shape must be (4, ny, nx)")
            _f64(U)
        self.grids, self.sc, self.states = list(grids), sc, list(states)
        garr = (Grid * n)(*grids)
        parr = (C.c_void_p * n)(*[U.data_ptr() for U in states])
        h = C.c_void_p()
        _check(lib().gd2_ensemble_create(n, garr, C.byref(sc), parr, C.byref(h)), "gd2_ensemble_create")
        self._h = h

    def set_tiling(self, rows=0, blocks=0):
        _check(lib().gd2_ensemble_set_tiling(self._h, rows, blocks), "gd2_ensemble_set_tiling")

    def tiling(self):
        r, b, t = C.c_int(), C.c_int(), C.c_int()
        _check(lib().gd2_ensemble_get_tiling(self._h, C.byref(r), C.byref(b), C.byref(t)), "gd2_ensemble_get_tiling")
        return r.value, b.value, t.value

    def start(self, stream=None):
        _check(lib().gd2_ensemble_start(self._h, _stream(stream)), "gd2_ensemble_start")

    def advance(self, nsteps, use_graph=False, stream=None):
        _check(lib().gd2_ensemble_advance(self._h, int(nsteps), 1 if use_graph else 0, _stream(stream)),
               "gd2_ensemble_advance")

    def time(self, p, stream=None):
        t, dt, n = C.c_double(), C.c_double(), C.c_longlong()
        _check(lib().gd2_ensemble_time(self._h, p, C.byref(t), C.byref(n), C.byref(dt), _stream(stream)),
               "gd2_ensemble_time")
        return t.value, n.value, dt.value

    def tune(self, stream=None):
        r, b, ms = C.c_int(), C.c_int(), C.c_double()
        _check(lib().gd2_ensemble_tune(self._h, _stream(stream), C.byref(r), C.byref(b), C.byref(ms)),
               "gd2_ensemble_tune")
        return r.value, b.value, ms.value

    def close(self):
        if self._h:
            lib().gd2_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------ on-disk formats (host buffers)

def _host(g, a, last):
    """Pointer to a C-contiguous float64 host array of shape (ny, nx, last)."""
    if a is None:
        return None
    if isinstance(a, torch.Tensor):
        if a.is_cuda:
            raise GD2Error("host array expected (copy the state to the host first)")
        a = a.numpy()
    if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"] or a.shape != (g.ny, g.nx, last):
        raise GD2Error(f"C-contiguous float64 array of shape {(g.ny, g.nx, last)} expected")
    return C.c_void_p(a.ctypes.data)


def write_state_text(g, s, u, b, path):
    """'x y rho p vx vy Bx By Bz' per cell, as the reference's end-of-run
    writer (kernel.cu:4257-4260); b may be None."""
    _check(lib().gd2_write_state_text(C.byref(g), _host(g, s, 2), _host(g, u, 2), _host(g, b, 3),
                                      os.fsencode(path)), "gd2_write_state_text")


def read_state_text(g, path, mhd=True):
    """Restart reader (kernel.cu:3944-3969); returns host arrays s, u, b."""
    s = np.empty((g.ny, g.nx, 2))
    u = np.empty((g.ny, g.nx, 2))
    b = np.empty((g.ny, g.nx, 3)) if mhd else None
    _check(lib().gd2_read_state_text(C.byref(g), os.fsencode(path), _host(g, s, 2), _host(g, u, 2),
                                     _host(g, b, 3)), "gd2_read_state_text")
    return s, u, b


def write_state_b3(g, s, u, b, path):
    """save_b3.dat: x, y, rho, p, vx, vy, Bz per cell (kernel.cu:4261-4267)."""
    _check(lib().gd2_write_state_b3(C.byref(g), _host(g, s, 2), _host(g, u, 2), _host(g, b, 3),
                                    os.fsencode(path)), "gd2_write_state_b3")


def write_tecplot(g, s, u, b, path, nmin=3, solution_time=None):
    """Tecplot FEPOINT zone of every nmin-th cell (print_file_mini2 /
    print_file_mini, kernel.cu:3682-3808)."""
    t = None if solution_time is None else C.byref(C.c_double(solution_time))
    _check(lib().gd2_write_tecplot(C.byref(g), _host(g, s, 2), _host(g, u, 2), _host(g, b, 3), int(nmin), t,
                                   os.fsencode(path)), "gd2_write_tecplot")


def write_tecplot_derived(g, s, u, b, path, nmin=3):
    """param_for_texplot_mini.txt: every nmin-th cell with PP, Fx, Fy, zav
    (kernel.cu:4384-4460)."""
    _check(lib().gd2_write_tecplot_derived(C.byref(g), _host(g, s, 2), _host(g, u, 2), _host(g, b, 3), int(nmin),
                                           os.fsencode(path)), "gd2_write_tecplot_derived")


def write_line_cuts(g, s, u, b, path_y0=None, path_x0=None):
    """param_y=0.txt / param_x=0.txt: the first row's and first column's
    cells 5 .. n-2 with PP, Fx, Fy, dP (kernel.cu:4283-4374)."""
    enc = lambda p: None if p is None else os.fsencode(p)
    _check(lib().gd2_write_line_cuts(C.byref(g), _host(g, s, 2), _host(g, u, 2), _host(g, b, 3), enc(path_y0),
                                     enc(path_x0)), "gd2_write_line_cuts")
