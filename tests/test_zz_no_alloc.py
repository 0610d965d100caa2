"""The allocation-free adjoint entry point gd2_stage_vjp_noalloc.

It runs the kernels of gd2_stage_vjp on the library's fixed per-device
scratch array instead of a caller's workspace. Checked through the C ABI
against gd2_stage_vjp with its own workspace: two calls in a row, the second
on a smaller grid (the array keeps the first call's contents), for the Euler
HLLC adjoint with periodic sides (its ghost cotangents then go through the
scratch array) and for the MHD HLLD / HLLC-MHD adjoint (all of its face-state
cotangents live there). Sizes beyond the array are refused before any device
work.
"""
import ctypes

import numpy as np
import pytest
import torch

from golikov_2d_amd import api
from golikov_2d_amd import problems as PB
from tests.gpu_util import say, to_dev, to_host

T, P, R = api.BC_TRANSMISSIVE, api.BC_PERIODIC, api.BC_REFLECTIVE


def euler_state(g, seed):
    rng = np.random.default_rng(seed)
    sh = (g.ny, g.nx)
    prim = [1.0 + 0.3 * rng.random(sh), 1.0 + 0.3 * rng.random(sh), rng.uniform(-0.5, 0.5, sh),
            rng.uniform(-0.5, 0.5, sh)]
    return np.ascontiguousarray(PB.euler_from_prim(g.gamma, *prim))


def mhd_state(g, seed):
    rng = np.random.default_rng(seed)
    sh = (g.ny, g.nx)
    r, p = 1.0 + 0.3 * rng.random(sh), 1.0 + 0.3 * rng.random(sh)
    u, v = rng.uniform(-0.5, 0.5, sh), rng.uniform(-0.5, 0.5, sh)
    bx, by, bz = (rng.uniform(-2.5, 2.5, sh) for _ in range(3))
    return np.ascontiguousarray(PB.mhd_from_prim(g.gamma, r, p, u, v, bx, by, bz))


def max_rel(got, ref):
    return float(np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-300))


# (physics, flux, order, bc, [(nx, ny) of the first call, then of the smaller second call])
CASES = [
    (api.EULER, api.FLUX_HLLC, 2, (P, P, P, P), [(130, 70), (37, 29)]),
    (api.EULER, api.FLUX_HLLC, 2, (P, P, R, T), [(96, 80), (40, 24)]),
    (api.EULER, api.FLUX_HLLC, 1, (T, T, T, T), [(70, 46), (33, 17)]),
    (api.MHD, api.FLUX_HLLD, 2, (P, P, P, P), [(72, 60), (40, 24)]),
    (api.MHD, api.FLUX_HLLC_MHD, 2, (R, T, T, R), [(64, 48), (31, 23)]),
    (api.MHD, api.FLUX_HLLD, 1, (T, T, P, P), [(50, 40), (21, 13)]),
]


@pytest.mark.gpu
@pytest.mark.parametrize("physics,flux,order,bc,sizes", CASES)
def test_noalloc_vjp_equals_workspace_vjp(physics, flux, order, bc, sizes):
    sc = api.make_scheme(physics, flux, order, api.SSPRK2, 0.4)
    gamma = 5.0 / 3.0 if physics == api.MHD else 1.4
    for call, (nx, ny) in enumerate(sizes):
        say(f"stage_vjp_noalloc physics={physics} flux={flux} order={order} bc={bc} {nx}x{ny} call {call}")
        g = api.make_grid(nx, ny, gamma=gamma, bc=bc)
        assert api.vjp_workspace_bytes(g, sc) <= api.NOALLOC_SCRATCH_BYTES
        seed = 100 * call + nx
        U = (mhd_state if physics == api.MHD else euler_state)(g, seed)
        rng = np.random.default_rng(seed + 1)
        V = np.ascontiguousarray(rng.standard_normal(U.shape))
        A = np.ascontiguousarray(rng.standard_normal(U.shape))
        Ud, Vd, Ad = to_dev(U), to_dev(V), to_dev(A)
        W_ref = torch.empty_like(Ud)
        api.stage_vjp(g, sc, Ud, Vd, W_ref, c=0.7, A=Ad, a=0.5)
        # poison the output so that an entry the call does not write shows
        W = torch.full_like(Ud, float("nan"))
        api.stage_vjp_noalloc(g, sc, Ud, Vd, W, c=0.7, A=Ad, a=0.5)
        ref, got = to_host(W_ref), to_host(W)
        assert np.isfinite(got).all()
        err = max_rel(got, ref)
        say(f"  max rel diff {err:.2e}")
        assert err <= 1e-13, err


def test_noalloc_vjp_refuses_sizes_beyond_the_scratch_array():
    """Host-side checks only: refused before any device work (no GPU needed)."""
    import os
    if not os.path.exists(api.SO_PATH):
        pytest.skip("libgd2.so not built")
    L = api.lib()
    x, y = ctypes.c_void_p(16), ctypes.c_void_p(32)
    euler = api.make_scheme(api.EULER, api.FLUX_HLLC, 2, api.SSPRK2, 0.4)
    mhd = api.make_scheme(api.MHD, api.FLUX_HLLD, 2, api.SSPRK2, 0.3)
    big_e = api.make_grid(8, 131073)
    assert api.vjp_workspace_bytes(big_e, euler) > api.NOALLOC_SCRATCH_BYTES
    assert L.gd2_stage_vjp_noalloc(ctypes.byref(big_e), ctypes.byref(euler), x, x, None, 0.0, None, 0.0, 1.0, y,
                                   None) == api.ERR_TOO_LARGE
    big_m = api.make_grid(460, 460, gamma=5.0 / 3.0)
    assert api.vjp_workspace_bytes(big_m, mhd) > api.NOALLOC_SCRATCH_BYTES
    assert L.gd2_stage_vjp_noalloc(ctypes.byref(big_m), ctypes.byref(mhd), x, x, None, 0.0, None, 0.0, 1.0, y,
                                   None) == api.ERR_TOO_LARGE
    # the same argument checks as gd2_stage_vjp
    g = api.make_grid(16, 16)
    assert L.gd2_stage_vjp_noalloc(ctypes.byref(g), ctypes.byref(euler), None, x, None, 0.0, None, 0.0, 1.0, y,
                                   None) == -1
    assert L.gd2_stage_vjp_noalloc(ctypes.byref(g), ctypes.byref(euler), x, x, None, 0.0, None, 0.0, 1.0, x,
                                   None) == -1
    hll = api.make_scheme(api.EULER, api.FLUX_HLL, 2, api.SSPRK2, 0.4)
    assert L.gd2_stage_vjp_noalloc(ctypes.byref(g), ctypes.byref(hll), x, x, None, 0.0, None, 0.0, 1.0, y,
                                   None) == -2
