"""Differentiable SSP-RK2 rollouts of the Euler HLLC and MHD HLLC-MHD / HLLD
MUSCL schemes.

`rollout` is a LibTorch autograd Function (gd2_torch, C++ over the libgd2 C
ABI): the forward pass runs the HIP stage kernels with a prescribed dt
schedule and keeps a checkpoint every `every` steps (and the stage states of
the last segment, where the reverse sweep starts); the backward pass
recomputes one segment at a time from its checkpoint, on a second stream
while the adjoint of the segment after it runs, and runs the stage VJP
kernels (gd2_stage_vjp) backwards through it. Gradients are with respect to
the initial conserved state, with the dt schedule held fixed.

`dt_schedule` records the CFL time steps that the native solver takes from a
given state, so a rollout can replay them.
"""
from __future__ import annotations

import os

import torch

from . import api
from . import build as _build

_mod = None


def module():
    """The compiled gd2_torch module (raises if it has not been built)."""
    global _mod
    if _mod is None:
        path = _build.torch_module_path()
        if not os.path.exists(path):
            raise api.GD2Error(f"{path} is missing: build it with golikov_2d_amd/build.py --torch "
                               "(or __graft_entry__.build()); there is no CPU fallback")
        api.lib()
        from . import gd2_torch
        _mod = gd2_torch
    return _mod


def _geom(g, sc):
    return [float(g.nx), float(g.ny), g.dx, g.dy, g.x0, g.y0, g.gamma, g.krit, sc.cfl]


def _check(g, sc):
    euler = sc.physics == api.EULER and sc.flux == api.FLUX_HLLC
    mhd = sc.physics == api.MHD and sc.flux in (api.FLUX_HLLC_MHD, api.FLUX_HLLD)
    if not (euler or mhd) or sc.integrator != api.SSPRK2:
        raise api.GD2Error("differentiable rollouts cover SSP-RK2 with Euler + HLLC or MHD + HLLC-MHD / HLLD")


def _scheme(sc):
    return [int(sc.physics), int(sc.flux), int(sc.order)]


def _stream(stream):
    s = torch.cuda.current_stream() if stream is None else stream
    return int(s.cuda_stream)


def rollout(g, sc, U0, dts, every=20, stream=None):
    """U_T after len(dts) SSP-RK2 steps from U0 (shape (nfields, ny, nx), float64,
    device). Differentiable with respect to U0 through torch autograd."""
    _check(g, sc)
    return module().rollout(U0, dts, _geom(g, sc), list(g.bc), _scheme(sc), int(every), _stream(stream))


def rollout_vjp(g, sc, U0, dts, lamT, every=20, stream=None):
    """(U_T, (dU_T/dU0)^T lamT) without building an autograd graph."""
    _check(g, sc)
    UT, lam0 = module().rollout_vjp(U0, dts, lamT, _geom(g, sc), list(g.bc), _scheme(sc), int(every),
                                    _stream(stream))
    return UT, lam0


def dt_schedule(g, sc, U0, nsteps):
    """CFL steps the native solver takes from U0: (host float64 tensor of dts, U after them)."""
    _check(g, sc)
    U = U0.detach().clone()
    log = torch.zeros(int(nsteps), dtype=torch.float64, device=U.device)
    s = api.Solver(g, sc)
    try:
        s.log_dt(log)
        s.start(U)
        s.advance(U, int(nsteps))
        s.log_dt(None)
        torch.cuda.synchronize()
    finally:
        s.close()
    return log.cpu(), U


def density_loss(UT, rho_target):
    """1/2 || rho(T) - rho_target ||^2"""
    d = UT[0] - rho_target
    return 0.5 * (d * d).sum()
