"""Synthetic initial states on the host (numpy), SoA conserved layout.

All random draws come from SplitMix64 (Steele, Lea & Flood 2014) so the CPU
and the GPU paths start from identical bits for a given seed.

Euler fields: rho, mx, my, E.  MHD fields: rho, mx, my, Bx, By, Bz, E with
Gaussian B (E = p/(gamma-1) + rho|v|^2/2 + |B|^2/(8 pi)), the state layout
of the reference's MHD kernels (kernel.cu:1877, Header.h:16).
"""
from __future__ import annotations

import math

import numpy as np

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))


def _centres(g):
    x = g.x0 + np.arange(g.nx) * g.dx
    y = g.y0 + np.arange(g.ny) * g.dy
    return np.meshgrid(x, y)  # (ny, nx)


def euler_from_prim(gamma, rho, p, u, v):
    U = np.empty((4,) + rho.shape)
    U[0] = rho
    U[1] = rho * u
    U[2] = rho * v
    U[3] = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    return U


def mhd_from_prim(gamma, rho, p, u, v, bx, by, bz):
    U = np.empty((7,) + rho.shape)
    U[0] = rho
    U[1] = rho * u
    U[2] = rho * v
    U[3], U[4], U[5] = bx, by, bz
    U[6] = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v) + (bx * bx + by * by + bz * bz) / 25.13274122871834590768
    return U


def uniform_euler(g, rho=1.0, p=1.0, u=0.0, v=0.0):
    one = np.ones((g.ny, g.nx))
    return np.ascontiguousarray(euler_from_prim(g.gamma, rho * one, p * one, u * one, v * one))


def uniform_mhd(g, rho=1.0, p=1.0, u=0.0, v=0.0, B=(0.0, 0.0, 0.0)):
    one = np.ones((g.ny, g.nx))
    return np.ascontiguousarray(mhd_from_prim(g.gamma, rho * one, p * one, u * one, v * one,
                                              B[0] * one, B[1] * one, B[2] * one))


def sod_x(g, x_split=0.5):
    """x-directed Sod problem: x < 0.5 (1, 0, 0, 1); else (0.125, 0, 0, 0.1)."""
    X, _ = _centres(g)
    left = X < x_split
    rho = np.where(left, 1.0, 0.125)
    p = np.where(left, 1.0, 0.1)
    z = np.zeros_like(rho)
    return np.ascontiguousarray(euler_from_prim(g.gamma, rho, p, z, z))


def four_quadrant_params(seed):
    """Split point in [0.3,0.7]^2, then per quadrant (SW, SE, NW, NE):
    rho ~ U[0.1,2], p ~ U[0.1,2], u, v ~ U[-1,1]."""
    r = SplitMix64(seed)
    xs, ys = r.uniform(0.3, 0.7), r.uniform(0.3, 0.7)
    quads = []
    for _ in range(4):
        quads.append((r.uniform(0.1, 2.0), r.uniform(0.1, 2.0), r.uniform(-1.0, 1.0), r.uniform(-1.0, 1.0)))
    return xs, ys, quads


def four_quadrant(g, seed=0, domain=(1.0, 1.0)):
    """Seeded four-quadrant Riemann problem on [0,Lx]x[0,Ly] (split in the unit square scaled)."""
    xs, ys, quads = four_quadrant_params(seed)
    X, Y = _centres(g)
    east = X >= xs * domain[0]
    north = Y >= ys * domain[1]
    idx = east.astype(int) + 2 * north.astype(int)
    tab = np.array(quads)  # (4, [rho, p, u, v])
    q = tab[idx]
    return np.ascontiguousarray(euler_from_prim(g.gamma, q[..., 0], q[..., 1], q[..., 2], q[..., 3]))


def smooth_perturbation(g, seed, amp, nmodes=4):
    """Sum of seeded smooth Fourier modes (integer wavenumbers 1..3), amplitude amp."""
    r = SplitMix64(seed ^ 0x5DEECE66D)
    X, Y = _centres(g)
    d = np.zeros((g.ny, g.nx))
    for _ in range(nmodes):
        kx = 1 + int(r.uniform(0, 3))
        ky = 1 + int(r.uniform(0, 3))
        ph = r.uniform(0, 2 * math.pi)
        d += np.sin(2 * math.pi * (kx * X + ky * Y) + ph)
    return amp * d / nmodes


def smooth_perturbed_euler(g, U, seed, amp):
    """Euler state U with seeded smooth perturbations of relative amplitude amp
    on rho and p and absolute amplitude amp on u and v."""
    gm1 = g.gamma - 1.0
    rho = U[0]
    u, v = U[1] / rho, U[2] / rho
    p = gm1 * (U[3] - 0.5 * rho * (u * u + v * v))
    rho = rho * (1.0 + smooth_perturbation(g, seed + 1, amp))
    p = p * (1.0 + smooth_perturbation(g, seed + 2, amp))
    u = u + smooth_perturbation(g, seed + 3, amp)
    v = v + smooth_perturbation(g, seed + 4, amp)
    return np.ascontiguousarray(euler_from_prim(g.gamma, rho, p, u, v))


def orszag_tang(g, seed=0, amp=1e-3):
    """Orszag-Tang vortex on [0,1]^2 (periodic), gamma from g:
    rho = 25/(36 pi), p = 5/(12 pi), v = (-sin 2 pi y, sin 2 pi x),
    Gaussian B = (-sin 2 pi y, sin 4 pi x) (= sqrt(4 pi) times the
    Heaviside-Lorentz field), Bz = vz = 0, each of rho, p, vx, vy, Bx, By
    plus seeded Fourier-phase perturbations of amplitude amp."""
    X, Y = _centres(g)
    r = SplitMix64(seed)
    modes = [(1, 0), (0, 1), (1, 1), (2, 1)]

    def pert():
        ph = [r.uniform(0, 2 * math.pi) for _ in modes]
        return amp * sum(np.sin(2 * math.pi * (kx * X + ky * Y) + p) for (kx, ky), p in zip(modes, ph)) / len(modes)

    rho = 25.0 / (36.0 * math.pi) + pert()
    p = 5.0 / (12.0 * math.pi) + pert()
    u = -np.sin(2 * math.pi * Y) + pert()
    v = np.sin(2 * math.pi * X) + pert()
    bx = -np.sin(2 * math.pi * Y) + pert()
    by = np.sin(4 * math.pi * X) + pert()
    bz = np.zeros_like(rho)
    return np.ascontiguousarray(mhd_from_prim(g.gamma, rho, p, u, v, bx, by, bz))


def ensemble_sizes(seed, count=256, lo=64, hi=1024):
    """Side lengths drawn log-uniform on [lo, hi] independently for nx and ny."""
    r = SplitMix64(seed)
    out = []
    for _ in range(count):
        nx = int(round(math.exp(r.uniform(math.log(lo), math.log(hi)))))
        ny = int(round(math.exp(r.uniform(math.log(lo), math.log(hi)))))
        out.append((nx, ny))
    return out
