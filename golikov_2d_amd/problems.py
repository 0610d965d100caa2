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


# ------------------------------------------------- the reference's production run
# Constants of the stellar-wind run (Header.h:25-37): wind speed phi_0, wind
# Mach number M_0, Alfven Mach number M_alf, reference radius rr_0, inner
# frozen sphere radius ddist; grid N x M on (0, 9] x (0, 5] with cell
# centres at (n + 1/2) dx, (m + 1/2) dy (kernel.cu:12-20).
WIND_PHI0, WIND_M0, WIND_MALF, WIND_RR0, WIND_DDIST = 4.0, 10.0, 8.0, 0.07, 0.3
WIND_NX, WIND_NY, WIND_XMAX, WIND_YMAX = 3226, 1792, 9.0, 5.0


def wind_grid_args(nx=WIND_NX, ny=WIND_NY, x_max=WIND_XMAX, y_max=WIND_YMAX):
    """Keyword arguments of make_grid for the production run: the four
    production side kinds (symmetry west, outflow east, axis south, inflow
    north), the frozen inner sphere and the axisymmetric sources."""
    return dict(dx=x_max / nx, dy=y_max / ny, x0=x_max / (2.0 * nx), y0=y_max / (2.0 * ny),
                gamma=5.0 / 3.0, krit=0.25, bc=(4, 6, 5, 7), r_frozen=WIND_DDIST, axisym=True)


def stellar_wind_prim(g):
    """Initial primitive state of the production run (kernel.cu:3901-3941):
    inside 1.01 ddist a radial wind rho ~ r^-2, p ~ r^-2gamma, |v| = phi_0 with
    an azimuthal field B_phi ~ sin(theta) / r; outside, rho = 1, p = 1/gamma,
    v = 0, B = 0. Returns AoS arrays s (ny, nx, 2) = (rho, p), u (ny, nx, 2),
    b (ny, nx, 3)."""
    gam = g.gamma
    x, y = _centres(g)
    dist = np.sqrt(x * x + y * y)
    ro_e = 1.0 / (WIND_PHI0 * WIND_PHI0 * WIND_RR0 * WIND_RR0)
    p_e = ro_e * WIND_PHI0 * WIND_PHI0 / (gam * WIND_M0 * WIND_M0)
    b_e = math.sqrt(4.0 * math.pi) / (WIND_MALF * WIND_RR0)
    inner = dist < WIND_DDIST * 1.01
    s = np.empty((g.ny, g.nx, 2))
    u = np.zeros((g.ny, g.nx, 2))
    b = np.zeros((g.ny, g.nx, 3))
    s[..., 0] = 1.0
    s[..., 1] = 1.0 / gam
    d = dist[inner]
    s[inner, 0] = ro_e / (d / WIND_RR0) ** 2
    s[inner, 1] = p_e * (WIND_RR0 / d) ** (2.0 * gam)
    u[inner, 0] = WIND_PHI0 * x[inner] / d
    u[inner, 1] = WIND_PHI0 * y[inner] / d
    # Cartesian image of the azimuthal component B_E (rr_0 / r) sin(theta)
    # (kernel.cu:155-173 at phi = 0): it lands in the third component
    bphi = b_e / (d / WIND_RR0) * np.sin(np.arccos(x[inner] / d))
    b[inner, 2] = np.where(np.abs(y[inner]) < 1e-6, 0.0, bphi)
    return s, u, b
