"""GPU: the reference's interface fluxes evaluated pointwise
(gd2_euler_face_fluxes, gd2_mhd_face_fluxes) against the oracle's
restatement of each function, face by face.

Every Euler flux function of the reference (HLL, HLLC_Aleksashov, the exact
Godunov solver, HLLC_Aleksashov2, HLLC_Korolkov, HLLC_Korolkov_2D and the
passive-scalar HLLCQ_Korolkov_2D, HLLCQ_Aleksashov) on random states with the
four axis normals and random unit normals, including supersonic pairs; the MHD
HLLDQ_Korolkov (metod 2 and 3) on random strong-field states with out-of-plane
velocity and field, all six axis normals, jumping normal field. Bar: each
output within 1e-12 of the oracle relative to the largest output of that face
(the GPU forms reciprocals and square roots from the hardware estimates plus
refinement steps: a few units of 2^-52), face time steps to 1e-13.
"""
import math

import numpy as np
import pytest
import torch

from golikov_2d_amd import api
from oracle import oracle as O
from tests.gpu_util import dev, say, to_dev, to_host

pytestmark = pytest.mark.gpu

KINDS = {
    api.FACE_HLL: "hll",
    api.FACE_HLLC_ALEKSASHOV: "hllc",
    api.FACE_GODUNOV: "godunov",
    api.FACE_HLLC_ALEKSASHOV2: "hllc_aleksashov2",
    api.FACE_HLLC_KOROLKOV: "hllc_korolkov",
    api.FACE_HLLC_KOROLKOV_2D: "hllc_korolkov_2d",
    api.FACE_HLLCQ_KOROLKOV_2D: "hllcq_korolkov_2d",
    api.FACE_HLLCQ_ALEKSASHOV: "hllcq_aleksashov",
}
Q_KINDS = (api.FACE_HLLCQ_KOROLKOV_2D, api.FACE_HLLCQ_ALEKSASHOV)


def euler_faces(n, seed):
    rng = np.random.default_rng(seed)
    Ls = np.stack([rng.uniform(0.1, 2.0, n), rng.uniform(0.1, 2.0, n)], 1)
    Rs = np.stack([rng.uniform(0.1, 2.0, n), rng.uniform(0.1, 2.0, n)], 1)
    Lu = rng.uniform(-1.0, 1.0, (n, 2))
    Ru = rng.uniform(-1.0, 1.0, (n, 2))
    shift = np.where(rng.random(n) < 0.1, rng.choice([-4.0, 4.0], n), 0.0)  # supersonic pairs
    Lu += shift[:, None]
    Ru += shift[:, None]
    axes = np.array([[1.0, 0.0], [0.0, -1.0], [-1.0, 0.0], [0.0, 1.0]])
    nrm = axes[np.arange(n) % 4]
    ang = rng.uniform(0, 2 * np.pi, n)
    general = np.arange(n) % 5 == 4  # every fifth face: a random unit normal
    nrm[general] = np.stack([np.cos(ang), np.sin(ang)], 1)[general]
    QL, QR = rng.uniform(0.1, 3.0, n), rng.uniform(0.1, 3.0, n)
    return Ls, Lu, Rs, Ru, np.ascontiguousarray(nrm), QL, QR


@pytest.mark.parametrize("kind", list(KINDS))
@pytest.mark.parametrize("gamma", [1.4, 5.0 / 3.0])
def test_euler_face_fluxes_match_oracle(kind, gamma):
    name = KINDS[kind]
    say(f"euler face fluxes {name} gamma={gamma:.4f}")
    n = 3000
    Ls, Lu, Rs, Ru, nrm, QL, QR = euler_faces(n, 100 + kind)
    krit, rad = 0.25, 0.01
    Ps, Pu, PQ, dt = api.euler_face_fluxes(kind, gamma, krit, to_dev(Ls), to_dev(Lu), to_dev(Rs), to_dev(Ru),
                                           to_dev(nrm), rad, to_dev(QL), to_dev(QR))
    Ps, Pu, dt = to_host(Ps), to_host(Pu), to_host(dt)
    PQ = to_host(PQ) if PQ is not None else None
    worst, worst_t = 0.0, 0.0
    for i in range(n):
        L, R = (*Ls[i], *Lu[i]), (*Rs[i], *Ru[i])
        if kind in Q_KINDS:
            F, q, t = O.euler_flux_q(name, gamma, krit, L, R, QL[i], QR[i], nrm[i], rad)
            exp, got = np.array([*F, q]), np.array([Ps[i, 0], Ps[i, 1], Pu[i, 0], Pu[i, 1], PQ[i]])
        else:
            F, t = O.euler_flux(name, gamma, krit, L, R, nrm[i], rad)
            exp, got = np.array(F), np.array([Ps[i, 0], Ps[i, 1], Pu[i, 0], Pu[i, 1]])
        worst = max(worst, float(np.max(np.abs(got - exp)) / (np.max(np.abs(exp)) + 1e-300)))
        worst_t = max(worst_t, abs(dt[i] - t) / t)
    say(f"  worst {worst:.2e}, dt {worst_t:.2e}")
    assert worst <= 1e-12
    assert worst_t <= 1e-13


def test_euler_face_fluxes_reject_unknown_kinds_and_empty_batches():
    x = torch.zeros((4, 2), dtype=torch.float64, device=dev())
    with pytest.raises(api.GD2Error):
        api.euler_face_fluxes(8, 1.4, 0.25, x, x, x, x, x)
    with pytest.raises(api.GD2Error):
        api.mhd_face_fluxes(1, 5 / 3, 0.25, torch.zeros((2, 8), dtype=torch.float64, device=dev()),
                            torch.zeros((2, 8), dtype=torch.float64, device=dev()),
                            torch.zeros((2, 3), dtype=torch.float64, device=dev()))
    e = torch.zeros((0, 2), dtype=torch.float64, device=dev())
    Ps, Pu, _, dt = api.euler_face_fluxes(api.FACE_HLL, 1.4, 0.25, e, e, e, e, e)
    assert Ps.shape == (0, 2) and dt.shape == (0,)


@pytest.mark.parametrize("metod", [2, 3])
def test_mhd_face_fluxes_match_oracle(metod):
    say(f"mhd face fluxes metod={metod}")
    n = 3000
    rng = np.random.default_rng(200 + metod)
    L = np.stack([rng.uniform(0.2, 2, n), rng.uniform(0.1, 2, n), *rng.uniform(-1, 1, (3, n)),
                  *rng.uniform(-2 * math.sqrt(4 * math.pi), 2 * math.sqrt(4 * math.pi), (3, n))], 1)
    R = np.stack([rng.uniform(0.2, 2, n), rng.uniform(0.1, 2, n), *rng.uniform(-1, 1, (3, n)),
                  *rng.uniform(-2 * math.sqrt(4 * math.pi), 2 * math.sqrt(4 * math.pi), (3, n))], 1)
    axes = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, 0, 0], [0, -1, 0], [0, 0, -1]], dtype=np.float64)
    nrm = np.ascontiguousarray(axes[np.arange(n) % 6])
    same = np.arange(n) % 3 == 0  # a third of the faces with a continuous normal field
    for i in np.nonzero(same)[0]:
        k = int(np.argmax(np.abs(nrm[i])))
        R[i, 5 + k] = L[i, 5 + k]
    krit, rad = 0.25, 0.02
    P, dt = api.mhd_face_fluxes(metod, 5.0 / 3.0, krit, to_dev(L), to_dev(R), to_dev(nrm), rad)
    P, dt = to_host(P), to_host(dt)
    worst, worst_t = 0.0, 0.0
    for i in range(n):
        exp, t = O.mhd_flux(5.0 / 3.0, krit, L[i], R[i], nrm[i], metod, rad)
        worst = max(worst, float(np.max(np.abs(P[i] - exp)) / (np.max(np.abs(exp)) + 1e-300)))
        worst_t = max(worst_t, abs(dt[i] - t) / t)
    say(f"  worst {worst:.2e}, dt {worst_t:.2e}")
    assert worst <= 1e-12
    assert worst_t <= 1e-13
