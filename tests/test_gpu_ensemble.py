"""Batched ensembles on the GPU: bitwise equal to the per-problem path, to
itself under hipGraph replay and under every tiling the tuner may choose
(including a persistent launch of one wave, which takes the whole work list
in order, and launches whose waves wait on each other's steps), with odd and
even step counts, and within the run tolerance of the oracle."""
import numpy as np
import pytest
import torch

from golikov_2d_amd import api
from golikov_2d_amd import problems as PB
from oracle import oracle as O
from tests.gpu_util import rel_l1, say, to_dev, to_host

pytestmark = pytest.mark.gpu

BCS = [(0, 0, 0, 0), (1, 1, 1, 1), (2, 0, 1, 1), (0, 2, 0, 2)]


def make_problems(count, lo, hi, seed):
    sizes = PB.ensemble_sizes(seed, count=count, lo=lo, hi=hi)
    grids, ogrids, states = [], [], []
    for p, (nx, ny) in enumerate(sizes):
        bc = BCS[p % len(BCS)]
        grids.append(api.make_grid(nx, ny, bc=bc))
        og = O.make_grid(nx, ny, bc=bc)
        ogrids.append(og)
        states.append(PB.four_quadrant(og, seed=seed * 1000 + p))
    return grids, ogrids, states


@pytest.mark.parametrize("flux,order,steps", [(api.FLUX_HLLC, 2, 10), (api.FLUX_HLL, 2, 10), (api.FLUX_HLLC, 2, 7),
                                              (api.FLUX_HLLC, 1, 5), (api.FLUX_EXACT, 2, 4), (api.FLUX_EXACT, 1, 3)])
def test_ensemble_bitwise_equals_individual_runs(flux, order, steps):
    say(f"ensemble vs individual flux={flux} order={order} steps={steps}")
    grids, _, states = make_problems(12, 16, 200, seed=3)
    sc = api.make_scheme(api.EULER, flux, order)
    ens_U = [to_dev(U) for U in states]
    ind_U = [to_dev(U) for U in states]
    e = api.Ensemble(grids, sc, ens_U)
    e.start()
    e.advance(steps)
    for g, U in zip(grids, ind_U):
        s = api.Solver(g, sc)
        s.start(U)
        s.advance(U, steps)
        s.close()
    torch.cuda.synchronize()
    for p, (a, b) in enumerate(zip(ens_U, ind_U)):
        assert torch.equal(a, b), p
        t, n, _ = e.time(p)
        assert n == steps and t > 0
    e.close()


def test_ensemble_waves_waiting_on_each_other():
    """40 problems, 20 of them single-tile (16..40 cells a side): with 96
    waves consecutive steps of one problem are taken by different waves, which
    must wait for each other; 12 steps, bitwise equal to per-problem runs."""
    say("ensemble dependent waves")
    g1, _, s1 = make_problems(20, 16, 40, seed=11)
    g2, _, s2 = make_problems(20, 60, 300, seed=12)
    grids, states = g1 + g2, s1 + s2
    sc = api.make_scheme()
    ens_U = [to_dev(U) for U in states]
    ind_U = [to_dev(U) for U in states]
    e = api.Ensemble(grids, sc, ens_U)
    e.set_tiling(32, 96)
    e.start()
    e.advance(12, use_graph=True)
    for g, U in zip(grids, ind_U):
        s = api.Solver(g, sc)
        s.start(U)
        s.advance(U, 12)
        s.close()
    torch.cuda.synchronize()
    for p, (a, b) in enumerate(zip(ens_U, ind_U)):
        assert torch.equal(a, b), p
        assert e.time(p)[1] == 12
    e.close()


def test_ensemble_graph_replays_match_eager():
    say("ensemble graph replay")
    grids, _, states = make_problems(9, 16, 150, seed=4)
    sc = api.make_scheme()
    Ua = [to_dev(U) for U in states]
    Ub = [to_dev(U) for U in states]
    ea, eb = api.Ensemble(grids, sc, Ua), api.Ensemble(grids, sc, Ub)
    ea.start()
    eb.start()
    for rep in range(3):
        ea.advance(5, use_graph=False)
        eb.advance(5, use_graph=True)
        torch.cuda.synchronize()
        junk = torch.randn(1000, device="cuda").sum().item()  # eager work between replays
        assert np.isfinite(junk)
        for p in range(len(grids)):
            assert torch.equal(Ua[p], Ub[p]), (rep, p)
    # odd step counts alternate the clock slots: still equal
    ea.advance(3)
    eb.advance(3, use_graph=True)
    eb.advance(3, use_graph=True)
    ea.advance(3)
    torch.cuda.synchronize()
    for p in range(len(grids)):
        assert torch.equal(Ua[p], Ub[p])
    ea.close()
    eb.close()


@pytest.mark.parametrize("rows,blocks", [(8, 0), (64, 0), (16, 37), (32, 1), (2, 5), (1, 3), (256, 0), (4, 100000)])
def test_ensemble_result_independent_of_tiling(rows, blocks):
    say(f"ensemble tiling rows={rows} blocks={blocks}")
    grids, _, states = make_problems(7, 16, 130, seed=5)
    sc = api.make_scheme()
    Ua = [to_dev(U) for U in states]
    Ub = [to_dev(U) for U in states]
    ea, eb = api.Ensemble(grids, sc, Ua), api.Ensemble(grids, sc, Ub)
    eb.set_tiling(rows, blocks)
    ea.start()
    eb.start()
    ea.advance(6)
    eb.advance(6)
    torch.cuda.synchronize()
    for p in range(len(grids)):
        assert torch.equal(Ua[p], Ub[p]), p
    ea.close()
    eb.close()


def test_ensemble_tuner_keeps_results():
    say("ensemble tuner")
    grids, _, states = make_problems(16, 32, 300, seed=6)
    sc = api.make_scheme()
    Ua = [to_dev(U) for U in states]
    Ub = [to_dev(U) for U in states]
    ea, eb = api.Ensemble(grids, sc, Ua), api.Ensemble(grids, sc, Ub)
    ea.start()
    eb.start()
    rows, blocks, ms = eb.tune()
    assert rows in (32, 64, 128, 256) and blocks > 0 and ms > 0
    assert eb.tiling()[:2] == (rows, blocks)
    for p in range(len(grids)):  # tuning writes only scratch
        assert torch.equal(Ub[p], Ua[p])
    ea.advance(4)
    eb.advance(4, use_graph=True)
    torch.cuda.synchronize()
    for p in range(len(grids)):
        assert torch.equal(Ua[p], Ub[p])
    ea.close()
    eb.close()


def test_ensemble_matches_oracle_50_steps():
    say("ensemble vs oracle 50 steps")
    grids, ogrids, states = make_problems(5, 64, 160, seed=7)
    sc = api.make_scheme()
    Ud = [to_dev(U) for U in states]
    e = api.Ensemble(grids, sc, Ud)
    e.start()
    e.advance(50, use_graph=True)
    torch.cuda.synchronize()
    osc = O.Scheme(O.PHYS_EULER, O.FLUX_HLLC, 2, 0.4)
    for p, (og, U) in enumerate(zip(ogrids, states)):
        Uo = U.copy()
        O.run(og, osc, 2, Uo, nsteps=50)
        err = rel_l1(to_host(Ud[p]), Uo)
        assert np.all(err <= 1e-10), (p, err)
    e.close()
