"""The production-run driver (golikov_2d_amd/gd2_production, the reference's
main step loop over the C ABI, kernel.cu:3810-4469) against the oracle's
loop of the same steps.

The driver runs as a child process. At a tenth of the production size it
advances 10 add2 and 10 Kernel_TVD steps (HLLC-MHD and HLLD), each followed by
funk_time, snapshotting every 4 steps of either phase. Its save_b3.dat must
match the oracle's loop to 1e-10, like the in-process 20-step wind test. Its
printed text files must match the oracle's writers on the oracle's state to
the 6 digits they carry, including the solution time of the second phase's
snapshots. With no steps, save_b3.dat must hold the stellar-wind initial
state.
"""
import os
import subprocess

import numpy as np
import pytest

from golikov_2d_amd import build as B
from golikov_2d_amd import problems as PB
from oracle import oracle as O
from tests.gpu_util import say

pytestmark = pytest.mark.gpu

NX, NY = 322, 179


def run_driver(out, *args):
    if not os.path.exists(B.PRODUCTION):
        raise FileNotFoundError(f"{B.PRODUCTION} is not built (golikov_2d_amd.build.build_production)")
    say(f"gd2_production {' '.join(args)}")
    r = subprocess.run([B.PRODUCTION, "--nx", str(NX), "--ny", str(NY), "--out", str(out), *args],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    return r.stdout


def read_b3(path):
    return np.fromfile(path, dtype=np.float64).reshape(NY, NX, 7)


def oracle_grid():
    kw = PB.wind_grid_args(NX, NY)
    return O.make_grid(NX, NY, **kw)


def read_table(path):
    """numeric rows of a Tecplot / column text file and its header lines"""
    head, rows = [], []
    with open(path) as f:
        for line in f:
            parts = line.split()
            try:
                rows.append([float(p) for p in parts])
            except ValueError:
                head.append(line)
    return head, rows


def same_table(got, ref, rtol=1e-5):
    """same header lines; numbers to rtol, with an absolute floor of 1e-7 of
    the column's largest magnitude for values that cross zero"""
    hg, rg = read_table(got)
    hr, rr = read_table(ref)
    assert hg == hr
    assert [len(r) for r in rg] == [len(r) for r in rr]
    for width in sorted({len(r) for r in rr}):
        a = np.array([r for r in rg if len(r) == width])
        b = np.array([r for r in rr if len(r) == width])
        if width == 0:
            continue
        floor = 1e-7 * np.abs(b).max(axis=0)
        bad = np.abs(a - b) > rtol * np.abs(b) + floor
        assert not bad.any(), (got, np.argwhere(bad)[:5], a[bad][:5], b[bad][:5])


def test_driver_initial_state(tmp_path):
    run_driver(tmp_path, "--steps1", "0", "--steps2", "0")
    og = oracle_grid()
    s, u, b = PB.stellar_wind_prim(og)
    a = read_b3(tmp_path / "save_b3.dat")
    x = og.x0 + np.arange(NX) * og.dx
    y = og.y0 + np.arange(NY) * og.dy
    np.testing.assert_array_equal(a[..., 0], np.broadcast_to(x, (NY, NX)))
    np.testing.assert_array_equal(a[..., 1], np.broadcast_to(y[:, None], (NY, NX)))
    np.testing.assert_allclose(a[..., 2:4], s, rtol=1e-14, atol=0)
    np.testing.assert_allclose(a[..., 4:6], u, rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(a[..., 6], b[..., 2], rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("method", [2, 3])
def test_driver_run_matches_oracle_loop(tmp_path, method):
    out = run_driver(tmp_path, "--steps1", "10", "--steps2", "10", "--snap1", "4", "--snap2", "4",
                     "--method", str(method))
    og = oracle_grid()
    s, u, b = PB.stellar_wind_prim(og)
    T, Tdo, TT, it = 1e7, 1e-8, 0.0, 0
    ref = tmp_path / "ref"
    ref.mkdir()
    tt_at = {}
    for k in range(20):
        kind = "add2" if k < 10 else "kernel_tvd"
        s, u, b, T = O.ref_mhd_step(kind, og, s, u, b, T, Tdo, it, method)
        T, Tdo, TT, it = O.funk_time(T, Tdo, TT, it)
        # the pass with counter i ends after step i + 2 of its phase
        i = (k % 10) - 1
        if i > 0 and i % 4 == 0 and k % 2 == 1:
            if k < 10:
                O.write_tecplot(og, s, u, b, str(ref / f"02_12_{i}.txt"))
            else:
                tt_at[i] = TT
                O.write_tecplot(og, s, u, b, str(ref / f"11_01_{i}.txt"), solution_time=TT - tt_at[4])
    assert sorted(os.listdir(ref)) == ["02_12_4.txt", "02_12_8.txt", "11_01_4.txt", "11_01_8.txt"]
    # final state
    a = read_b3(tmp_path / "save_b3.dat")
    assert np.max(np.abs(a[..., 2:4] - s) / np.abs(s)) <= 1e-10
    assert np.max(np.abs(a[..., 4:6] - u)) <= 1e-10
    assert np.max(np.abs(a[..., 6] - b[..., 2])) <= 1e-10
    assert f"TT = {TT:g}" in out
    # snapshots and end-of-run text files against the oracle's writers
    for name in sorted(os.listdir(ref)):
        same_table(tmp_path / name, ref / name)
    O.write_state_text(og, s, u, b, str(ref / "instable_for_inner_b3.txt"))
    O.write_tecplot_derived(og, s, u, b, str(ref / "param_for_texplot_mini.txt"))
    O.write_line_cuts(og, s, u, b, str(ref / "param_y=0.txt"), str(ref / "param_x=0.txt"))
    for name in ["instable_for_inner_b3.txt", "param_for_texplot_mini.txt", "param_y=0.txt", "param_x=0.txt"]:
        same_table(tmp_path / name, ref / name, rtol=1e-4)


def test_driver_restart_reimposes_wind(tmp_path):
    """--restart: the text state is read back (6 significant digits) and the
    wind imposed again inside 1.001 ddist (kernel.cu:3944-3996); with no steps
    save_b3.dat holds exactly that."""
    og = oracle_grid()
    s, u, b = PB.stellar_wind_prim(og)
    T, Tdo, TT, it = 1e7, 1e-8, 0.0, 0
    for _ in range(4):
        s, u, b, T = O.ref_mhd_step("add2", og, s, u, b, T, Tdo, it, 2)
        T, Tdo, TT, it = O.funk_time(T, Tdo, TT, it)
    path = str(tmp_path / "instable_cp14.txt")
    O.write_state_text(og, s, u, b, path)
    rs, ru, rb = O.read_state_text(og, path)
    run_driver(tmp_path, "--steps1", "0", "--steps2", "0", "--restart", path)
    ws, wu, wb = PB.stellar_wind_prim(og)
    x = og.x0 + np.arange(NX) * og.dx
    y = og.y0 + np.arange(NY) * og.dy
    X, Y = np.meshgrid(x, y)
    inner = np.sqrt(X * X + Y * Y) < PB.WIND_DDIST * 1.001
    assert inner.any() and not inner.all()
    es = np.where(inner[..., None], ws, rs)
    eu = np.where(inner[..., None], wu, ru)
    eb = np.where(inner, wb[..., 2], rb[..., 2])
    a = read_b3(tmp_path / "save_b3.dat")
    np.testing.assert_array_equal(a[~inner][:, 2:4], es[~inner])
    np.testing.assert_array_equal(a[~inner][:, 4:6], eu[~inner])
    np.testing.assert_array_equal(a[~inner][:, 6], eb[~inner])
    np.testing.assert_allclose(a[inner][:, 2:4], es[inner], rtol=1e-14, atol=0)
    np.testing.assert_allclose(a[inner][:, 4:6], eu[inner], rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(a[inner][:, 6], eb[inner], rtol=1e-14, atol=1e-15)
    # the ring 1.001 ddist <= r < 1.01 ddist keeps the file's (evolved) values
    ring = (np.sqrt(X * X + Y * Y) < PB.WIND_DDIST * 1.01) & ~inner
    assert ring.any()
