"""The reference's on-disk formats (host side, no GPU): text state / restart,
save_b3.dat, Tecplot FEPOINT (kernel.cu:3682-3808, 3944-3969, 4252-4267).

* The library's writers produce files byte-identical to the oracle's stdio
  restatement, on states with signed zeros, integers, tiny and huge
  magnitudes, for Euler and MHD, several grid shapes and subsampling steps.
* The oracle's text is pinned independently: every number equals Python's
  own '%g' rendering (the reference's std::ofstream default is printf %g,
  6 significant digits; a NaN keeps its sign, as glibc prints it), and the
  Tecplot header, node count, element count
  and 1-based quadrilateral connectivity are rebuilt here from the
  definition of an FEPOINT quadrilateral zone.
* Round trips: text -> reader gives each value rounded to 6 significant
  digits (the reference's restart precision); the library reader equals the
  oracle's fscanf reader; save_b3.dat read back with numpy equals the state
  bit for bit.
"""
import os

import numpy as np
import pytest

from golikov_2d_amd import api
from golikov_2d_amd import problems as PB
from oracle import oracle as O

pytestmark = pytest.mark.skipif(not os.path.exists(api.SO_PATH), reason="libgd2.so not built")


def state(ny, nx, seed, mhd=True):
    rng = np.random.default_rng(seed)
    s = rng.uniform(0.1, 3.0, (ny, nx, 2))
    u = rng.normal(0.0, 2.0, (ny, nx, 2))
    b = rng.normal(0.0, 1.0, (ny, nx, 3)) if mhd else None
    flat = [s.reshape(-1), u.reshape(-1)] + ([b.reshape(-1)] if mhd else [])
    specials = [0.0, -0.0, 1.0, -2.0, 1e-300, -3.5e-310, 1e300, 123456.5, 999999.5, 0.1, 1e-5, 12345678.0]
    for a in flat:
        idx = rng.choice(a.size, size=min(a.size, len(specials)), replace=False)
        a[idx] = specials[:len(idx)]
    s[0, 0, 0] = 0.0     # rho = 0: the Tecplot writer's Max/T guard
    s[-1, -1, 0] = -1.0
    return s, u, b


def fmt(v):
    """printf %g as glibc writes it (Python's '%g' drops the sign of a NaN)."""
    v = float(v)
    if np.isnan(v):
        return "-nan" if np.copysign(1.0, v) < 0 else "nan"
    return "%g" % v


def grids(nx, ny, **kw):
    return api.make_grid(nx, ny, **kw), O.make_grid(nx, ny, **kw)


SHAPES = [(7, 5), (1, 4), (10, 1), (31, 17)]


@pytest.mark.parametrize("nx,ny", SHAPES)
@pytest.mark.parametrize("mhd", [True, False])
def test_text_state_matches_oracle_and_percent_g(tmp_path, nx, ny, mhd):
    g, og = grids(nx, ny, dx=0.3, dy=0.7, x0=-1.25, y0=0.35)
    s, u, b = state(ny, nx, nx * 31 + ny, mhd)
    a, o = tmp_path / "lib.txt", tmp_path / "oracle.txt"
    api.write_state_text(g, s, u, b, a)
    O.write_state_text(og, s, u, b, o)
    assert a.read_bytes() == o.read_bytes()
    lines = o.read_text().splitlines()
    assert len(lines) == nx * ny
    bz = b if mhd else np.zeros((ny, nx, 3))
    for k in (0, 1, nx * ny // 2, nx * ny - 1):
        n, m = k % nx, k // nx
        x, y = -1.25 + n * 0.3, 0.35 + m * 0.7
        vals = [x, y, *s[m, n], *u[m, n], *bz[m, n]]
        assert lines[k] == " ".join(fmt(v) for v in vals)


@pytest.mark.parametrize("mhd", [True, False])
def test_text_round_trip_and_reader_parity(tmp_path, mhd):
    nx, ny = 23, 19
    g, og = grids(nx, ny)
    s, u, b = state(ny, nx, 5, mhd)
    path = tmp_path / "restart.txt"
    api.write_state_text(g, s, u, b, path)
    s1, u1, b1 = api.read_state_text(g, path, mhd=mhd)
    s2, u2, b2 = O.read_state_text(og, path, mhd=mhd)
    r6 = np.vectorize(lambda v: float("%g" % v))
    np.testing.assert_array_equal(s1, r6(s))
    np.testing.assert_array_equal(u1, r6(u))
    np.testing.assert_array_equal(s1, s2)
    np.testing.assert_array_equal(u1, u2)
    if mhd:
        np.testing.assert_array_equal(b1, r6(b))
        np.testing.assert_array_equal(b1, b2)


def test_text_reader_rejects_short_and_bad_files(tmp_path):
    g = api.make_grid(4, 3)
    p = tmp_path / "short.txt"
    p.write_text("0 0 1 1 0 0 0 0 0\n" * 11)
    with pytest.raises(api.GD2Error):
        api.read_state_text(g, p)
    p.write_text("0 0 1 1 0 0 0 0 x\n" * 12)
    with pytest.raises(api.GD2Error):
        api.read_state_text(g, p)
    with pytest.raises(api.GD2Error):
        api.read_state_text(g, tmp_path / "missing.txt")
    # the reference's reader is whitespace-, not line-, separated
    p.write_text("\n".join(["+0 0", "1 1 0", "0 0 0 0"] * 12) + "\n")
    s, u, b = api.read_state_text(g, p)
    assert np.all(s == (1.0, 1.0)) and np.all(u == 0.0) and np.all(b == 0.0)


def test_b3_binary_matches_oracle_bitwise(tmp_path):
    nx, ny = 29, 13
    g, og = grids(nx, ny, dx=0.125, dy=0.25)
    s, u, b = state(ny, nx, 9)
    a, o = tmp_path / "lib.dat", tmp_path / "oracle.dat"
    api.write_state_b3(g, s, u, b, a)
    O.write_state_b3(og, s, u, b, o)
    assert a.read_bytes() == o.read_bytes()
    rec = np.fromfile(a, dtype=np.float64).reshape(ny, nx, 7)
    x = g.x0 + np.arange(nx) * g.dx
    y = g.y0 + np.arange(ny) * g.dy
    np.testing.assert_array_equal(rec[..., 0], np.broadcast_to(x, (ny, nx)))
    np.testing.assert_array_equal(rec[..., 1], np.broadcast_to(y[:, None], (ny, nx)))
    np.testing.assert_array_equal(rec[..., 2:4], s)
    np.testing.assert_array_equal(rec[..., 4:6], u)
    np.testing.assert_array_equal(rec[..., 6], b[..., 2])


@pytest.mark.parametrize("nx,ny,nmin", [(10, 7, 3), (9, 9, 1), (31, 17, 4), (5, 2, 2), (1, 6, 3)])
@pytest.mark.parametrize("mhd", [True, False])
@pytest.mark.parametrize("t", [None, 12.75])
def test_tecplot_matches_oracle_and_fepoint_definition(tmp_path, nx, ny, nmin, mhd, t):
    g, og = grids(nx, ny, gamma=5.0 / 3.0)
    s, u, b = state(ny, nx, nx + 7 * ny + nmin, mhd)
    a, o = tmp_path / "lib.dat", tmp_path / "oracle.dat"
    api.write_tecplot(g, s, u, b, a, nmin=nmin, solution_time=t)
    O.write_tecplot(og, s, u, b, o, nmin=nmin, solution_time=t)
    assert a.read_bytes() == o.read_bytes()
    lines = o.read_text().splitlines()
    nn, mm = -(-nx // nmin), -(-ny // nmin)
    head = lines[0]
    assert head.startswith('TITLE = "HP"  VARIABLES = "X", "Y", "Ro", "P", "Vx", "Vy", ')
    assert (', "Bx", "By", "Bz", ' in head) == mhd
    assert f'ZONE T = "HP", N = {nn * mm} , E = {(nn - 1) * (mm - 1)}, F = FEPOINT, ET = quadrilateral' in head
    assert head.endswith(f", SOLUTIONTIME = {t:g}" if t is not None else "ET = quadrilateral")
    nodes = lines[1:1 + nn * mm]
    elems = lines[1 + nn * mm:]
    assert len(elems) == (nn - 1) * (mm - 1)
    # nodes in row-major order of the subsampled cells, each on its own
    # coordinates; Max = |v| / c, T = p / rho (0, 0 where rho <= 0)
    x_max, y_max = g.x0 + (nx - 0.5) * g.dx, g.y0 + (ny - 0.5) * g.dy
    q = 0
    for m in range(0, ny, nmin):
        for n in range(0, nx, nmin):
            r, p = s[m, n]
            vx, vy = u[m, n]
            with np.errstate(all="ignore"):
                mx, tt = (np.sqrt((vx * vx + vy * vy) / (5.0 / 3.0 * p / r)), p / r) if r > 0 else (0.0, 0.0)
            x = g.x0 + n * (x_max - g.x0) / (nx - 1) if nx > 1 else g.x0
            y = g.y0 + m * (y_max - g.y0) / (ny - 1) if ny > 1 else g.y0
            vals = [x, y, r, p, vx, vy] + (list(b[m, n]) if mhd else []) + [mx, tt]
            assert nodes[q] == " ".join(fmt(v) for v in vals), (m, n)
            q += 1
    # every element is the counter-clockwise quad of 4 neighbouring nodes
    for e, line in enumerate(elems):
        i, j = e % (nn - 1), e // (nn - 1)
        n1 = j * nn + i + 1
        assert line == f"{n1} {n1 + 1} {n1 + nn + 1} {n1 + nn}"


def test_tecplot_node_coordinates_span_the_domain(tmp_path):
    """print_file_mini2's (N - 1) spacing puts the first cell at x_min and the
    last at the domain's edge x_max (the production grid: 9 x 5)."""
    kw = PB.wind_grid_args(40, 24, x_max=9.0, y_max=5.0)
    g = api.make_grid(40, 24, **kw)
    s, u, b = PB.stellar_wind_prim(g)
    p = tmp_path / "wind.dat"
    api.write_tecplot(g, s, u, b, p, nmin=1)
    rows = np.loadtxt(p, skiprows=1, max_rows=40 * 24)
    assert rows[0, 0] == pytest.approx(9.0 / 80) and rows[39, 0] == pytest.approx(9.0)
    assert rows[-1, 1] == pytest.approx(5.0)


def test_io_arguments_rejected(tmp_path):
    g = api.make_grid(4, 3)
    s, u, b = state(3, 4, 1)
    with pytest.raises(api.GD2Error):
        api.write_tecplot(g, s, u, b, tmp_path / "x.dat", nmin=0)
    with pytest.raises(api.GD2Error):
        api.write_state_b3(g, s, u, None, tmp_path / "x.dat")
    with pytest.raises(api.GD2Error):
        api.write_state_text(g, s[:, :2], u, b, tmp_path / "x.txt")
    with pytest.raises(api.GD2Error):
        api.write_state_text(g, s, u, b, tmp_path / "no" / "such" / "dir.txt")


def _smooth(ny, nx, seed):
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.5, 2.0, 6)
    X, Y = np.meshgrid(np.arange(nx) * 0.1, np.arange(ny) * 0.1)
    s = np.ascontiguousarray(np.stack([1 + 0.3 * np.sin(a[0] * X) * np.cos(a[1] * Y), 0.8 + 0.2 * np.cos(a[2] * X + Y)], -1))
    u = np.ascontiguousarray(np.stack([0.4 * np.sin(a[3] * Y), 0.3 * np.cos(a[4] * X)], -1))
    b = np.ascontiguousarray(np.stack([0.2 * np.cos(Y), 0.1 * np.sin(X), 0.5 * np.sin(a[5] * X + Y) + 0.6], -1))
    return s, u, b


@pytest.mark.parametrize("nx,ny,nmin", [(23, 17, 3), (12, 9, 1), (7, 4, 2)])
def test_derived_tecplot_matches_oracle_and_finite_differences(tmp_path, nx, ny, nmin):
    g, og = grids(nx, ny, dx=0.1, dy=0.1, gamma=5.0 / 3.0)
    s, u, b = _smooth(ny, nx, nx + ny)
    a, o = tmp_path / "lib.dat", tmp_path / "oracle.dat"
    api.write_tecplot_derived(g, s, u, b, a, nmin=nmin)
    O.write_tecplot_derived(og, s, u, b, o, nmin=nmin)
    assert a.read_bytes() == o.read_bytes()
    lines = o.read_text().splitlines()
    nn, mm = -(-nx // nmin), -(-ny // nmin)
    assert lines[0].endswith(f'"zav",  ZONE T = "HP", N = {nn * mm} , E = {(nn - 1) * (mm - 1)}, '
                             'F = FEPOINT, ET = quadrilateral')
    rows = np.array([[float(t) for t in l.split()] for l in lines[1:1 + nn * mm]])
    bz, vx, vy = b[..., 2], u[..., 0], u[..., 1]
    y = g.y0 + np.arange(ny) * g.dy
    q = 0
    for m in range(0, ny, nmin):
        for n in range(0, nx, nmin):
            Fx = Fy = zav = 0.0
            if m < ny - 1 and n < nx - 1:
                dz = (bz[m, n + 1] - bz[m, n - 1]) / 0.2 if n > 0 else (bz[m, n + 1] - bz[m, n]) / 0.1
                dr = (bz[m + 1, n] - bz[m - 1, n]) / 0.2 if m > 0 else (bz[m + 1, n] - bz[m, n]) / 0.1
                Fx = bz[m, n] * dz / (4 * np.pi)
                Fy = (-bz[m, n] ** 2 / y[m] - bz[m, n] * dr) / (4 * np.pi)
                if n > 0 and m > 0:
                    zav = (vy[m, n + 1] - vy[m, n - 1]) / 0.2 - (vx[m + 1, n] - vx[m - 1, n]) / 0.2
            PP = s[m, n, 1] + bz[m, n] ** 2 / (8 * np.pi)
            assert np.allclose(rows[q, 11:], [PP, Fx, Fy, zav], rtol=1e-5, atol=1e-6), (m, n)
            q += 1


def test_line_cuts_match_oracle_and_finite_differences(tmp_path):
    nx, ny = 19, 14
    g, og = grids(nx, ny, dx=0.1, dy=0.1, gamma=5.0 / 3.0)
    s, u, b = _smooth(ny, nx, 3)
    ay, ax, oy, ox = (tmp_path / n for n in ("ly", "lx", "oy", "ox"))
    api.write_line_cuts(g, s, u, b, ay, ax)
    O.write_line_cuts(og, s, u, b, oy, ox)
    assert ay.read_bytes() == oy.read_bytes() and ax.read_bytes() == ox.read_bytes()
    ry = np.array([[float(t) for t in l.split()] for l in oy.read_text().splitlines()[1:]])
    rx = np.array([[float(t) for t in l.split()] for l in ox.read_text().splitlines()[1:]])
    assert len(ry) == nx - 6 and len(rx) == ny - 6
    pp = s[..., 1] + b[..., 2] ** 2 / (8 * np.pi)
    for i, n in enumerate(range(5, nx - 1)):
        assert np.isclose(ry[i, 0], g.x0 + n * g.dx, rtol=1e-5)
        assert np.isclose(ry[i, 10], (pp[0, n + 1] - pp[0, n]) / 0.1, rtol=1e-5, atol=1e-6)
    for i, m in enumerate(range(5, ny - 1)):
        assert np.isclose(rx[i, 0], g.y0 + m * g.dy, rtol=1e-5)
        assert np.isclose(rx[i, 10], (pp[m + 1, 0] - pp[m, 0]) / 0.1, rtol=1e-5, atol=1e-6)
    # either file alone
    api.write_line_cuts(g, s, u, b, path_x0=tmp_path / "only_x")
    assert (tmp_path / "only_x").read_bytes() == ox.read_bytes()
