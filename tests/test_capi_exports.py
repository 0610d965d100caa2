"""libgd2.so loads and exports every entry point include/gd2.h declares (no GPU calls)."""
import ctypes
import os
import re

import pytest

from golikov_2d_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "gd2.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gd2_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_expected_entry_points():
    syms = declared_symbols()
    for name in ["gd2_ker_dekard", "gd2_kernel_tvd_old", "gd2_add2", "gd2_kernel_tvd", "gd2_funk_time",
                 "gd2_stage", "gd2_ctx_advance"]:
        assert name in syms


@pytest.mark.skipif(not os.path.exists(api.SO_PATH), reason="libgd2.so not built")
def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(api.SO_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    assert set(api.EXPORTED) <= set(declared_symbols())


@pytest.mark.skipif(not os.path.exists(api.SO_PATH), reason="libgd2.so not built")
def test_pure_host_entry_points():
    L = api.lib()
    assert L.gd2_nfields(api.EULER) == 4 and L.gd2_nfields(api.MHD) == 7
    assert L.gd2_nfields(9) < 0
    assert b"gfx950" in L.gd2_version()
    # argument validation happens before any device work
    g = api.make_grid(8, 8)
    sc = api.make_scheme()
    assert L.gd2_stage(ctypes.byref(g), ctypes.byref(sc), None, None, None, 0.0, 1.0, 0.0, None, None, None) == -1
    bad = api.make_grid(8, 8, bc=(api.BC_PERIODIC, api.BC_TRANSMISSIVE, 0, 0))
    x = ctypes.c_void_p(16)
    assert L.gd2_stage(ctypes.byref(bad), ctypes.byref(sc), x, None, ctypes.c_void_p(32), 0.0, 1.0, 0.0, None,
                       None, None) == -1
    assert L.gd2_ker_dekard(ctypes.byref(g), x, x, ctypes.c_void_p(32), ctypes.c_void_p(48), x, x, 7, None) == -2
    # production options: MHD reference-layout steps only, each side kind on
    # its own side, axisym with y0 > 0, r_frozen >= 0
    from golikov_2d_amd import problems as PB
    prod = api.make_grid(8, 8, **PB.wind_grid_args(8, 8))
    y = ctypes.c_void_p(64)
    assert L.gd2_stage(ctypes.byref(prod), ctypes.byref(sc), x, None, y, 0.0, 1.0, 0.0, None, None, None) == -1
    assert L.gd2_ker_dekard(ctypes.byref(prod), x, x, y, y, x, x, 1, None) == -1
    for bad in (dict(bc=(api.BC_OUTFLOW, api.BC_SYMMETRY, api.BC_AXIS, api.BC_INFLOW)),
                dict(bc=(api.BC_SYMMETRY, api.BC_OUTFLOW, api.BC_INFLOW, api.BC_AXIS)),
                dict(y0=0.0), dict(r_frozen=-1.0)):
        gb = api.make_grid(8, 8, **dict(PB.wind_grid_args(8, 8), **bad))
        for fn in (L.gd2_add2, L.gd2_kernel_tvd):
            assert fn(ctypes.byref(gb), x, x, x, y, y, y, x, x, 0, 2, None) == -1


@pytest.mark.skipif(not os.path.exists(api.SO_PATH), reason="libgd2.so not built")
def test_step_path_policy():
    """AUTO takes the one-pass SSP-RK2 kernel from 2^21 cells up; the forced
    modes override it; non-Euler schemes, forward Euler and grids under 2 x 2
    never take it; a bad mode is rejected (host logic only)."""
    sc = api.make_scheme()
    big, small = api.make_grid(2048, 1024), api.make_grid(1024, 1024)
    prev = api.set_step_path(api.STEP_AUTO)
    try:
        assert api.one_pass_preferred(big, sc) and not api.one_pass_preferred(small, sc)
        assert api.set_step_path(api.STEP_ONE_PASS) == api.STEP_AUTO
        assert api.one_pass_preferred(small, sc)
        assert not api.one_pass_preferred(api.make_grid(1, 64), sc)
        assert not api.one_pass_preferred(big, api.make_scheme(integrator=api.FORWARD_EULER))
        assert not api.one_pass_preferred(big, api.make_scheme(api.MHD, api.FLUX_HLLD))
        api.set_step_path(api.STEP_TWO_PASS)
        assert not api.one_pass_preferred(big, sc)
        assert api.lib().gd2_set_step_path(7) == -1
    finally:
        api.set_step_path(prev)


@pytest.mark.skipif(not os.path.exists(api.SO_PATH), reason="libgd2.so not built")
def test_header_is_plain_c_and_links(tmp_path):
    """A C99 program that includes gd2.h compiles with -pedantic, links
    against libgd2.so and runs its host-only calls (no device work): argument
    validation of a drop-in step and the text-state writer."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    src = tmp_path / "use_gd2.c"
    src.write_text(r'''
#include <stdio.h>
#include <string.h>
#include "gd2.h"
int main(int argc, char** argv) {
    gd2_grid g;
    memset(&g, 0, sizeof g);
    g.nx = 2; g.ny = 1; g.dx = 0.5; g.dy = 1.0; g.x0 = 0.25; g.y0 = 0.5; g.gamma = 1.4; g.krit = 0.25;
    gd2_double2 s[2] = {{1.0, 2.5}, {0.125, 0.1}}, u[2] = {{0.0, 0.0}, {-1.0, 3.0}};
    /* a bad method is rejected before any device work */
    if (gd2_ker_dekard(&g, s, u, s, u, NULL, NULL, 7, NULL) != GD2_ERR_ARG) return 2;
    if (gd2_write_state_text(&g, s, u, NULL, argv[1]) != GD2_OK) return 3;
    printf("%s\n", gd2_version());
    return argc == 2 ? 0 : 4;
}
''')
    exe = tmp_path / "use_gd2"
    inc = os.path.join(ROOT, "include")
    libdir = os.path.dirname(api.SO_PATH)
    subprocess.run([cc, "-std=c99", "-pedantic", "-Wall", "-Werror", "-I", inc, str(src), "-o", str(exe),
                    f"-L{libdir}", "-lgd2", f"-Wl,-rpath,{libdir}"], check=True)
    out = tmp_path / "state.txt"
    r = subprocess.run([str(exe), str(out)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, (r.returncode, r.stderr)
    assert "gfx950" in r.stdout
    assert out.read_text() == "0.25 0.5 1 2.5 0 0 0 0 0\n0.75 0.5 0.125 0.1 -1 3 0 0 0\n"


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(api.SO_PATH), "gd2_production")),
                    reason="gd2_production not built")
def test_production_driver_links_and_rejects_bad_arguments():
    """The production-run driver resolves libgd2 and HIP at load and refuses an
    unknown option or a degenerate grid with its usage line and status 2,
    before any device call."""
    import subprocess
    exe = os.path.join(os.path.dirname(api.SO_PATH), "gd2_production")
    for args in (["--bogus", "1"], ["--nx", "1"], ["--steps1"]):
        r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=60)
        assert r.returncode == 2 and "usage:" in r.stderr, (args, r.returncode, r.stderr)


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(api.SO_PATH), "gd2_production")),
                    reason="gd2_production not built")
def test_production_driver_missing_restart_file(tmp_path):
    """A restart file that cannot be read stops the driver with status 1 and
    the failing call named, before any device call."""
    import subprocess
    exe = os.path.join(os.path.dirname(api.SO_PATH), "gd2_production")
    r = subprocess.run([exe, "--nx", "8", "--ny", "8", "--restart", str(tmp_path / "missing.txt"),
                        "--out", str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "gd2_read_state_text" in r.stderr, (r.returncode, r.stderr)


def test_ctypes_structs_match_the_headers(tmp_path):
    """The ctypes mirrors of gd2_grid / gd2_scheme (golikov_2d_amd.api) and of
    o_grid / o_scheme (the oracle's) have the C headers' sizes and field
    offsets, so no call reads past or beside what Python passes."""
    import shutil
    import subprocess
    import ctypes as C
    from oracle import oracle as O
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = [("gd2_grid", api.Grid), ("gd2_scheme", api.Scheme), ("o_grid", O.Grid), ("o_scheme", O.Scheme)]
    lines = []
    for cname, py in structs:
        lines.append(f'    printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'    printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    src = tmp_path / "layout.c"
    src.write_text("#include <stdio.h>\n#include <stddef.h>\n#include \"gd2.h\"\n#include \"gd2_oracle.h\"\n"
                   "int main(void) {\n" + "\n".join(lines) + "\n    return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-I",
                    os.path.join(ROOT, "oracle"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines():
        *key, val = line.split()
        got[tuple(key)] = int(val)
    for cname, py in structs:
        assert got[(cname, "size")] == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)
