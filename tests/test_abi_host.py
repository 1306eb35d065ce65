"""The C ABI library without a GPU: it loads, exports every entry point that
include/spilady_hip.h declares, and its host-only planning (link cells,
groups, sweep levels) agrees with the reference's geometry and ordering."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from oracle import orcpy
from spilady_amd import abi, configs, engine

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_library_exports_every_declared_symbol():
    lib = engine.load()
    names = engine.exported_symbols()
    assert "sld_create" in names and "sld_step" in names and "sld_plan" in names
    for nm in names:
        assert hasattr(lib, nm), nm
    out = subprocess.run(["nm", "-D", "--defined-only", engine.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert set(names) <= exported
    assert lib.sld_version() == abi.ABI_VERSION == 2


def test_ctypes_structs_match_header_sizes():
    """Field layout of the plain-data structs: sizes fixed by the header."""
    assert C.sizeof(abi.SldParams) == 4 * 2 + 8 * 6 + 8 * 11 + 8 + 4 * 6 + 4 * 8 + 4 * 4 + 8
    assert C.sizeof(abi.SldAtoms) == 15 * 8
    assert C.sizeof(abi.SldObs) == 8 + 8 + 8 * 5 + 8 * 4 + 8 * 8 + 8 + 8 * 6 + 8


STRUCTS = (("sld_params", abi.SldParams), ("sld_atoms", abi.SldAtoms), ("sld_obs", abi.SldObs),
           ("sld_cells", abi.SldCells), ("sld_shapes", abi.SldShapes), ("sld_barostat", abi.SldBarostat))


def test_ctypes_structs_match_compiled_header(tmp_path):
    """Every field's offset and every struct's size as gcc lays out
    include/spilady_hip.h, against the ctypes mirror in spilady_amd/abi.py."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "spilady_hip.h"', "int main(void) {"]
    for cname, py in STRUCTS:
        lines.append(f'    printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'    printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(root, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for ln in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines():
        cname, f, val = ln.split()
        got[(cname, f)] = int(val)
    for cname, py in STRUCTS:
        assert got[(cname, "size")] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)


def test_plan_matches_reference_groups():
    g = np.load(os.path.join(GOLD, "groups.npz"), allow_pickle=False)
    for nu in g["nus"]:
        L = int(nu) * 2.8665
        pl = engine.plan([L, 0, L, 0, 0, L], 4.11)
        assert pl["ncell3"] == tuple(g[f"nu{nu}_ncell3"])
        np.testing.assert_array_equal(pl["group_sizes"], g[f"nu{nu}_sizes"])
        np.testing.assert_array_equal(pl["group_cells"], g[f"nu{nu}_members"])


@pytest.mark.parametrize("nu", [8, 10, 16])
def test_sweep_levels_reproduce_group_order(nu):
    """A level schedule is exact if (a) cells of one level never touch each
    other through the stencil of the sweep and (b) every pair of touching
    cells is ordered as the groups order them (forward: lower group first;
    backward: higher group first)."""
    L = nu * 2.8665
    pl = engine.plan([L, 0, L, 0, 0, L], 4.11)
    nx, ny, nz = pl["ncell3"]
    C_ = nx * ny * nz
    group_of = np.empty(C_, dtype=int)
    off = np.concatenate([[0], np.cumsum(pl["group_sizes"])])
    for gi in range(len(pl["group_sizes"])):
        group_of[pl["group_cells"][off[gi]:off[gi + 1]]] = gi
    cells = np.arange(C_)
    x, y, z = cells % nx, (cells // nx) % ny, cells // (nx * ny)
    offs26 = [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)
              if (dx, dy, dz) != (0, 0, 0)]
    offs6 = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    for sch, offs in enumerate((offs26, offs26, offs6, offs6)):
        lv = pl["levels"][sch]
        assert lv.max() + 1 == pl["nlevels"][sch]
        backward = sch % 2 == 1
        for dx, dy, dz in offs:
            nb = ((x + dx) % nx) + ((y + dy) % ny) * nx + ((z + dz) % nz) * nx * ny
            assert np.all(lv != lv[nb]), (sch, dx, dy, dz)
            earlier = group_of[nb] > group_of if backward else group_of[nb] < group_of
            assert np.all(lv[nb][earlier] < lv[earlier]), (sch, dx, dy, dz)


def test_plan_levels_even_box_merge_face_sweep():
    """With an even number of link cells per side the 8 groups form the 2x2x2
    parity pattern: the face-only electron sweep needs 4 levels, the spin
    sweep 8."""
    L = 10 * 2.8665
    pl = engine.plan([L, 0, L, 0, 0, L], 4.11)
    assert pl["ncell3"] == (6, 6, 6)
    assert list(pl["nlevels"]) == [8, 8, 4, 4]


def _batch_args(params_list):
    from spilady_amd import abi as A
    n = len(params_list)
    ats = [configs.make_bcc_fe(6, 300.0, 1) for _ in range(n)]
    bufs = [A.AtomBuffers.from_state(a) for a in ats]
    pv = (A.SldParams * n)(*params_list)
    av = (A.SldAtoms * n)(*[b.view() for b in bufs])
    return pv, av, bufs


def test_batch_rejects_mismatched_or_eltemp_replicas():
    """Argument checks of sld_create_batch that run before any device call:
    replicas must share dt and the switches, and eltemp and changestep need
    one replica."""
    lib = engine.load()
    at = configs.make_bcc_fe(6, 300.0, 1)
    v = configs.fe_variables(300.0, 1, thermostats=True)
    p0 = abi.params_from_variables(v, at.box, at.natom)
    p1 = abi.params_from_variables(v, at.box, at.natom, replica=1)
    p1.step = 2e-15
    pv, av, bufs = _batch_args([p0, p1])
    h = C.c_void_p()
    rc = lib.sld_create_batch(2, pv, av, 0, C.byref(h))
    assert rc == 1 and b"must equal replica 0" in lib.sld_last_error(h)
    lib.sld_destroy(h)
    ve = configs.fe_variables(300.0, 1, thermostats=True, eltemp=True)
    q0 = abi.params_from_variables(ve, at.box, at.natom)
    q1 = abi.params_from_variables(ve, at.box, at.natom, replica=1)
    pv, av, bufs = _batch_args([q0, q1])
    h = C.c_void_p()
    rc = lib.sld_create_batch(2, pv, av, 0, C.byref(h))
    assert rc == 1 and b"single-system" in lib.sld_last_error(h)
    lib.sld_destroy(h)
    # changestep adapts one time step per context
    c0 = abi.params_from_variables(dict(v, changestep=1), at.box, at.natom)
    c1 = abi.params_from_variables(dict(v, changestep=1), at.box, at.natom, replica=1)
    pv, av, bufs = _batch_args([c0, c1])
    h = C.c_void_p()
    rc = lib.sld_create_batch(2, pv, av, 0, C.byref(h))
    assert rc == 1 and b"changestep" in lib.sld_last_error(h)
    lib.sld_destroy(h)


def test_entry_points_refuse_a_null_context():
    """Every entry point that takes a context returns SLD_ERR_ARG (or 0 for
    the time accessors) for a null one, without touching a device."""
    lib = engine.load()
    null = C.c_void_p(None)
    buf = (C.c_double * 8)()
    assert lib.sld_links(null) == 1
    assert lib.sld_calculate_rho(null) == 1
    assert lib.sld_calculate_force_energy(null) == 1
    assert lib.sld_core_ds(null, C.c_double(1e-15)) == 1
    assert lib.sld_core_dr(null, C.c_double(1e-15)) == 1
    assert lib.sld_check(null, 0, None) == 1
    assert lib.sld_step(null, 0, 1, None, None) == 1
    assert lib.sld_get_box(null, buf) == 1
    assert lib.sld_select_replica(null, 0) == 1
    assert lib.sld_set_cells_te(null, buf) == 1
    assert lib.sld_destroy(null) == 0
    assert lib.sld_nreplicas(null) == 0
    lib.sld_total_time.restype = C.c_double
    assert lib.sld_total_time(null) == 0.0
