"""The plain-C oracle against the compiled SPILADY reference.

Fixtures in tests/golden/ were written by tests/golden/make_golden.py from the
reference built out of its own sources (oracle/Makefile.ref). The oracle runs
the reference's loop order on one thread with the reference's MWC noise
stream, so agreement is bit for bit. When oracle/_ref is present the same
comparison also runs live.
"""
import os
import tempfile

import numpy as np
import pytest

from oracle import orcpy, refpy
from spilady_amd import abi, configs
from spilady_amd import constants as K
from spilady_amd.io import AtomState

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def _variables(g):
    keys = [str(k) for k in g["variables_keys"]]
    vals = [str(v) for v in g["variables_values"]]
    out = {}
    for k, v in zip(keys, vals):
        try:
            out[k] = int(v)
        except ValueError:
            try:
                out[k] = float(v)
            except ValueError:
                out[k] = v
    return out


def _oracle_from_fixture(g):
    v = _variables(g)
    n = g["in_r"].shape[0]
    st = {k: g[f"in_{k}"] for k in ("r", "p", "f", "s", "stress", "rho", "ke", "pe", "me", "me0")}
    # the reference build has every control.h switch on (refpy.REF_SWITCHES)
    p = abi.params_from_variables(dict(v, **refpy.REF_SWITCHES), g["box"], n)
    return v, p, orcpy.Oracle(p, st, compute_initial_forces=False, rng=orcpy.RNG_MWC)


def test_tables_match_reference():
    g = _load("tables.npz")
    at = configs.make_bcc_fe(6, 300.0, 1)
    v = configs.fe_variables(300.0, 1, thermostats=False)
    p = abi.params_from_variables(v, at.box, at.natom)
    o = orcpy.Oracle(p, at, compute_initial_forces=True)
    for w in range(16):
        xs = g["x_rho"] if w in (0, 1) or w >= 8 else g["x_r"]
        got = np.array([o.table(w, float(x)) for x in xs])
        np.testing.assert_array_equal(got, g[f"t{w}"], err_msg=f"table {w}")
        raw = np.array([o.table_raw(w, int(i)) for i in g["raw_index"]])
        np.testing.assert_array_equal(raw, g["raw"][w], err_msg=f"raw table {w}")
    # density of the perfect lattice through the reference's own evaluation
    np.testing.assert_array_equal(o.download()["rho"], g["perfect_rho"])
    o.close()


def test_groups_match_reference():
    g = _load("groups.npz")
    for nu in g["nus"]:
        at = configs.make_bcc_fe(int(nu), 300.0, 1)
        v = configs.fe_variables(300.0, 1, thermostats=False)
        p = abi.params_from_variables(v, at.box, at.natom)
        at.rho[:] = 0.48
        o = orcpy.Oracle(p, at)
        assert o.ncell3 == tuple(g[f"nu{nu}_ncell3"])
        sizes, members = o.groups()
        np.testing.assert_array_equal(sizes, g[f"nu{nu}_sizes"])
        np.testing.assert_array_equal(members, g[f"nu{nu}_members"])
        o.close()


@pytest.mark.parametrize("name", ["step_nve.npz", "step_lang.npz"])
def test_entry_points_match_reference(name):
    """Each reference entry point of one core() step, bit for bit."""
    g = _load(name)
    v, p, o = _oracle_from_fixture(g)
    dt = p.step
    args = {"core_dTe": (dt / 2,), "core_dp": (dt / 2,), "core_ds": (dt / 2,), "core_dr": (dt,)}
    for i, nm in enumerate(str(x) for x in g["seq"]):
        getattr(o, nm)(*args.get(nm, ()))
        a, c = o.download(), o.cells()
        keys = [k[len(f"s{i}_"):] for k in g.files if k.startswith(f"s{i}_")]
        for k in keys:
            ref = g[f"s{i}_{k}"]
            if k == "cell_counts":
                got = o.cell_lists(32)[0]
            elif k == "cell_lists":
                got = o.cell_lists(32)[1]
            elif k.startswith("cell_"):
                got = c[k[5:]]
            else:
                got = a[k]
            np.testing.assert_array_equal(got, ref, err_msg=f"{name} step {i} ({nm}) field {k}")
    o.close()


@pytest.mark.parametrize("name", ["step_nve.npz", "step_lang.npz"])
def test_trajectory_matches_reference(name):
    """nsteps of core() + check() (renormalizeEnergy on): the enr-/tmp-/spn-/
    str-/prs- columns as printed (%.15e) and the final state, bit for bit."""
    g = _load(name)
    v, p, o = _oracle_from_fixture(g)
    n = int(g["nsteps"])
    obs = o.step(0, n)
    enr = np.array([[x["ave_pe"], x["ave_ke"], x["ave_me"], x["ave_me0"], x["ave_ee"]] for x in obs])
    tmp = np.array([[x["tl"], x["ts_r"], x["ts_l"], x["te"]] for x in obs])
    spn = np.array([list(x["ave_s"]) + [x["ave_s0"]] + list(x["ave_m"]) + [x["ave_m0"]] for x in obs])
    as_printed = lambda a: np.array([[float("%.15e" % y) for y in row] for row in a])  # noqa: E731
    np.testing.assert_array_equal(as_printed(enr), g["enr"][1:, 2:])
    np.testing.assert_array_equal(as_printed(tmp), g["tmp"][1:, 2:])
    np.testing.assert_array_equal(as_printed(spn), g["spn"][1:, 2:])
    # str- and prs- (check_stress_CPU.cpp:68-81, check_pressure_CPU.cpp:49-55)
    box = [float(x) for x in g["box"]]
    dens = g["in_r"].shape[0] / (box[0] * box[2] * box[5])
    st = np.array([box + [dens, sum(x["ave_stress"][:3]) / 3e0] + list(x["ave_stress"]) for x in obs])
    pr = np.array([box + [dens, x["pressure"]] for x in obs])
    np.testing.assert_array_equal(as_printed(st), g["str"][1:, 2:])
    np.testing.assert_array_equal(as_printed(pr), g["prs"][1:, 2:])
    a = o.download()
    for k in ("r", "p", "f", "s", "rho", "pe", "ke", "me", "me0"):
        np.testing.assert_array_equal(a[k], g[f"final_{k}"], err_msg=k)
    np.testing.assert_array_equal(o.cells()["te"], g["final_cell_te"])
    counts, lists = o.cell_lists(32)
    np.testing.assert_array_equal(counts, g["final_cell_counts"])
    np.testing.assert_array_equal(lists, g["final_cell_lists"])
    o.close()


@pytest.mark.skipif(not refpy.available(), reason="compiled reference not built (oracle/Makefile.ref)")
def test_live_reference_agrees_hot_lattice():
    """Live run of the compiled reference against the oracle on a hotter,
    displaced lattice (movers between link cells), thermostats on."""
    at = configs.make_bcc_fe(8, 900.0, 5, displace_sigma=0.05, tilt_max_deg=30.0)
    v = configs.fe_variables(900.0, 5, thermostats=True, eltemp=True, localcolmot=True,
                             renormalize=False)
    st = refpy.consistent_start(v, at)
    d = tempfile.mkdtemp()
    ref = refpy.Reference()
    ref.init_memory(d, v, st)
    ref.run(0, 8, False)
    a = ref.atoms()
    counts_r, lists_r = ref.cell_lists(32)
    ref.free()
    p = abi.params_from_variables(dict(v, **refpy.REF_SWITCHES), st.box, st.natom)
    p.renormalize_energy = 0
    o = orcpy.Oracle(p, st, rng=orcpy.RNG_MWC)
    for s in range(8):
        o.core(s)
    b = o.download()
    for k in ("r", "p", "f", "s", "rho", "pe", "me", "me0"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    counts_o, lists_o = o.cell_lists(32)
    np.testing.assert_array_equal(counts_r, counts_o)
    np.testing.assert_array_equal(lists_r, lists_o)
    o.close()


@pytest.mark.skipif(not refpy.available(), reason="compiled reference not built (oracle/Makefile.ref)")
def test_live_reference_changestep():
    """changestep (scale_step_CPU.cpp:33-97): 16 steps of core() + scale()
    with the reference's limits (0.01 A, 2 pi / 10 rad), where the time step
    shrinks by 0.8 three times and grows by 1.05 otherwise; the dt sequence
    and the final state bit for bit. check() is left out: the reference build
    renormalises the electron energy in it, which this small system does not
    survive."""
    at = configs.make_bcc_fe(8, 600.0, 5, tilt_max_deg=10.0)
    v = configs.fe_variables(600.0, 5, thermostats=True, eltemp=True, localcolmot=True, renormalize=False)
    v["displace_limit"] = 0.01
    v["phi_limit"] = 0.628319
    st = refpy.consistent_start(v, at)
    d = tempfile.mkdtemp()
    ref = refpy.Reference()
    ref.init_memory(d, v, st)
    p = abi.params_from_variables(dict(v, **refpy.REF_SWITCHES), st.box, st.natom)
    p.renormalize_energy = 0
    p.changestep = 1
    o = orcpy.Oracle(p, st, rng=orcpy.RNG_MWC)
    dt_ref, dt_orc = [], []
    for cur in range(16):
        ref.run(cur, 1, False)
        ref.lib.ref_scale(cur)
        dt_ref.append(ref.ref_step())
        o.core(cur)
        o.scale_step()
        dt_orc.append(o.dt)
    a = ref.atoms()
    ref.free()
    b = o.download()
    o.close()
    assert dt_ref == dt_orc
    steps = np.array(dt_ref)
    assert np.sum(steps[1:] < steps[:-1]) >= 2 and np.sum(steps[1:] > steps[:-1]) >= 2
    for k in ("r", "p", "f", "s", "rho", "pe", "me", "me0"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


@pytest.mark.skipif(not refpy.available_ext(),
                    reason="reference with extfield/extforce not built (oracle/Makefile.ref OUT=_ref/ext)")
def test_live_reference_external_field_and_force():
    """extfield and extforce: per-atom fields (read by the reference from
    extfield.in in tesla, converted to eV as -g muB B) and forces (extforce.in,
    eV/A); 6 steps of core() with thermostats and the electron temperature,
    bit for bit. The in-place f += fext of core_dp is part of what is pinned."""
    at = configs.make_bcc_fe(8, 600.0, 7, tilt_max_deg=20.0)
    v = configs.fe_variables(600.0, 7, thermostats=True, eltemp=True, localcolmot=True, renormalize=False)
    st = refpy.consistent_start(v, at)
    n = st.natom
    rng = np.random.default_rng(4)
    b_tesla = rng.normal(0.0, 10.0, size=(n, 3))
    fext = rng.normal(0.0, 0.1, size=(n, 3))
    d = tempfile.mkdtemp()
    for name, arr in (("extfield.in", b_tesla), ("extforce.in", fext)):
        with open(os.path.join(d, name), "w") as fh:
            fh.write(f"{n}\n")
            for i in range(n):
                fh.write("%d %.17g %.17g %.17g\n" % (i, arr[i, 0], arr[i, 1], arr[i, 2]))
    ref = refpy.Reference(refpy.REF_LIB_EXT)
    ref.init_memory(d, v, st)
    for cur in range(6):
        ref.run(cur, 1, False)
    a = ref.atoms()
    ref.free()
    # external_field_CPU.cpp:50: Hext = (-el_g*muB) * B
    hext = (-2.0023193043622e0 * 5.7883817555e-5) * b_tesla
    p = abi.params_from_variables(dict(v, **refpy.REF_SWITCHES), st.box, n)
    p.renormalize_energy = 0
    p.extfield = 1
    p.extforce = 1
    o = orcpy.Oracle(p, st, rng=orcpy.RNG_MWC)
    o.set_external(hext, fext)
    # initial.cpp reads the field before its calculate_temperature()
    o.calculate_temperature()
    for cur in range(6):
        o.core(cur)
    b = o.download()
    o.close()
    for k in ("r", "p", "f", "s", "rho", "pe", "me", "me0", "heff_h"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_barostats_match_reference_fixture():
    """barostat_fe1k.npz (the reference built with -DSTRESS -DPRESSURE):
    12 steps of core(), check() and scale() on a compressed lattice; the
    box and time step after every step, the str-/prs- rows (box at the
    check, density, stresses, pressure) and the final r, p, s bit for bit."""
    g = _load("barostat_fe1k.npz")
    v = _variables(g)
    n = g["in_r"].shape[0]
    st = {k: g[f"in_{k}"] for k in ("r", "p", "f", "s", "stress", "rho", "ke", "pe", "me", "me0")}
    p = abi.params_from_variables(dict(v, **refpy.REF_SWITCHES), g["box"], n)
    p.changestep = 1  # compiled into the shipped control.h
    o = orcpy.Oracle(p, st, compute_initial_forces=False, rng=orcpy.RNG_MWC)
    o.set_barostat(abi.barostat_from_variables(v))
    fmt = lambda x: float("%.15e" % x)  # noqa: E731
    box = o.box()
    for cur in range(12):
        ob = o.step(cur, 1)[0]
        srow, prow = g["str"][cur + 1], g["prs"][cur + 1]
        assert int(srow[0]) == cur and srow[1] == fmt(ob["total_time"])
        sv = list(ob["ave_stress"])
        dens = n / (box[0] * box[2] * box[5])
        assert list(srow[2:]) == [fmt(x) for x in list(box) + [dens, sum(sv[:3]) / 3e0] + sv], cur
        assert list(prow[2:]) == [fmt(x) for x in list(box) + [dens, ob["pressure"]]], cur
        box = o.box()
        np.testing.assert_array_equal(box, g["boxes"][cur])
        assert o.dt == g["dts"][cur]
    b = o.download()
    o.close()
    for k in ("r", "p", "s"):
        np.testing.assert_array_equal(b[k], g[f"final_{k}"], err_msg=k)


@pytest.mark.skipif(not refpy.available_baro(),
                    reason="reference with STRESS/PRESSURE not built (oracle/Makefile.ref OUT=_ref/baro)")
def test_live_reference_barostats():
    """The Berendsen barostats (scale_stress_CPU.cpp:33-86, scale_pressure_CPU.cpp:
    33-61) on a compressed lattice (a = 2.80 A, about 13 GPa), stress every 2
    steps and pressure every 3, with check() after every step as their input
    and the shipped changestep after them (scale.cpp:31-44): 12 steps, the box
    matrix after every step and the final state bit for bit. Thermostats are
    off: the reference build renormalises the electron energy in check()."""
    at = configs.make_bcc_fe(8, 300.0, 5, a=2.80, tilt_max_deg=10.0)
    v = configs.fe_variables(300.0, 5, thermostats=False)
    v.update({"STRESS": 1, "PRESSURE": 1, "interval_of_scale_stress": 2, "interval_of_scale_pressure": 3,
              "baro_damping_time": 2e-14, "stress_xx": 1.0, "stress_yx": 0.5, "stress_yy": -1.0,
              "stress_zx": 0.2, "stress_zy": -0.3, "stress_zz": 2.0, "pressure": 5.0})
    st = refpy.consistent_start(v, at)
    d = tempfile.mkdtemp()
    ref = refpy.Reference(refpy.REF_LIB_BARO)
    ref.init_memory(d, v, st)
    p = abi.params_from_variables(dict(v, **refpy.REF_SWITCHES), st.box, st.natom)
    p.changestep = 1  # compiled into the shipped control.h
    o = orcpy.Oracle(p, st, rng=orcpy.RNG_MWC)
    o.set_barostat(abi.barostat_from_variables(v))
    boxes_r, boxes_o = [], []
    for cur in range(12):
        ref.run(cur, 1, True)
        ref.lib.ref_scale(cur)
        boxes_r.append(ref.box())
        o.step(cur, 1)
        boxes_o.append(o.box())
    a = ref.atoms()
    dt_r = ref.ref_step()
    ref.free()
    b = o.download()
    dt_o = o.dt
    o.close()
    np.testing.assert_array_equal(np.array(boxes_r), np.array(boxes_o))
    assert dt_r == dt_o
    # both barostats moved the box: the diagonal and the tilts changed
    bx = np.array(boxes_o)
    assert np.all(np.abs(bx[-1] - st.box) > 1e-3)
    for k in ("r", "p", "f", "s", "rho", "pe", "me", "me0"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def _quantum_case():
    at = configs.make_bcc_fe(6, 300.0, 5, tilt_max_deg=10.0)
    v = configs.fe_variables(300.0, 5, thermostats=True, eltemp=True, localcolmot=True)
    v.update({"Msteps_quantum": 3, "Nfrequency_quantum": 8})
    return at, v


def _quantum_oracle(v, st, box, n):
    p = abi.params_from_variables(dict(v, **refpy.REF_SWITCHES), box, n)
    p.renormalize_energy = 0
    p.quantumnoise = 1
    return orcpy.Oracle(p, st, rng=orcpy.RNG_MWC)


@pytest.mark.skipif(not refpy.available_qn(),
                    reason="reference with quantumnoise not built (oracle/Makefile.ref OUT=_ref/qn)")
def test_live_reference_quantum_noise():
    """The quantum thermostat (quantum_noise_CPU.cpp, core_dp_CPU.cpp:137-146)
    with M = 3 and Nf = 8 on 432 atoms, eltemp and localcolmot on: the
    16-tap filter H and 8 steps of core() (five redraws of every degree of
    freedom's noise, each appending the MWC stream's next normal to its
    history) bit for bit."""
    at, v = _quantum_case()
    st = refpy.consistent_start(v, at)
    d = tempfile.mkdtemp()
    ref = refpy.Reference(refpy.REF_LIB_QN)
    ref.init_memory(d, v, st)
    o = _quantum_oracle(v, st, st.box, st.natom)
    np.testing.assert_array_equal(ref.quantum_H(), o.quantum_H())
    for cur in range(8):
        ref.run(cur, 1, False)
        o.core(cur)
    a = ref.atoms()
    ref.free()
    b = o.download()
    o.close()
    for k in ("r", "p", "f", "s", "rho", "pe", "me", "me0"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_quantum_noise_matches_reference_fixture():
    """quantum_fe432.npz (the reference built with -Dquantumnoise): the filter
    and the state after 8 steps, bit for bit."""
    g = _load("quantum_fe432.npz")
    v = _variables(g)
    n = g["in_r"].shape[0]
    st = {k: g[f"in_{k}"] for k in ("r", "p", "f", "s", "stress", "rho", "ke", "pe", "me", "me0")}
    o = _quantum_oracle(v, st, g["box"], n)
    np.testing.assert_array_equal(o.quantum_H(), g["H"])
    for cur in range(int(g["nsteps"])):
        o.core(cur)
    b = o.download()
    o.close()
    for k in ("r", "p", "s", "f"):
        np.testing.assert_array_equal(b[k], g[f"final_{k}"], err_msg=k)


@pytest.mark.skipif(not refpy.available_lv(),
                    reason="reference with localvol not built (oracle/Makefile.ref OUT=_ref/lv)")
def test_live_reference_local_volume():
    """localvol (calculate_force_energy_CPU.cpp:80-90,266-293): each atom's
    volume from its sums of 1/r and 1/r^2 within rcut_vol, scaled so the
    volumes fill the box, and the per-atom stress over it. On a lattice with
    displaced atoms (volumes that differ from atom to atom) and rcut_vol
    between the first and second shells, 4 steps of core() with thermostats
    and the electron temperature: stress, local volumes and state bit for bit."""
    at = configs.make_bcc_fe(8, 600.0, 9, tilt_max_deg=15.0, displace_sigma=0.05)
    v = configs.fe_variables(600.0, 9, thermostats=True, eltemp=True, localcolmot=True, renormalize=False)
    v["rcut_vol"] = 2.7
    st = refpy.consistent_start(v, at)
    d = tempfile.mkdtemp()
    ref = refpy.Reference(refpy.REF_LIB_LV)
    ref.init_memory(d, v, st)
    for cur in range(4):
        ref.run(cur, 1, False)
    a = ref.atoms()
    lv_ref = ref.local_volume()
    ref.free()
    p = abi.params_from_variables(dict(v, **refpy.REF_SWITCHES), st.box, st.natom)
    p.renormalize_energy = 0
    p.localvol = 1
    o = orcpy.Oracle(p, st, rng=orcpy.RNG_MWC)
    for cur in range(4):
        o.core(cur)
    b = o.download()
    o.close()
    assert np.ptp(lv_ref) > 0.01 * lv_ref.mean()   # the volumes do differ between atoms
    # the volume factor makes them fill the box (box_volume = d.xx d.yy d.zz)
    np.testing.assert_allclose(lv_ref.sum(), np.prod(np.asarray(st.box)[[0, 2, 5]]), rtol=1e-12)
    np.testing.assert_array_equal(b["local_volume"], lv_ref)
    for k in ("r", "p", "f", "s", "rho", "pe", "me", "me0", "stress"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


@pytest.mark.skipif(not refpy.available_mm(), reason="reference magmom build absent")
def test_live_reference_magmom():
    """magmom (control.h:36): the spin column carries the moments m; the spins
    follow as s = m / (-g) at the start of each spin pass and m = -g s after it
    (core_ds_CPU.cpp:37-81, read_config.cpp:136-139), and the exchange and
    Landau tables carry g^2, g^4, g^6, g^8 (potential_CPU.cpp:125-160). 4 steps
    of core() with thermostats and the electron temperature: moments and
    state bit for bit, and a trajectory that differs from the same start read
    as spins."""
    at = configs.make_bcc_fe(8, 600.0, 9, tilt_max_deg=15.0, displace_sigma=0.05)
    at.s = at.s * -K.EL_G   # the moments of the same magnetic state
    v = configs.fe_variables(600.0, 9, thermostats=True, eltemp=True, localcolmot=True, renormalize=False)
    st = refpy.consistent_start(v, at, lib=refpy.REF_LIB_MM)
    d = tempfile.mkdtemp()
    ref = refpy.Reference(refpy.REF_LIB_MM)
    ref.init_memory(d, v, st)
    for cur in range(4):
        ref.run(cur, 1, False)
    a = ref.atoms()
    m_ref = ref.moments()
    ref.free()
    p = abi.params_from_variables(dict(v, **refpy.REF_SWITCHES), st.box, st.natom)
    p.renormalize_energy = 0
    p.magmom = 1
    o = orcpy.Oracle(p, st, rng=orcpy.RNG_MWC)
    for cur in range(4):
        o.core(cur)
    b = o.download()
    o.close()
    # the moments are -g times the spins the reference keeps beside them
    np.testing.assert_array_equal(m_ref, -K.EL_G * a["s"])
    np.testing.assert_array_equal(b["s"], m_ref)
    for k in ("r", "p", "f", "rho", "pe", "me", "me0", "stress"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    # the same column read as spins: the same energies (J g^2 |m/g|^2 = J |m|^2)
    # but precession g times slower, so another trajectory
    p.magmom = 0
    o = orcpy.Oracle(p, st, rng=orcpy.RNG_MWC)
    for cur in range(4):
        o.core(cur)
    c = o.download()
    o.close()
    assert np.abs(c["s"] - b["s"]).max() > 1e-3 * np.abs(b["s"]).max()
