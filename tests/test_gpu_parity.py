"""GPU parity: the HIP engine against the plain-C oracle on the same seeded
inputs, through the C ABI (spilady_amd/engine.py -> libspilady_hip.so).

Both sides draw Langevin noise from the same counter-based Philox stream, so
thermostatted runs are compared step by step too. Floating-point fields are
held to a normwise relative error of 1e-10 per array (helpers.TOL); integer
data (cell lists, groups) must be identical.
"""
import numpy as np
import pytest

from helpers import TOL, assert_close, consistent_state, fe_params, relerr
from oracle import orcpy
from spilady_amd import constants as K

pytestmark = pytest.mark.gpu

# the effective fields are compared with the state: Heff_H and Heff_L as the
# last spin update or temperature pass left them, and sum_j J|s_j|
# (calculate_spin_CPU.cpp:111-194, calculate_temperature_CPU.cpp:94-110)
ATOM_FIELDS = ("r", "p", "f", "s", "heff_h", "heff_l", "sum_jij_sj", "rho", "pe", "ke", "me", "me0",
               "stress")
CELL_FIELDS = ("te", "tl", "ts_r", "ts_l", "gel", "ges")


def _pair(nu, T, seed, **kw):
    from spilady_amd.engine import Engine
    v, p, at = fe_params(nu, T, seed, **kw)
    st = consistent_state(p, at)
    eng = Engine(p, st, compute_initial_forces=False)
    orc = orcpy.Oracle(p, st, compute_initial_forces=False, rng=orcpy.RNG_PHILOX)
    return p, eng, orc


def _compare(eng, orc, tag, fields=ATOM_FIELDS, tol=TOL, cells=False):
    a, b = orc.download(), eng.download()
    worst = {}
    for k in fields:
        worst[k] = assert_close(a[k], b[k], f"{tag}:{k}", tol)
    if cells:
        ca, cb = orc.cells(), eng.cells()
        for k in CELL_FIELDS:
            worst["cell_" + k] = assert_close(ca[k], cb[k], f"{tag}:cell {k}", tol)
    return worst


def test_initial_force_energy():
    """calculate_rho + calculate_force_energy from the same positions and
    spins (no thermostats)."""
    from spilady_amd.engine import Engine
    v, p, at = fe_params(10, 300.0, 1, thermostats=False)
    st = consistent_state(p, at)
    eng = Engine(p, at, compute_initial_forces=True)
    b = eng.download()
    for k in ("rho", "f", "pe", "me", "me0", "stress"):
        assert_close(st[k], b[k], k)
    eng.close()


def test_cell_geometry_and_groups():
    p, eng, orc = _pair(8, 300.0, 3, thermostats=False)
    assert eng.ncell3 == orc.ncell3
    sizes, _ = orc.groups()
    np.testing.assert_array_equal(eng.groups(), sizes)
    c1, l1 = orc.cell_lists(32)
    c2, l2 = eng.cell_lists(32)
    np.testing.assert_array_equal(c1, c2)
    np.testing.assert_array_equal(l1, l2)
    eng.close(); orc.close()


@pytest.mark.parametrize("nu", [10, 16])
def test_entry_points_one_by_one_full_physics(nu):
    """Every reference entry point in core() order, thermostats, eltemp and
    local collective motion on; 10^3 unit cells (6 link cells per side,
    8 groups) and 16^3 (11 per side, 18 groups)."""
    p, eng, orc = _pair(nu, 600.0, 7, thermostats=True, eltemp=True, localcolmot=True,
                        renormalize=False)
    dt = p.step
    seq = [("links", ()), ("core_dTe", (dt / 2,)), ("core_dp", (dt / 2,)), ("core_ds", (dt / 2,)),
           ("core_dr", (dt,)), ("calculate_rho", ()), ("core_ds", (dt / 2,)),
           ("calculate_force_energy", ()), ("core_dp", (dt / 2,)),
           ("calculate_temperature", ()), ("core_dTe", (dt / 2,))]
    for name, args in seq:
        getattr(orc, name)(*args)
        getattr(eng, name)(*args)
        _compare(eng, orc, name, cells=True)
    eng.close(); orc.close()


def test_steps_nve():
    """Thermostats off: 20 full steps, observables of every step and the
    final state."""
    p, eng, orc = _pair(10, 300.0, 1, thermostats=False)
    oa = orc.step(0, 20)
    ob = eng.step(0, 20)
    assert len(oa) == len(ob) == 20
    for x, y in zip(oa, ob):
        for k in ("ave_pe", "ave_ke", "ave_me", "ave_me0", "tl", "ts_r", "ts_l", "ave_s0", "virial", "pressure"):
            assert abs(x[k] - y[k]) <= TOL * max(abs(x[k]), 1e-300), (x["step"], k, x[k], y[k])
        # off-diagonal stresses are small sums of large per-atom terms: normwise
        assert_close(x["ave_stress"], y["ave_stress"], f"stress at step {x['step']}")
    _compare(eng, orc, "nve-20")
    eng.close(); orc.close()


def test_steps_langevin_eltemp_renormalize():
    """All switches on, renormalizeEnergy included, 5 steps (the reference's
    own Te stays finite that long from this start)."""
    p, eng, orc = _pair(10, 300.0, 1, thermostats=True, eltemp=True, localcolmot=True,
                        renormalize=True)
    oa = orc.step(0, 5)
    ob = eng.step(0, 5)
    for x, y in zip(oa, ob):
        for k in ("ave_pe", "ave_ke", "ave_me", "ave_me0", "ave_ee", "tl", "ts_l", "te"):
            assert abs(x[k] - y[k]) <= 1e-9 * max(abs(x[k]), 1e-300), (x["step"], k, x[k], y[k])
    _compare(eng, orc, "lang-5", cells=True, tol=1e-9)
    eng.close(); orc.close()


def test_cell_lists_after_movers():
    """Hot lattice (1400 K) for 30 steps: atoms change link cells and the
    per-cell list order must follow the reference's relink rule exactly."""
    p, eng, orc = _pair(8, 1400.0, 11, thermostats=True)
    orc.step(0, 30)
    eng.step(0, 30)
    c1, l1 = orc.cell_lists(32)
    c2, l2 = eng.cell_lists(32)
    np.testing.assert_array_equal(c1, c2)
    np.testing.assert_array_equal(l1, l2)
    eng.close(); orc.close()


@pytest.mark.parametrize("nu,peak", [(16, 2400.0), (24, 2400.0), (31, 2400.0), (16, 9400.0)])
def test_core_dte_hot_spot(nu, peak):
    """core_dTe on a non-uniform electron temperature (a hot sphere over
    cell-to-cell noise), three calls of 50 sub-steps each. 16^3 and 31^3 unit
    cells give odd link-cell counts per side (11, 21: 18 sweep levels), 24^3
    an even one (16: 4 levels). A 10,000 K peak puts the sphere's core above
    6,250 K, where the kernels' heat-capacity functions leave their fitted
    polynomials for the library route, with cooler cells in the same waves."""
    p, eng, orc = _pair(nu, 600.0, 5, thermostats=True, eltemp=True, localcolmot=True,
                        renormalize=False)
    for o in (orc, eng):
        o.links()
        o.calculate_temperature()
    nx, ny, nz = orc.ncell3
    c = np.arange(nx * ny * nz)
    x, y, z = c % nx, (c // nx) % ny, c // (nx * ny)
    d2 = (x - nx / 2) ** 2 + (y - ny / 2) ** 2 + (z - nz / 2) ** 2
    te_k = 600.0 + peak * np.exp(-d2 / (0.1 * nx * nx)) + np.random.default_rng(nu).uniform(0, 50, c.size)
    te = te_k * K.BOLTZ   # cells hold Te in eV
    for o in (orc, eng):
        o.set_cells_te(te)
    dt = p.step
    for k in range(3):
        orc.core_dTe(dt / 2)
        eng.core_dTe(dt / 2)
        a, b = orc.cells()["te"], eng.cells()["te"]
        assert_close(a, b, f"dTe call {k}", TOL)
    # diffusion moved heat out of the hot spot, and the total is kept
    assert b.max() < te.max()
    eng.close(); orc.close()


def test_core_dte_graph_replays_equal_eager_launches(monkeypatch):
    """core_dTe is captured into a graph on its first call and replayed after:
    four replays, with eager momentum, temperature and download work between
    them, bitwise against an engine that launches every kernel (SLD_GRAPHS=0)."""
    from spilady_amd.engine import Engine
    v, p, at = fe_params(16, 600.0, 9, thermostats=True, eltemp=True, localcolmot=True, renormalize=False)
    st = consistent_state(p, at)
    engines = []
    for flag in ("1", "0"):
        monkeypatch.setenv("SLD_GRAPHS", flag)
        engines.append(Engine(p, st, compute_initial_forces=False))
    dt = p.step
    for k in range(4):
        for e in engines:
            e.core_dTe(dt / 2)
            e.core_dp(dt / 2)
            e.calculate_temperature()
        a, b = (e.cells() for e in engines)
        for f in ("te", "tl", "gel", "ts_l"):
            np.testing.assert_array_equal(a[f], b[f], err_msg=f"replay {k} {f}")
        np.testing.assert_array_equal(engines[0].download()["p"], engines[1].download()["p"])
    for e in engines:
        e.close()


def test_changestep_matches_oracle():
    """changestep: 16 steps of core(), check() and scale_step() with
    thermostats and the electron temperature (renormalisation off); the time
    step sequence exactly, observables and final state at the tolerance."""
    from spilady_amd.engine import Engine
    v, p, at = fe_params(8, 600.0, 5, thermostats=True, eltemp=True, localcolmot=True)
    p.changestep = 1
    p.displace_limit = 0.01
    p.phi_limit = 0.628319
    st = consistent_state(p, at)
    eng = Engine(p, st, compute_initial_forces=False)
    orc = orcpy.Oracle(p, st, compute_initial_forces=False, rng=orcpy.RNG_PHILOX)
    dts_e, dts_o = [], []
    for s in range(16):
        x = orc.step(s, 1)[0]
        y = eng.step(s, 1)[0]
        dts_o.append(orc.dt)
        dts_e.append(eng.dt)
        for k in ("ave_pe", "ave_ke", "ave_me", "tl", "ts_r", "te", "pressure"):
            assert abs(x[k] - y[k]) <= TOL * max(abs(x[k]), 1e-300), (s, k, x[k], y[k])
    assert dts_e == dts_o
    steps = np.array(dts_e)
    assert np.sum(steps[1:] < steps[:-1]) >= 1 and np.sum(steps[1:] > steps[:-1]) >= 1
    _compare(eng, orc, "changestep", cells=True)
    eng.close(); orc.close()


def test_barostats_match_oracle():
    """The stress and pressure barostats with changestep after them, on a
    compressed lattice (a = 2.80 A, about 13 GPa) with thermostats and the
    electron temperature: 12 steps, stress every 2 and pressure every 3, each
    fed by the check() of the step before it. The box after every step and
    the final state at the tolerance, the time step sequence exactly; a
    batched context refuses barostats."""
    from spilady_amd import abi, configs
    from spilady_amd.engine import BatchEngine, Engine, SldError
    at = configs.make_bcc_fe(8, 600.0, 5, a=2.80, tilt_max_deg=10.0)
    v = configs.fe_variables(600.0, 5, thermostats=True, eltemp=True, localcolmot=True)
    v.update({"STRESS": 1, "PRESSURE": 1, "interval_of_scale_stress": 2, "interval_of_scale_pressure": 3,
              "baro_damping_time": 2e-14, "stress_xx": 1.0, "stress_yx": 0.5, "stress_yy": -1.0,
              "stress_zx": 0.2, "stress_zy": -0.3, "stress_zz": 2.0, "pressure": 5.0, "changestep": 1})
    p = abi.params_from_variables(v, at.box, at.natom)
    st = consistent_state(p, at)
    baro = abi.barostat_from_variables(v)
    eng = Engine(p, st, compute_initial_forces=False)
    orc = orcpy.Oracle(p, st, compute_initial_forces=False, rng=orcpy.RNG_PHILOX)
    eng.set_barostat(baro)
    orc.set_barostat(baro)
    dts_e, dts_o = [], []
    for s in range(12):
        x = orc.step(s, 1)[0]
        y = eng.step(s, 1)[0]
        dts_o.append(orc.dt)
        dts_e.append(eng.dt)
        bo, be = orc.box(), eng.box()
        assert relerr(bo, be) <= TOL, (s, bo, be)
        for k in ("ave_pe", "ave_ke", "tl", "te", "pressure"):
            assert abs(x[k] - y[k]) <= TOL * max(abs(x[k]), 1e-300), (s, k, x[k], y[k])
        assert relerr(x["ave_stress"], y["ave_stress"]) <= TOL, (s, x["ave_stress"], y["ave_stress"])
    assert dts_e == dts_o
    assert np.all(np.abs(eng.box() - np.asarray(at.box)) > 1e-3)
    _compare(eng, orc, "barostats", cells=True)
    eng.close(); orc.close()
    vb = dict(v, eltemp=0, localcolmot=0, changestep=0)
    q = [abi.params_from_variables(vb, at.box, at.natom, replica=r) for r in range(2)]
    b = BatchEngine(q, [at, at], compute_initial_forces=True)
    with pytest.raises(SldError):
        b.set_barostat(baro)
    b.close()


def test_quantum_noise_matches_oracle():
    """The quantum thermostat (coloured lattice noise from a per-degree-of-
    freedom history filtered by H) with M = 3 and Nf = 8, thermostats, the
    electron temperature and localcolmot: 10 steps (seven redraws), the
    observables and the final state at the tolerance. The history ring on
    the GPU and the oracle's shifted arrays hold the same Philox elements."""
    from spilady_amd.engine import Engine
    v, p, at = fe_params(8, 300.0, 9, thermostats=True, eltemp=True, localcolmot=True)
    p.quantumnoise, p.msteps_quantum, p.nfrequency_quantum = 1, 3, 8
    st = consistent_state(p, at)
    eng = Engine(p, st, compute_initial_forces=False)
    orc = orcpy.Oracle(p, st, compute_initial_forces=False, rng=orcpy.RNG_PHILOX)
    oa = orc.step(0, 10)
    ea = eng.step(0, 10)
    for x, y in zip(oa, ea):
        for k in ("ave_pe", "ave_ke", "tl", "te", "ts_r"):
            assert abs(x[k] - y[k]) <= TOL * max(abs(x[k]), 1e-300), (x["step"], k, x[k], y[k])
    _compare(eng, orc, "quantumnoise", cells=True)
    eng.close(); orc.close()


def test_local_volume_matches_oracle():
    """localvol (calculate_force_energy_CPU.cpp:80-90,266-293): per-atom
    volumes from the sums of 1/r and 1/r^2 within rcut_vol (2.7 A, between the
    first and second shells), scaled to fill the box, and the stress over
    them; displaced atoms so the volumes differ. The initial evaluation, then
    6 steps with thermostats and the electron temperature: the local volumes,
    the stress, the str- averages and the state at the tolerance."""
    from spilady_amd.engine import Engine
    v, p, at = fe_params(8, 600.0, 9, thermostats=True, eltemp=True, localcolmot=True, tilt=15.0,
                         displace=0.05)
    p.localvol = 1
    p.rcut_vol = 2.7
    st = consistent_state(p, at)
    eng = Engine(p, st, compute_initial_forces=True)
    orc = orcpy.Oracle(p, st, compute_initial_forces=True, rng=orcpy.RNG_PHILOX)
    a, b = orc.download(), eng.download()
    assert np.ptp(a["local_volume"]) > 0.01 * a["local_volume"].mean()
    assert_close(a["local_volume"], b["local_volume"], "localvol:initial volume")
    assert_close(a["stress"], b["stress"], "localvol:initial stress")
    oa = orc.step(0, 6)
    ob = eng.step(0, 6)
    for x, y in zip(oa, ob):
        for k in ("ave_pe", "ave_ke", "pressure"):
            assert abs(x[k] - y[k]) <= TOL * max(abs(x[k]), 1e-300), (x["step"], k, x[k], y[k])
        assert relerr(x["ave_stress"], y["ave_stress"]) <= TOL, x["step"]
    _compare(eng, orc, "localvol", fields=ATOM_FIELDS + ("local_volume",), cells=True)
    eng.close(); orc.close()


def test_magmom_matches_oracle():
    """magmom (control.h:36): the spin column holds the moments m, the spins
    are s = m / (-g) inside each spin pass and m = -g s after it
    (core_ds_CPU.cpp:37-81), with the exchange and Landau tables scaled by
    powers of g (potential_CPU.cpp:125-160). The initial evaluation and 6
    steps with thermostats and the electron temperature, then new moments
    uploaded into both and 3 more steps: the moments and the state at the
    tolerance."""
    from spilady_amd.engine import Engine
    v, p, at = fe_params(8, 600.0, 21, thermostats=True, eltemp=True, localcolmot=True)
    at.s = at.s * -K.EL_G
    p.magmom = 1
    st = consistent_state(p, at)
    np.testing.assert_array_equal(st["s"], at.s)   # the column comes back as moments
    eng = Engine(p, st, compute_initial_forces=True)
    orc = orcpy.Oracle(p, st, compute_initial_forces=True, rng=orcpy.RNG_PHILOX)
    _compare(eng, orc, "magmom:initial")
    orc.step(0, 6)
    eng.step(0, 6)
    _compare(eng, orc, "magmom:6 steps", cells=True)
    # a moment field rotated by a fixed random rotation, uploaded into both
    b = orc.download()
    q, _ = np.linalg.qr(np.random.default_rng(5).standard_normal((3, 3)))
    b["s"] = b["s"] @ q
    orc.upload(b)
    eng.upload(b)
    orc.step(6, 3)
    eng.step(6, 3)
    _compare(eng, orc, "magmom:after upload", cells=True)
    eng.close(); orc.close()


def test_external_field_and_force_match_oracle():
    """extfield and extforce: random per-atom fields (about 10 T) and forces
    (0.1 eV/A), 8 steps with thermostats and the electron temperature; the
    observables and the final state at the tolerance."""
    from spilady_amd.engine import Engine
    v, p, at = fe_params(8, 600.0, 7, thermostats=True, eltemp=True, localcolmot=True, tilt=20.0)
    p.extfield = 1
    p.extforce = 1
    st = consistent_state(p, at)
    n = p.natom
    rng = np.random.default_rng(4)
    hext = (-2.0023193043622e0 * 5.7883817555e-5) * rng.normal(0.0, 10.0, size=(n, 3))
    fext = rng.normal(0.0, 0.1, size=(n, 3))
    eng = Engine(p, st, compute_initial_forces=False)
    orc = orcpy.Oracle(p, st, compute_initial_forces=False, rng=orcpy.RNG_PHILOX)
    for x in (eng, orc):
        x.set_external(hext, fext)
        x.calculate_temperature()
    oa = orc.step(0, 8)
    ob = eng.step(0, 8)
    for x, y in zip(oa, ob):
        for k in ("ave_pe", "ave_ke", "ave_me", "ave_me0", "tl", "ts_r", "ts_l", "te"):
            assert abs(x[k] - y[k]) <= TOL * max(abs(x[k]), 1e-300), (x["step"], k, x[k], y[k])
    _compare(eng, orc, "external", cells=True)
    eng.close(); orc.close()
