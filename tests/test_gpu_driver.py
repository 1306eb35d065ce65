"""The C++ host driver (spilady_amd/spilady_hip, spilady.cpp's loop): its
enr-/tmp-/spn-/str-/prs- logs and con- files against the in-process engine started
from the same con- file."""
import os
import subprocess

import numpy as np
import pytest

from helpers import consistent_state, fe_params
from spilady_amd import configs, io

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "spilady_amd",
                      "spilady_hip")


def _setup(tmp, eltemp, **extra):
    v, p, at = fe_params(8, 600.0, 21, thermostats=True, eltemp=eltemp, localcolmot=eltemp)
    st = consistent_state(p, at)
    state = io.AtomState(box=np.asarray(at.box), r=st["r"], p=st["p"], f=st["f"], s=st["s"],
                         stress=st["stress"], rho=st["rho"], ke=st["ke"], pe=st["pe"], me=st["me"],
                         me0=st["me0"])
    io.write_con(os.path.join(tmp, "con-start.dat"), state)
    v = dict(v, in_config="con-start.dat", out_body="drv", no_of_production_steps=7,
             interval_of_print_out=2, interval_of_config_out=3)
    v.update(extra)
    io.write_variables(os.path.join(tmp, "variables.in"), v)
    return v


@pytest.mark.parametrize("eltemp", [False, True])
def test_driver_logs_match_engine(tmp_path, eltemp):
    from spilady_amd import abi
    from spilady_amd.engine import Engine
    tmp = str(tmp_path)
    v = _setup(tmp, eltemp)
    res = subprocess.run([DRIVER, "variables.in"], cwd=tmp, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    # the same start, read back from the file the driver read
    st = io.read_con(os.path.join(tmp, "con-start.dat"))
    p = abi.params_from_variables(io.read_variables(os.path.join(tmp, "variables.in")), st.box, st.natom)
    e = Engine(p, st, compute_initial_forces=False)
    obs = [e.check(-1)] + e.step(0, 7)
    fmt = lambda x: float("%.15e" % x)  # noqa: E731
    enr = io.read_log(os.path.join(tmp, "enr-drv.dat"))
    tmpl = io.read_log(os.path.join(tmp, "tmp-drv.dat"))
    spn = io.read_log(os.path.join(tmp, "spn-drv.dat"))
    assert enr.shape[0] == len(obs) == 4          # steps -1, 1, 3, 5
    for row_e, row_t, row_s, o in zip(enr, tmpl, spn, obs):
        assert int(row_e[0]) == o["step"]
        want_e = [o["ave_pe"], o["ave_ke"], o["ave_me"], o["ave_me0"]] + ([o["ave_ee"]] if eltemp else [])
        assert list(row_e[2:]) == [fmt(x) for x in want_e]
        want_t = [o["tl"], o["ts_r"], o["ts_l"]] + ([o["te"]] if eltemp else [])
        assert list(row_t[2:]) == [fmt(x) for x in want_t]
        want_s = list(o["ave_s"]) + [o["ave_s0"]] + list(o["ave_m"]) + [o["ave_m0"]]
        assert list(row_s[2:]) == [fmt(x) for x in want_s]
    # str- and prs-: the box matrix, the density, the stresses or the pressure
    strl = io.read_log(os.path.join(tmp, "str-drv.dat"))
    prsl = io.read_log(os.path.join(tmp, "prs-drv.dat"))
    assert strl.shape == (len(obs), 16) and prsl.shape == (len(obs), 10)
    box = [float(x) for x in st.box]
    dens = st.natom / (box[0] * box[2] * box[5])
    for row_st, row_p, o in zip(strl, prsl, obs):
        sv = list(o["ave_stress"])
        assert list(row_st[2:]) == [fmt(x) for x in box + [dens, sum(sv[:3]) / 3e0] + sv]
        assert list(row_p[2:]) == [fmt(x) for x in box + [dens, o["pressure"]]]
    # configurations after steps 3 and 6, and the final one (write(-2))
    names = sorted(f for f in os.listdir(tmp) if f.startswith("con-drv_"))
    assert names == ["con-drv_0000.dat", "con-drv_0001.dat", "con-drv_0002.dat", "con-drv_9999.dat"]
    fin = io.read_con(os.path.join(tmp, "con-drv_9999.dat"))
    a = e.download()
    for k in ("r", "p", "f", "s", "rho", "pe", "me", "me0"):
        np.testing.assert_array_equal(getattr(fin, k), np.vectorize(fmt)(a[k]), err_msg=k)
    # local_volume = box_volume / natom (calculate_force_energy_CPU.cpp:276)
    assert np.all(fin.local_volume == fmt(box[0] * box[2] * box[5] / st.natom))
    if eltemp:
        assert os.path.exists(os.path.join(tmp, "cel-drv_9999.dat"))
    e.close()


def test_driver_time_loop_with_changestep(tmp_path):
    """The shipped control.h loop: runstep off (run while total_time <=
    total_production_time) with changestep on; the driver's step count, times
    and energies against the in-process engine stepped the same way."""
    from spilady_amd import abi
    from spilady_amd.engine import Engine
    tmp = str(tmp_path)
    _setup(tmp, True, runstep=0, changestep=1, total_production_time=6e-15, displace_limit=0.01,
           phi_limit=0.628319, interval_of_print_out=1)
    res = subprocess.run([DRIVER, "variables.in"], cwd=tmp, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    st = io.read_con(os.path.join(tmp, "con-start.dat"))
    v = io.read_variables(os.path.join(tmp, "variables.in"))
    p = abi.params_from_variables(v, st.box, st.natom)
    assert p.changestep == 1
    e = Engine(p, st, compute_initial_forces=False)
    obs = [e.check(-1)]
    dts = [e.dt]
    s = 0
    while e.total_time <= 6e-15 + 1e-18:
        obs += e.step(s, 1)
        dts.append(e.dt)
        s += 1
    e.close()
    fmt = lambda x: float("%.15e" % x)  # noqa: E731
    enr = io.read_log(os.path.join(tmp, "enr-drv.dat"))
    assert enr.shape[0] == len(obs)
    assert len(set(dts)) > 2                          # the step did change
    for row, o in zip(enr, obs):
        assert int(row[0]) == o["step"]
        assert row[1] == fmt(o["total_time"])
        assert list(row[2:6]) == [fmt(o[k]) for k in ("ave_pe", "ave_ke", "ave_me", "ave_me0")]


def test_driver_reads_external_field_and_force(tmp_path):
    """extfield.in (tesla) and extforce.in (eV/A) in the run directory, as
    external_field_CPU.cpp / external_force_CPU.cpp read them: the driver's
    logs against the in-process engine given the same values."""
    from spilady_amd import abi
    from spilady_amd.engine import Engine
    tmp = str(tmp_path)
    _setup(tmp, True, extfield=1, extforce=1)
    st = io.read_con(os.path.join(tmp, "con-start.dat"))
    n = st.natom
    rng = np.random.default_rng(8)
    b = rng.normal(0.0, 5.0, size=(n, 3))
    f = rng.normal(0.0, 0.05, size=(n, 3))
    for name, arr in (("extfield.in", b), ("extforce.in", f)):
        with open(os.path.join(tmp, name), "w") as fh:
            fh.write(f"{n}\n")
            for i in range(n):
                fh.write("%d %.17g %.17g %.17g\n" % (i, arr[i, 0], arr[i, 1], arr[i, 2]))
    res = subprocess.run([DRIVER, "variables.in"], cwd=tmp, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    p = abi.params_from_variables(io.read_variables(os.path.join(tmp, "variables.in")), st.box, n)
    assert p.extfield == 1 and p.extforce == 1
    e = Engine(p, st, compute_initial_forces=False)
    e.set_external((-2.0023193043622e0 * 5.7883817555e-5) * b, f)
    e.calculate_temperature()
    obs = [e.check(-1)] + e.step(0, 7)
    e.close()
    fmt = lambda x: float("%.15e" % x)  # noqa: E731
    enr = io.read_log(os.path.join(tmp, "enr-drv.dat"))
    tmpl = io.read_log(os.path.join(tmp, "tmp-drv.dat"))
    assert enr.shape[0] == len(obs)
    for row_e, row_t, o in zip(enr, tmpl, obs):
        assert list(row_e[2:]) == [fmt(o[k]) for k in ("ave_pe", "ave_ke", "ave_me", "ave_me0", "ave_ee")]
        assert list(row_t[2:]) == [fmt(o[k]) for k in ("tl", "ts_r", "ts_l", "te")]


def test_driver_barostats_log_the_moving_box(tmp_path):
    """STRESS and PRESSURE from variables.in: the box in every str-/prs- line
    is the box at that check (before the step's scale()), and the last con-
    file carries the final box and its local volume; all against the
    in-process engine with the same barostat settings."""
    from spilady_amd import abi
    from spilady_amd.engine import Engine
    tmp = str(tmp_path)
    _setup(tmp, False, interval_of_print_out=1, STRESS=1, PRESSURE=1, interval_of_scale_stress=2,
           interval_of_scale_pressure=3, baro_damping_time=2e-14, stress_xx=1.0, stress_yx=0.5,
           stress_yy=-1.0, stress_zx=0.2, stress_zy=-0.3, stress_zz=2.0, pressure=5.0)
    res = subprocess.run([DRIVER, "variables.in"], cwd=tmp, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    st = io.read_con(os.path.join(tmp, "con-start.dat"))
    v = io.read_variables(os.path.join(tmp, "variables.in"))
    p = abi.params_from_variables(v, st.box, st.natom)
    e = Engine(p, st, compute_initial_forces=False)
    e.set_barostat(abi.barostat_from_variables(v))
    obs = [e.check(-1)]
    boxes = [e.box()]
    for s in range(7):
        boxes.append(e.box())
        obs += e.step(s, 1)
    final_box = e.box()
    a = e.download()
    e.close()
    fmt = lambda x: float("%.15e" % x)  # noqa: E731
    strl = io.read_log(os.path.join(tmp, "str-drv.dat"))
    prsl = io.read_log(os.path.join(tmp, "prs-drv.dat"))
    assert strl.shape == (8, 16) and prsl.shape == (8, 10)
    for row_st, row_p, o, b in zip(strl, prsl, obs, boxes):
        b = [float(x) for x in b]
        dens = st.natom / (b[0] * b[2] * b[5])
        sv = list(o["ave_stress"])
        assert list(row_st[2:]) == [fmt(x) for x in b + [dens, sum(sv[:3]) / 3e0] + sv]
        assert list(row_p[2:]) == [fmt(x) for x in b + [dens, o["pressure"]]]
    assert np.all(np.abs(final_box - np.asarray(st.box)) > 1e-4)
    fin = io.read_con(os.path.join(tmp, "con-drv_9999.dat"))
    np.testing.assert_array_equal(np.asarray(fin.box), np.vectorize(fmt)(final_box))
    for k in ("r", "p", "s"):
        np.testing.assert_array_equal(getattr(fin, k), np.vectorize(fmt)(a[k]), err_msg=k)
    b = final_box
    assert np.all(fin.local_volume == fmt(b[0] * b[2] * b[5] / st.natom))


def test_driver_reads_cell_temperatures(tmp_path):
    """readTe: the cells start from a cel- file's electron temperatures (K)
    with Ee = E(Te); the driver's logs equal the in-process engine's given
    the same Te through sld_set_cells_te."""
    from spilady_amd import abi
    from spilady_amd import constants as K
    from spilady_amd.engine import Engine
    tmp = str(tmp_path)
    v = _setup(tmp, True, readTe=1, in_eltemp="cel-start.dat", interval_of_print_out=1)
    st = io.read_con(os.path.join(tmp, "con-start.dat"))
    p = abi.params_from_variables(io.read_variables(os.path.join(tmp, "variables.in")), st.box, st.natom)
    e = Engine(p, st, compute_initial_forces=False)
    C = e.ncells
    rng = np.random.default_rng(12)
    te = rng.uniform(300.0, 1500.0, C) * K.BOLTZ
    zero = np.zeros(C)
    io.write_cel(os.path.join(tmp, "cel-start.dat"), st.box, e.ncell3, te, zero, zero, zero, 0.0)
    # what the driver reads back: Te in K at 15 digits
    _, _, cols = io.read_cel(os.path.join(tmp, "cel-start.dat"))
    nx, ny, nz = e.ncell3
    te_read = np.zeros(C)
    k = 0
    for ix in range(nx):
        for iy in range(ny):
            for iz in range(nz):
                te_read[ix + iy * nx + iz * nx * ny] = cols[k, 3] * K.BOLTZ
                k += 1
    e.set_cells_te(te_read)
    obs = [e.check(-1)] + e.step(0, 7)
    e.close()
    res = subprocess.run([DRIVER, "variables.in"], cwd=tmp, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    fmt = lambda x: float("%.15e" % x)  # noqa: E731
    enr = io.read_log(os.path.join(tmp, "enr-drv.dat"))
    tmpl = io.read_log(os.path.join(tmp, "tmp-drv.dat"))
    assert enr.shape[0] == len(obs) == 8
    assert obs[0]["te"] == pytest.approx(te_read.mean() / K.BOLTZ, rel=1e-12)
    for row_e, row_t, o in zip(enr, tmpl, obs):
        assert list(row_e[2:]) == [fmt(x) for x in (o["ave_pe"], o["ave_ke"], o["ave_me"], o["ave_me0"], o["ave_ee"])]
        assert list(row_t[2:]) == [fmt(x) for x in (o["tl"], o["ts_r"], o["ts_l"], o["te"])]


def test_driver_variables_all_matches_reference(tmp_path):
    """variables.all (read_variables.cpp:278-418): for tests/golden/variables_in.txt
    (the shipped control.h's switches, writevsim included) the driver's echo
    equals the one the compiled reference wrote, line for line."""
    from spilady_amd import abi
    tmp = str(tmp_path)
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    with open(os.path.join(gold, "variables_in.txt")) as fh, open(os.path.join(tmp, "variables.in"), "w") as out:
        out.write(fh.read())
    v = io.read_variables(os.path.join(tmp, "variables.in"))
    at = configs.make_bcc_fe(8, 600.0, 21)
    st = consistent_state(abi.params_from_variables(v, at.box, at.natom), at)
    state = io.AtomState(box=np.asarray(at.box), r=st["r"], p=st["p"], f=st["f"], s=st["s"],
                         stress=st["stress"], rho=st["rho"], ke=st["ke"], pe=st["pe"], me=st["me"],
                         me0=st["me0"])
    io.write_con(os.path.join(tmp, v["in_config"]), state)
    res = subprocess.run([DRIVER, "variables.in"], cwd=tmp, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    with open(os.path.join(tmp, "variables.all")) as fh:
        ours = fh.read().splitlines()
    with open(os.path.join(gold, "variables_all_ref.txt")) as fh:
        ref = fh.read().splitlines()
    assert ours == ref


def test_driver_vsim_matches_reference(tmp_path):
    """writevsim (write_vsim.cpp:33-134): from the configuration of
    tests/golden/vsim_fe1k.npz and the variables of variables_in.txt, the
    driver's vsm-drv_0000.ascii/.spin/.dat equal the files the compiled
    reference wrote, character for character; later files follow the
    interval (steps 2, 4 of 5) and the final state (_9999)."""
    tmp = str(tmp_path)
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    g = np.load(os.path.join(gold, "vsim_fe1k.npz"), allow_pickle=False)
    with open(os.path.join(gold, "variables_in.txt")) as fh:
        text = fh.read()
    v = io.read_variables(os.path.join(gold, "variables_in.txt"))
    with open(os.path.join(tmp, "variables.in"), "w") as out:
        out.write(text + "runstep 1\nno_of_production_steps 5\n")
    state = io.AtomState(box=g["box"], **{k: g[k] for k in ("r", "p", "f", "s", "stress", "rho", "ke", "pe",
                                                             "me", "me0")})
    io.write_con(os.path.join(tmp, v["in_config"]), state)
    res = subprocess.run([DRIVER, "variables.in"], cwd=tmp, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    for ext in ("ascii", "spin", "dat"):
        with open(os.path.join(tmp, f"vsm-drv_0000.{ext}")) as fh:
            assert fh.read() == str(g[f"vsm_{ext}"]), ext
    names = sorted(f for f in os.listdir(tmp) if f.startswith("vsm-drv_") and f.endswith(".ascii"))
    assert names == ["vsm-drv_0000.ascii", "vsm-drv_0001.ascii", "vsm-drv_0002.ascii", "vsm-drv_9999.ascii"]
    fin = io.read_con(os.path.join(tmp, "con-drv_9999.dat"))
    with open(os.path.join(tmp, "vsm-drv_9999.ascii")) as fh:
        rows = fh.read().splitlines()[3:]
    assert len(rows) == fin.natom
    assert [float(x) for x in rows[5].split()[:3]] == [float("%.4e" % x) for x in fin.r[5]]


def test_driver_checkpoint_resume(tmp_path):
    """Checkpoint and resume through the con- files (write_config.cpp:31-131,
    read_config.cpp:32-147): 6 NVE steps in one run against 3 steps, then 3
    more started from the con- file written after step 3 with start_time set
    to its time (inittime). The con- file holds 16 significant digits, so the
    resumed run agrees to 1e-10, not bitwise: positions, momenta, spins and
    the energies of the last three printed steps."""
    from helpers import fe_params
    tmp = str(tmp_path)
    v, p, at = fe_params(8, 600.0, 23, thermostats=False)
    st = consistent_state(p, at)
    state = io.AtomState(box=np.asarray(at.box), r=st["r"], p=st["p"], f=st["f"], s=st["s"],
                         stress=st["stress"], rho=st["rho"], ke=st["ke"], pe=st["pe"], me=st["me"],
                         me0=st["me0"])
    full, part = os.path.join(tmp, "full"), os.path.join(tmp, "part")
    for d in (full, part):
        os.makedirs(d)
        io.write_con(os.path.join(d, "con-start.dat"), state)
    base = dict(v, in_config="con-start.dat", out_body="drv", interval_of_print_out=1, interval_of_config_out=3)
    io.write_variables(os.path.join(full, "variables.in"), dict(base, no_of_production_steps=6))
    io.write_variables(os.path.join(part, "variables.in"), dict(base, no_of_production_steps=3))
    for d in (full, part):
        res = subprocess.run([DRIVER, "variables.in"], cwd=d, capture_output=True, text=True, timeout=120)
        assert res.returncode == 0, res.stderr
    ck = io.read_con(os.path.join(part, "con-drv_0001.dat"))
    t3 = float(io.read_log(os.path.join(part, "enr-drv.dat"))[-1, 1])
    resume = os.path.join(tmp, "resume")
    os.makedirs(resume)
    io.write_con(os.path.join(resume, "con-ck.dat"), ck)
    io.write_variables(os.path.join(resume, "variables.in"),
                       dict(base, in_config="con-ck.dat", no_of_production_steps=3, start_time=t3))
    res = subprocess.run([DRIVER, "variables.in"], cwd=resume, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    a = io.read_con(os.path.join(full, "con-drv_9999.dat"))
    b = io.read_con(os.path.join(resume, "con-drv_9999.dat"))
    for k in ("r", "p", "s"):
        x, y = getattr(a, k), getattr(b, k)
        assert np.linalg.norm(x - y) <= 1e-10 * np.linalg.norm(x), k
    ea = io.read_log(os.path.join(full, "enr-drv.dat"))[-3:]
    eb = io.read_log(os.path.join(resume, "enr-drv.dat"))[-3:]
    np.testing.assert_allclose(eb[:, 1], ea[:, 1], rtol=1e-12)           # total time continues
    np.testing.assert_allclose(eb[:, 2:], ea[:, 2:], rtol=1e-10, atol=1e-12)
