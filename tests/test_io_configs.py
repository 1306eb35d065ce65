"""SPILADY file formats and the seeded bcc Fe generator."""
import os

import numpy as np

from spilady_amd import abi, configs, io
from spilady_amd import constants as K


def test_bcc100_order_and_offset():
    """bcc100.cpp:55-87: for z, y, x unit cells the corner then the body
    centre, everything shifted by a/4."""
    a = 2.0
    r = configs.bcc100_positions(3, 2, 2, a)
    assert r.shape == (24, 3)
    np.testing.assert_allclose(r[0], [0.5, 0.5, 0.5])
    np.testing.assert_allclose(r[1], [1.5, 1.5, 1.5])
    np.testing.assert_allclose(r[2], [2.5, 0.5, 0.5])          # next x cell
    np.testing.assert_allclose(r[6], [0.5, 2.5, 0.5])          # next y row
    np.testing.assert_allclose(r[12], [0.5, 0.5, 2.5])         # next z layer
    q = r / np.array([3 * a, 2 * a, 2 * a])
    assert q.min() >= 0.25 / 3 - 1e-12 and q.max() < 1.0


def test_generator_statistics():
    at = configs.make_bcc_fe(12, 600.0, 7)
    m = K.mass_internal(K.FE_MASS_U)
    np.testing.assert_allclose(at.p.sum(axis=0), 0.0, atol=1e-25)
    t = (at.p ** 2).sum() / (3 * at.natom) / m / K.BOLTZ
    assert abs(t - 600.0) < 600.0 * 0.05
    s0 = np.linalg.norm(at.s, axis=1)
    np.testing.assert_allclose(s0, 2.2 / K.EL_G, rtol=1e-14)
    tilt = np.degrees(np.arccos(at.s[:, 2] / s0))
    assert tilt.max() <= 10.0 and tilt.min() >= 0.0
    # seeded: same seed, same atoms
    np.testing.assert_array_equal(configs.make_bcc_fe(12, 600.0, 7).p, at.p)


def test_con_file_round_trip(tmp_path):
    at = configs.make_bcc_fe(3, 300.0, 2)
    at.f[:] = np.random.default_rng(0).normal(size=at.f.shape)
    at.heff_h = np.random.default_rng(1).normal(size=(at.natom, 3))
    path = tmp_path / "con-test_0000.dat"
    io.write_con(str(path), at)
    back = io.read_con(str(path))
    for k in ("r", "p", "f", "s", "stress", "rho", "ke", "pe", "me", "me0"):
        np.testing.assert_allclose(getattr(back, k), getattr(at, k), rtol=1e-15, atol=0)
    np.testing.assert_array_equal(back.box, at.box)
    # the Heff y column repeats x (write_config.cpp:117)
    np.testing.assert_allclose(back.heff_h[:, 1], at.heff_h[:, 0], rtol=1e-15)
    with open(path) as fh:
        head = fh.readline().split()
    assert head[0] == str(at.natom) and head[2] == "SLDHL"
    assert io.con_name("run", 3) == "con-run_0003.dat"


def test_variables_round_trip(tmp_path):
    v = configs.fe_variables(450.0, 17, thermostats=True, eltemp=True)
    path = tmp_path / "variables.in"
    io.write_variables(str(path), v)
    back = io.read_variables(str(path))
    for k, x in v.items():
        assert back[k] == x, k
    p = abi.params_from_variables(back, [10, 0, 10, 0, 0, 10], 100)
    assert abs(p.temperature - 450.0 * K.BOLTZ) < 1e-18
    assert p.eltemp == 1 and p.spinlang == 1 and p.random_seed == 17


def test_min_length_link_cell_raised_to_rcut():
    """read_variables.cpp:262: rcut_max above min_length_link_cell raises it."""
    v = configs.fe_variables(300.0, 1, thermostats=False)
    v["min_length_link_cell"] = 3.0
    p = abi.params_from_variables(v, [10, 0, 10, 0, 0, 10], 10)
    assert abs(p.min_length_link_cell - (4.1001 + 0.001)) < 1e-12


def test_driver_reports_missing_inputs(tmp_path):
    """The C++ driver fails with a message, before touching any device, when
    variables.in or the initial configuration is missing."""
    import subprocess
    drv = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "spilady_amd",
                       "spilady_hip")
    r = subprocess.run([drv, "nope.in"], cwd=tmp_path, capture_output=True, text=True, timeout=30)
    assert r.returncode == 1 and "cannot open nope.in" in r.stderr
    (tmp_path / "variables.in").write_text("in_config con-missing.dat\n")
    r = subprocess.run([drv], cwd=tmp_path, capture_output=True, text=True, timeout=30)
    assert r.returncode == 1 and "con-missing.dat" in r.stderr


def test_cel_file_round_trip(tmp_path):
    """cel- files as write_config.cpp:160-200 writes them and the readTe start
    reads them (map_cells.cpp:33-60,125-160): Te in K, cells x outermost."""
    rng = np.random.default_rng(3)
    n3 = (4, 3, 5)
    C = int(np.prod(n3))
    box = [12.0, 0.0, 9.0, 0.0, 0.0, 15.0]
    te = rng.uniform(300.0, 900.0, C) * K.BOLTZ
    tl, tr, ts = (rng.uniform(200.0, 800.0, C) for _ in range(3))
    path = str(tmp_path / "cel-x_0001.dat")
    io.write_cel(path, box, n3, te, tl, tr, ts, 2.5e-13)
    t, b, cols = io.read_cel(path)
    assert t == 2.5e-13 and np.array_equal(b, box) and cols.shape == (C, 7)
    order = [ix + iy * n3[0] + iz * n3[0] * n3[1] for ix in range(n3[0]) for iy in range(n3[1]) for iz in range(n3[2])]
    np.testing.assert_allclose(cols[:, 3] * K.BOLTZ, te[order], rtol=1e-14)
    np.testing.assert_array_equal(cols[:, 4], [float("%.15e" % x) for x in tl[order]])
    # cell centres from the box matrix
    ix, iy, iz = order[7] % n3[0], (order[7] // n3[0]) % n3[1], order[7] // (n3[0] * n3[1])
    np.testing.assert_allclose(cols[7, :3], [(ix + 0.5) / n3[0] * 12.0, (iy + 0.5) / n3[1] * 9.0,
                                             (iz + 0.5) / n3[2] * 15.0], rtol=1e-14)
