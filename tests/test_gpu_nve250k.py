"""Microcanonical parity at scale: 50^3 cubic cells = 250,000 atoms, 300 K,
thermostats off, 500 steps on the GPU against the compiled reference's own
run of the same start (tests/golden/nve250k.npz, 8 OpenMP threads).

Every step's energies, temperatures and mean spin, and r, p, f, s, H_eff of
every 250th atom after steps 1, 10, 100, 250 and 500, must agree to 1e-10
(normwise relative per column or array; the fixture holds the values printed
with 16 significant digits). The start is rebuilt from the same generator
and seed with the oracle's initial forces, which equal the reference's to
rounding."""
import os

import numpy as np
import pytest

from helpers import TOL, assert_close, consistent_state, fe_params

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "nve250k.npz")


@pytest.mark.skipif(not os.path.exists(GOLD), reason="nve250k.npz not generated")
def test_nve_250k_500_steps_against_reference():
    from spilady_amd.engine import Engine
    g = np.load(GOLD, allow_pickle=False)
    v, p, at = fe_params(50, 300.0, 1, thermostats=False)
    st = consistent_state(p, at)
    e = Engine(p, st, compute_initial_forces=False)
    obs = [e.check(-1)]
    done = 0
    idx = g["atoms"]
    for s in g["snap_steps"]:
        obs += e.step(done, int(s) - done)
        done = int(s)
        a = e.download()
        for k in ("r", "p", "f", "s", "heff_h"):
            assert_close(g[f"step{s}_{k}"], a[k][idx], f"step {s} {k}", TOL)
    e.close()
    assert len(obs) == g["enr"].shape[0] == 501
    cols = {
        "ave_pe": g["enr"][:, 2], "ave_ke": g["enr"][:, 3], "ave_me": g["enr"][:, 4],
        "ave_me0": g["enr"][:, 5], "tl": g["tmp"][:, 2], "ts_r": g["tmp"][:, 3], "ts_l": g["tmp"][:, 4],
    }
    for k, ref in cols.items():
        assert_close(ref, np.array([o[k] for o in obs]), k, TOL)
    for c, k in ((2, 0), (3, 1), (4, 2)):
        assert_close(g["spn"][:, c], np.array([o["ave_s"][k] for o in obs]), f"ave_s[{k}]", TOL)
    assert_close(g["spn"][:, 9], np.array([o["ave_m0"] for o in obs]), "ave_m0", TOL)
    # total energy per atom, kinetic + potential + magnetic: the reference's
    # own value swings by about 9e-4 eV/atom over these 500 steps (the
    # splitting error of the 1 fs step while the tilted spins precess); the
    # GPU follows the same curve
    etot_ref = cols["ave_pe"] + cols["ave_ke"] + cols["ave_me"]
    etot = np.array([o["ave_pe"] + o["ave_ke"] + o["ave_me"] for o in obs])
    assert np.max(np.abs(etot - etot_ref)) <= TOL * np.max(np.abs(etot_ref))
