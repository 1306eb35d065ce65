"""Batched replicas (sld_create_batch): independent systems of different
size, temperature and seed advanced in the same kernel launches. Each replica
must follow exactly the trajectory it follows alone: noise is keyed by
(seed, replica, the replica's own atom index), reductions are per replica in
the single-system order, and replicas never share a link cell."""
import numpy as np
import pytest

from helpers import TOL, assert_close, consistent_state, fe_params
from oracle import orcpy

pytestmark = pytest.mark.gpu

FIELDS = ("r", "p", "f", "s", "rho", "pe", "ke", "me", "me0", "stress", "heff_h", "heff_l", "sum_jij_sj")
# compared with the oracle: the state and the effective fields
ORACLE_FIELDS = ("r", "p", "f", "s", "heff_h", "heff_l", "sum_jij_sj", "rho", "pe", "ke", "me", "me0")
# bcc Fe unit cells per side, temperature (K), seed: 5, 6 and 4 link cells per
# side, i.e. 20, 8 and 8 cell groups
REPLICAS = [(8, 300.0, 101), (10, 900.0, 102), (7, 1400.0, 103)]


def _obs_equal(a, b):
    return a.keys() == b.keys() and all(np.array_equal(a[k], b[k]) for k in a)


def _replicas(thermostats=True):
    out = []
    for r, (nu, T, seed) in enumerate(REPLICAS):
        v, p, at = fe_params(nu, T, seed, thermostats=thermostats, replica=r)
        out.append((p, consistent_state(p, at)))
    return out


@pytest.mark.parametrize("thermostats", [True, False])
def test_batch_equals_each_replica_alone(thermostats):
    from spilady_amd.engine import BatchEngine, Engine
    reps = _replicas(thermostats)
    nsteps = 12
    b = BatchEngine([p for p, _ in reps], [st for _, st in reps])
    assert b.nrep == len(reps)
    obs_b = b.step(0, nsteps)
    assert len(obs_b) == nsteps and all(len(x) == len(reps) for x in obs_b)
    for r, (p, st) in enumerate(reps):
        e = Engine(p, st)
        obs_s = e.step(0, nsteps)
        for k in range(nsteps):
            assert _obs_equal(obs_b[k][r], obs_s[k]), (r, k, obs_b[k][r], obs_s[k])
        b.select(r)
        a1, a2 = b.download(), e.download()
        for f in FIELDS:
            np.testing.assert_array_equal(a1[f], a2[f], err_msg=f"replica {r} {f}")
        c1, l1 = b.cell_lists(32)
        c2, l2 = e.cell_lists(32)
        np.testing.assert_array_equal(c1, c2)
        np.testing.assert_array_equal(l1, l2)
        e.close()
    b.close()


def test_batch_replica_matches_oracle():
    """The first replica of a batch against the oracle at the parity
    tolerance (Philox noise on both sides)."""
    from spilady_amd.engine import BatchEngine
    reps = _replicas(True)
    b = BatchEngine([p for p, _ in reps], [st for _, st in reps])
    b.step(0, 8)
    p, st = reps[1]
    o = orcpy.Oracle(p, st, compute_initial_forces=False, rng=orcpy.RNG_PHILOX)
    o.step(0, 8)
    b.select(1)
    x, y = o.download(), b.download()
    for f in ORACLE_FIELDS:
        assert_close(x[f], y[f], f"replica 1 {f}", TOL)
    o.close()
    b.close()


def test_batch_quantum_noise_matches_oracle():
    """The quantum thermostat in a batch: each replica's filter follows its
    own temperature (300, 900 and 1,400 K), and its coloured noise its own
    seed and stream id. M = 3 and Nf = 8, 8 steps (six redraws); every
    replica against an oracle run of it alone, at the parity tolerance."""
    from spilady_amd.engine import BatchEngine
    reps = []
    for r, (nu, T, seed) in enumerate(REPLICAS):
        v, p, at = fe_params(nu, T, seed, thermostats=True, replica=r)
        p.quantumnoise, p.msteps_quantum, p.nfrequency_quantum = 1, 3, 8
        reps.append((p, consistent_state(p, at)))
    b = BatchEngine([p for p, _ in reps], [st for _, st in reps])
    b.step(0, 8)
    for r, (p, st) in enumerate(reps):
        o = orcpy.Oracle(p, st, compute_initial_forces=False, rng=orcpy.RNG_PHILOX)
        o.step(0, 8)
        b.select(r)
        x, y = o.download(), b.download()
        for f in ORACLE_FIELDS:
            assert_close(x[f], y[f], f"replica {r} {f}", TOL)
        o.close()
    b.close()


def test_batch_local_volume_and_moments_equal_each_replica_alone():
    """localvol and magmom in a batch: each replica's volumes are normalised
    to its own box (one per-replica sum), and its moments convert to spins and
    back on its own slots; every replica equals its single-system run bit for
    bit, the local volumes included."""
    from spilady_amd import constants as K
    from spilady_amd.engine import BatchEngine, Engine
    reps = []
    for r, (nu, T, seed) in enumerate(REPLICAS):
        v, p, at = fe_params(nu, T, seed, thermostats=True, replica=r, displace=0.04)
        at.s = at.s * -K.EL_G
        p.localvol = 1
        p.rcut_vol = 2.7
        p.magmom = 1
        reps.append((p, consistent_state(p, at)))
    nsteps = 6
    b = BatchEngine([p for p, _ in reps], [st for _, st in reps])
    obs_b = b.step(0, nsteps)
    for r, (p, st) in enumerate(reps):
        e = Engine(p, st)
        obs_s = e.step(0, nsteps)
        for k in range(nsteps):
            assert _obs_equal(obs_b[k][r], obs_s[k]), (r, k)
        b.select(r)
        a1, a2 = b.download(), e.download()
        for f in FIELDS + ("local_volume",):
            np.testing.assert_array_equal(a1[f], a2[f], err_msg=f"replica {r} {f}")
        box = np.asarray(p.box[:])
        np.testing.assert_allclose(a1["local_volume"].sum(), box[0] * box[2] * box[5], rtol=1e-12)
        e.close()
    b.close()
