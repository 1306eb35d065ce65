"""The production configuration at its full size: bcc Fe, 100^3 cubic cells =
2,000,000 atoms, 600 K, seed 7, Langevin lattice and spin thermostats, the
electron temperature and localcolmot, dt = 1 fs. The oracle is too slow at
this size, so these checks are properties that hold at any size (the same
kernels are compared with the oracle element by element at small sizes):
- the link-cell lists partition the atoms, and after links() every atom sits
  in the cell its position bins to (links_CPU.cpp:96-112);
- the forces sum to zero (each pair's force is evaluated with exactly
  negated arguments on both sides, so only the summation rounds);
- core_dp's rescale leaves zero net momentum (core_dp_CPU.cpp:270-302);
- the observables are the means of the per-atom fields they summarise;
- eager launches and the per-step graph give bitwise the same trajectory."""
import numpy as np
import pytest

from spilady_amd import abi, configs

import helpers

pytestmark = pytest.mark.gpu

NU, NSTEPS = 100, 3


@pytest.fixture(scope="module")
def production():
    at = configs.make_bcc_fe(NU, 600.0, 7)
    v = configs.fe_variables(600.0, 7, thermostats=True, eltemp=True, localcolmot=True, renormalize=False)
    return at, abi.params_from_variables(v, at.box, at.natom)


def test_production_size_invariants(production):
    from spilady_amd.engine import Engine
    at, p = production
    n = at.natom
    e = Engine(p, at, compute_initial_forces=True)
    obs = e.step(0, NSTEPS)
    assert len(obs) == NSTEPS
    a = e.download()
    # forces: sum zero to rounding
    f = a["f"]
    assert np.linalg.norm(f.sum(axis=0)) <= 1e-12 * np.abs(f).sum()
    # net momentum zero after the rescale
    pm = a["p"]
    assert np.linalg.norm(pm.sum(axis=0)) <= 1e-12 * np.abs(pm).sum()
    # observables against the per-atom fields
    o = obs[-1]
    for key, fld in (("ave_pe", "pe"), ("ave_ke", "ke"), ("ave_me", "me")):
        assert o[key] == pytest.approx(a[fld].mean(), rel=1e-12, abs=1e-15), key
    np.testing.assert_allclose(o["ave_s"], a["s"].mean(axis=0), rtol=0, atol=1e-13)
    # link cells: a partition of the atoms, consistent with the positions
    e.links()
    a = e.download()
    counts, lists = e.cell_lists(32)
    members = lists[lists >= 0]
    assert counts.sum() == n and members.size == n
    assert np.array_equal(np.sort(members), np.arange(n))
    nx, ny, nz = e.ncell3
    box = np.asarray(at.box)
    q = a["r"] / np.array([box[0], box[2], box[5]])
    idx = np.clip(np.trunc(q * np.array([nx, ny, nz])).astype(np.int64), 0, np.array([nx, ny, nz]) - 1)
    cell_of = idx[:, 0] + idx[:, 1] * nx + idx[:, 2] * nx * ny
    owner = np.repeat(np.arange(e.ncells), counts)
    listed = np.concatenate([lists[c, :counts[c]] for c in range(e.ncells)])
    assert np.array_equal(cell_of[listed], owner)
    e.close()


ATOMS = ("r", "p", "f", "s", "heff_h", "heff_l", "rho", "pe", "ke", "me", "me0", "sum_jij_sj", "stress")


def _run(at, p, mode):
    from spilady_amd.engine import Engine
    e = Engine(p, at, compute_initial_forces=True)
    e.set_graph_mode(mode)
    obs = e.step(0, 2)
    a = {k: v for k, v in e.download().items() if k in ATOMS}
    cells = e.cells()
    e.close()
    return obs, a, cells


def _diffs(ref, other):
    (o0, a0, c0), (o1, a1, c1) = ref, other
    d = []
    for i, (x, y) in enumerate(zip(o0, o1)):
        d += [f"step {i} obs {m}" for m in helpers.bitwise_diffs(x, y)]
    d += [f"atoms {m}" for m in helpers.bitwise_diffs(a0, a1)]
    d += [f"cells {m}" for m in helpers.bitwise_diffs(c0, c1)]
    return d


def test_production_size_graph_replay_is_bitwise_eager(production):
    """Two full steps eagerly, then as per-step graphs, then eagerly again in
    a process where two 2M-atom engines have already been created and freed:
    every observable, every per-atom field and every cell field must be
    bitwise the same in all three. Each comparison reports every array that
    differs, not only the first.

    (A read-while-write race on the stored |s| payload made the thermostatted
    fields differ in their last bits between runs: the temperature pass at
    set-up rewrote |s| with the device's rounding while neighbouring cells'
    lanes read the host's. Now one device function alone writes |s|.)"""
    at, p = production
    eager = _run(at, p, 0)
    graph = _run(at, p, 2)
    eager2 = _run(at, p, 0)
    bad = [f"graph vs eager: {m}" for m in _diffs(eager, graph)]
    bad += [f"second eager vs eager: {m}" for m in _diffs(eager, eager2)]
    assert not bad, "\n".join(bad)


def test_production_size_local_volume_and_moments(production):
    """The production configuration with localvol (rcut_vol 2.7 A, between the
    first and second shells) and magmom on: the local volumes fill the box,
    every moment comes back as -g times its spin (each spin pass ends with
    m = -g s, core_ds_CPU.cpp:73-81), and the stress average of the last
    printed step equals the mean of the per-atom stress, in GPa."""
    from spilady_amd import constants as K
    from spilady_amd.engine import Engine
    at, p = production
    # moments -s: the spins s = m / (-g) are then 1/g of the plain run's, and
    # the g^2-scaled exchange (J g^2 |s|^2) keeps the plain run's energies.
    # Moments -g s put the Landau terms far from their minimum: at this size
    # the oracle's state turns non-finite in the second step.
    at = at.copy()
    at.s = -at.s
    p.localvol = 1
    p.rcut_vol = 2.7
    p.magmom = 1
    try:
        e = Engine(p, at, compute_initial_forces=True)
        obs = e.step(0, 2)
        a = e.download()
        e.close()
    finally:
        p.localvol = 0
        p.magmom = 0
    lv = a["local_volume"]
    box = np.asarray(at.box)
    assert np.all(np.isfinite(lv)) and lv.min() > 0.0
    assert lv.sum() == pytest.approx(box[0] * box[2] * box[5], rel=1e-12)
    m0 = np.linalg.norm(a["s"], axis=1)
    assert np.all(np.isfinite(m0)) and 0.3 < m0.mean() < 3.0
    # str- columns in GPa (check_stress_CPU.cpp:51-56)
    np.testing.assert_allclose(obs[-1]["ave_stress"], a["stress"].mean(axis=0) * 160.217653, rtol=1e-10,
                               atol=1e-12)
