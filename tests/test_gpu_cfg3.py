"""The batched-replica workload at its full size: 64 independent bcc Fe
replicas of L^3 cubic cells, L uniform in {8,12,...,32} (1,024 to 65,536
atoms each, 1,378,560 in all), temperatures evenly spaced over 100-1,400 K,
replica r seeded 100 + r, Langevin lattice and spin thermostats, dt = 1 fs,
all advanced in one batched context (configs.batch_replicas).

- After 3 steps the smallest (L = 8) and the largest (L = 32) replica are
  bitwise equal, in every observable of every step, every per-atom field and
  the cell lists, to the same replica run alone.
- After 3 steps the L = 32 replica (65,536 atoms, 22 link cells per side)
  matches the oracle at the parity tolerance, state and effective fields
  (the oracle draws the same counter-based noise).
- Over 2,000 steps the batch gives the magnetisation curve |<m>|(T), the mean
  over steps 1,000-2,000 per replica (check_spin_CPU.cpp:66-79). Three
  replicas (the coldest, the largest, and the one nearest 1,000 K where the
  magnetisation fluctuates most) run alone for the same 2,000 steps must give
  bitwise the same |<m>|, T_s,R and T_l at every step, and the curve must
  fall from the ordered state at 100 K to a disordered one at 1,400 K.
"""
import numpy as np
import pytest

import helpers
from helpers import TOL, assert_close
from oracle import orcpy
from spilady_amd import abi, configs

pytestmark = pytest.mark.gpu

NSTEPS, WINDOW0 = 2000, 1000
FIELDS = ("r", "p", "f", "s", "heff_h", "heff_l", "sum_jij_sj", "rho", "pe", "ke", "me", "me0", "stress")


@pytest.fixture(scope="module")
def replicas():
    reps = configs.batch_replicas(64, 3, nsteps=NSTEPS)
    out = []
    for r, (v, at) in enumerate(reps):
        out.append((abi.params_from_variables(v, at.box, at.natom, replica=r), at, float(v["temperature"])))
    return out


def _side(p):
    return int(round((p.natom / 2) ** (1.0 / 3.0)))


def _alone(p, at, nsteps):
    from spilady_amd.engine import Engine
    e = Engine(p, at, compute_initial_forces=True)
    obs = e.step(0, nsteps)
    return e, obs


def test_batch_first_steps_equal_alone_and_oracle(replicas):
    from spilady_amd.engine import BatchEngine
    sides = [_side(p) for p, _, _ in replicas]
    small, large = sides.index(8), sides.index(32)
    b = BatchEngine([p for p, _, _ in replicas], [a for _, a, _ in replicas], compute_initial_forces=True)
    assert b.nrep == 64
    ob = b.step(0, 3)
    for r in (small, large):
        p, at, T = replicas[r]
        print(f"replica {r}: L = {sides[r]}, {p.natom} atoms, {T:.1f} K", flush=True)
        e, os_ = _alone(p, at, 3)
        bad = []
        for k in range(3):
            bad += [f"step {k} obs {m}" for m in helpers.bitwise_diffs(ob[k][r], os_[k])]
        b.select(r)
        bad += [f"atoms {m}" for m in helpers.bitwise_diffs(b.download(), e.download(), FIELDS)]
        bad += [f"cells {m}" for m in helpers.bitwise_diffs(b.cells(), e.cells())]
        c1, l1 = b.cell_lists(32)
        c2, l2 = e.cell_lists(32)
        bad += [f"cell lists {m}" for m in helpers.bitwise_diffs({"n": c1, "l": l1}, {"n": c2, "l": l2})]
        e.close()
        assert not bad, f"replica {r}:\n" + "\n".join(bad)
    # the largest replica against the oracle
    p, at, T = replicas[large]
    o = orcpy.Oracle(p, at, compute_initial_forces=True, rng=orcpy.RNG_PHILOX)
    oo = o.step(0, 3)
    b.select(large)
    x, y = o.download(), b.download()
    for f in FIELDS:
        assert_close(x[f], y[f], f"L=32 replica {f}", TOL)
    for k in range(3):
        for key in ("ave_pe", "ave_ke", "ave_me", "ave_me0", "tl", "ts_r", "ts_l"):
            a_, g_ = oo[k][key], ob[k][large][key]
            assert abs(a_ - g_) <= TOL * max(abs(a_), 1e-300), (k, key, a_, g_)
        assert_close(oo[k]["ave_m"], ob[k][large]["ave_m"], f"L=32 replica <m> step {k}", TOL)
    o.close()
    b.close()


def _series(obs, r):
    return {"m": np.array([np.linalg.norm(o[r]["ave_m"]) for o in obs]),
            "ts_r": np.array([o[r]["ts_r"] for o in obs]),
            "tl": np.array([o[r]["tl"] for o in obs])}


def test_magnetisation_curve_2000_steps(replicas):
    from spilady_amd.engine import BatchEngine
    b = BatchEngine([p for p, _, _ in replicas], [a for _, a, _ in replicas], compute_initial_forces=True)
    obs = b.step(0, NSTEPS)
    b.close()
    assert len(obs) == NSTEPS
    temps = np.array([T for _, _, T in replicas])
    sides = np.array([_side(p) for p, _, _ in replicas])
    win = slice(WINDOW0, NSTEPS)
    curve = []
    for r in range(64):
        m = _series(obs, r)["m"][win]
        # standard error from 10 blocks of 100 steps
        blocks = m.reshape(10, -1).mean(axis=1)
        curve.append((temps[r], sides[r], m.mean(), blocks.std(ddof=1) / np.sqrt(blocks.size)))
    for T, L, m, se in curve:
        print(f"T {T:7.1f} K  L {L:2d}  |<m>| {m:.4f} +- {se:.4f} muB", flush=True)
    mm = np.array([c[2] for c in curve])
    # ordered at the cold end, disordered at the hot end (|<m>| of a
    # disordered finite system is the O(N^-1/2) residue)
    assert mm[0] > 1.9 and mm[-1] < 0.35, (mm[0], mm[-1])
    # three replicas run alone: bitwise the same series
    r_hot = int(np.argmin(np.abs(temps - 1000.0)))
    for r in sorted({0, int(np.argmax(sides)), r_hot}):
        p, at, T = replicas[r]
        e, os_ = _alone(p, at, NSTEPS)
        e.close()
        one = _series([[o] for o in os_], 0)
        bat = _series(obs, r)
        for k in one:
            assert np.array_equal(one[k], bat[k]), (r, k)
        print(f"replica {r} ({T:.1f} K, L {sides[r]}): {NSTEPS} steps alone equal the batch bitwise", flush=True)
