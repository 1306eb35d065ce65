"""Magnetisation curve M(T) of batched replicas against the CPU restatement.

Twenty replicas (five temperatures from 100 to 1,300 K, four seeds each) of
8^3 cubic cells run in one batched context with the counter-based noise; the
fixture tests/golden/mt_fe1k.npz holds the same protocol run by the oracle
with the reference's MWC noise and other seeds (tests/golden/make_mt.py).
The two noise streams are independent, so the comparison is statistical:
per temperature the seed-averaged window means of |<m>|, T_s,R and T_l must
agree within 4 standard errors of the difference (from the seed-to-seed
spread on both sides) plus 0.5 % of the value.
"""
import os

import numpy as np
import pytest

from spilady_amd import abi, configs

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GPU_SEEDS = (501, 502, 503, 504)


def test_magnetisation_curve_matches_cpu_restatement():
    from spilady_amd.engine import BatchEngine
    ref = np.load(os.path.join(HERE, "golden", "mt_fe1k.npz"))
    temps = ref["temps"]
    nu, nsteps, every, w0 = (int(ref[k]) for k in ("nu", "nsteps", "every", "window0"))
    plist, alist = [], []
    for T in temps:
        for s in GPU_SEEDS:
            at = configs.make_bcc_fe(nu, float(T), s)
            v = configs.fe_variables(float(T), s, thermostats=True, print_every=every)
            plist.append(abi.params_from_variables(v, at.box, at.natom, replica=len(plist)))
            alist.append(at)
    b = BatchEngine(plist, alist, compute_initial_forces=True)
    obs = b.step(0, nsteps)
    b.close()
    win = [o for o in obs if o[0]["step"] >= w0]
    assert len(win) == (nsteps - w0) // every
    R = len(plist)
    m = np.array([np.mean([np.linalg.norm(o[r]["ave_m"]) for o in win]) for r in range(R)])
    ts = np.array([np.mean([o[r]["ts_r"] for o in win]) for r in range(R)])
    tl = np.array([np.mean([o[r]["tl"] for o in win]) for r in range(R)])
    ns = len(GPU_SEEDS)
    rows = []
    for k, T in enumerate(temps):
        for name, g, c in (("M", m[k * ns:(k + 1) * ns], ref["m"][k]), ("Ts", ts[k * ns:(k + 1) * ns], ref["ts"][k]),
                           ("Tl", tl[k * ns:(k + 1) * ns], ref["tl"][k])):
            d = abs(g.mean() - c.mean())
            se = np.sqrt(g.var(ddof=1) / g.size + c.var(ddof=1) / c.size)
            bound = 4.0 * se + 0.005 * abs(c.mean())
            rows.append((T, name, g.mean(), c.mean(), se, d <= bound))
    for r in rows:
        print(f"T {r[0]:6.0f} {r[1]:>2s}  gpu {r[2]:10.4f}  cpu {r[3]:10.4f}  se {r[4]:.4f}  {'ok' if r[5] else 'FAIL'}")
    assert all(r[5] for r in rows), rows
