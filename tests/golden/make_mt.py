#!/usr/bin/env python3
"""Writes tests/golden/mt_fe1k.npz: the magnetisation curve M(T) of small
bcc Fe systems from the CPU restatement (oracle/liborc.so) with the
reference's own MWC noise stream, one single-threaded run per
(temperature, seed), run in parallel processes:

    python tests/golden/make_mt.py

Protocol, shared with tests/test_gpu_mt.py: 8^3 cubic cells (1,024 atoms),
Maxwell-Boltzmann velocities and spins tilted <= 10 deg from the seed,
Langevin lattice and spin thermostats at T (no electron temperature),
dt = 1 fs, 2,000 steps, observables every 10 steps. Per run: the means of
|<m>|, T_s,R and T_l over steps 1,000-2,000.
"""
from __future__ import annotations

import os
import sys
from concurrent.futures import ProcessPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

TEMPS = (100.0, 400.0, 700.0, 1000.0, 1300.0)
SEEDS = (301, 302, 303, 304)
NU, NSTEPS, EVERY, WINDOW0 = 8, 2000, 10, 1000


def run(args):
    T, seed = args
    from oracle import orcpy
    from spilady_amd import abi, configs
    at = configs.make_bcc_fe(NU, T, seed)
    v = configs.fe_variables(T, seed, thermostats=True, print_every=EVERY)
    p = abi.params_from_variables(v, at.box, at.natom)
    o = orcpy.Oracle(p, at, compute_initial_forces=True, rng=orcpy.RNG_MWC)
    obs = o.step(0, NSTEPS)
    o.close()
    w = [x for x in obs if x["step"] >= WINDOW0]
    return (T, seed, np.mean([np.linalg.norm(x["ave_m"]) for x in w]), np.mean([x["ts_r"] for x in w]),
            np.mean([x["tl"] for x in w]))


def main(path=os.path.join(HERE, "mt_fe1k.npz")):
    from oracle import orcpy
    orcpy.build()
    jobs = [(T, s) for T in TEMPS for s in SEEDS]
    with ProcessPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        res = list(ex.map(run, jobs))
    m = np.array([r[2] for r in res]).reshape(len(TEMPS), len(SEEDS))
    ts = np.array([r[3] for r in res]).reshape(len(TEMPS), len(SEEDS))
    tl = np.array([r[4] for r in res]).reshape(len(TEMPS), len(SEEDS))
    np.savez_compressed(path, temps=np.array(TEMPS), seeds=np.array(SEEDS), m=m, ts=ts, tl=tl,
                        nu=NU, nsteps=NSTEPS, every=EVERY, window0=WINDOW0)
    for k, T in enumerate(TEMPS):
        print(f"T {T:6.0f}  |M| {m[k].mean():.4f} +- {m[k].std(ddof=1):.4f}  Ts {ts[k].mean():7.1f}  "
              f"Tl {tl[k].mean():7.1f}")


if __name__ == "__main__":
    main()
