#!/usr/bin/env python3
"""Throughput of the spin-lattice time step on one MI355X.

Workload (production size, all couplings): bcc Fe, 100x100x100 cubic cells =
2,000,000 atoms, a = 2.8665 A, 600 K Maxwell-Boltzmann velocities from seed 7,
spins 2.2 muB along +z tilted by up to 10 degrees; DD05 magnetic EAM with Jij
Heisenberg exchange and Landau terms; Langevin lattice and spin thermostats,
electron-temperature diffusion on the link-cell mesh (Te0 = 600 K) with local
collective motion; renormalizeEnergy off (with it the electron temperature
of this start diverges within five steps in the reference itself);
dt = 1 fs, FP64. Observables are evaluated every step on the device.

A step is one full core() sequence plus the per-step observables. The timed
region starts with every input resident in HBM and ends after a device
synchronisation; `value` = atoms * steps / seconds.

Prints one JSON line (see the keys below). --gpus above 1 is not supported:
this engine runs one GPU per process and the run uses one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "atom-steps/s at 2M-atom bcc Fe SLD (FP64); batched-vs-serial replica speedup"
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0

# Algorithmic FP64 work per unit, counted from the reference formulas with
# transcendentals, square roots and divisions as one operation each:
# W_F = 192 per ordered pair (force), 26 pairs per atom in bcc Fe within
# rcut_pot, plus a 70-operation one-body epilogue; W_H = 31 and W_J = 33 per
# magnetic pair (14 per atom) and U = 300 for the Langevin SLDHL spin update
# (57 without thermostat); rho + neighbour build W_rho = 26 per pair.
PAIRS_POT = 26
PAIRS_MAG = 14
FLOP = {
    "k_force": PAIRS_POT * 192 + 70,                 # per atom
    "k_spin_phase": PAIRS_MAG * (31 + 33) + 300,     # per atom visit (4 visits per step)
    "k_nbuild": PAIRS_POT * 26 + PAIRS_MAG * 7,      # per atom
}
# Compulsory HBM bytes per unit for the same kernels (FP64 state read and
# written once; gathered neighbour data assumed on chip).
BYTES = {
    "k_force": 32 * 3 + 8 * 2 + 4 * PAIRS_POT + 4 + 32 + 8 * 4 + 8 * 6,
    "k_spin_phase": 32 + 8 + 4 * PAIRS_MAG + 8 * PAIRS_MAG + 4 + 32 * 3 + 8,
    "k_nbuild": 32 + 4 * PAIRS_POT + 12 * PAIRS_MAG + 8 + 8,
}
VISITS_PER_STEP = {"k_force": 1, "k_spin_phase": 4, "k_nbuild": 1}


def build_workload(n_unit: int):
    from spilady_amd import abi, configs
    at = configs.make_bcc_fe(n_unit, 600.0, 7)
    v = configs.fe_variables(600.0, 7, thermostats=True, eltemp=True, localcolmot=True,
                             renormalize=False, print_every=1, out_body="fe2m")
    p = abi.params_from_variables(v, at.box, at.natom)
    return v, p, at


def cpu_baseline(v, st_dict, box, natom, steps: int):
    """The compiled reference (oracle/_ref) on the same atoms, timed on the
    host's cores for a bounded number of steps; falls back to the
    single-threaded plain-C oracle when the reference library is absent."""
    from oracle import refpy
    threads = int(os.environ.get("SLD_CPU_THREADS", min(16, os.cpu_count() or 1)))
    vv = dict(v)
    vv["OMP_threads"] = threads
    if refpy.available():
        from spilady_amd.io import AtomState
        st = AtomState(box=np.asarray(box), r=st_dict["r"], p=st_dict["p"], f=st_dict["f"],
                       s=st_dict["s"], stress=st_dict["stress"], rho=st_dict["rho"],
                       ke=st_dict["ke"], pe=st_dict["pe"], me=st_dict["me"], me0=st_dict["me0"])
        d = tempfile.mkdtemp(prefix="sldref_bench_")
        ref = refpy.Reference()
        ref.init_memory(d, vv, st)
        t0 = time.perf_counter()
        ref.run(0, steps, False)
        dt = time.perf_counter() - t0
        ref.free()
        return {"value": natom * steps / dt, "unit": "atom-steps/s", "cores": threads,
                "kind": "reference",
                "sample": f"{steps} full SLDHL steps of the same {natom}-atom workload "
                          f"(compiled reference, OpenMP {threads} threads), {dt:.1f} s"}
    from oracle import orcpy
    from spilady_amd import abi
    p = abi.params_from_variables(v, box, natom)
    o = orcpy.Oracle(p, st_dict, compute_initial_forces=False, rng=orcpy.RNG_PHILOX)
    t0 = time.perf_counter()
    o.step(0, steps)
    dt = time.perf_counter() - t0
    o.close()
    return {"value": natom * steps / dt, "unit": "atom-steps/s", "cores": 1, "kind": "port",
            "sample": f"{steps} full steps of the same workload, single-threaded C oracle, {dt:.1f} s"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n-unit", type=int, default=100, help="cubic cells per side (100 = 2M atoms)")
    ap.add_argument("--timed-kernel", default="k_spin_phase")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: the spin-lattice engine is single-device in this build "
              "(see DESIGN.md for how the path would shard).", file=sys.stderr)
        return 2

    from spilady_amd.engine import Engine, device_info
    v, p, at = build_workload(args.n_unit)
    n = at.natom
    eng = Engine(p, at, compute_initial_forces=True)
    dev, ncu = device_info()
    eng.step(0, args.warmup)
    eng.time_kernel(args.timed_kernel)
    t0 = time.perf_counter()
    eng.step(args.warmup, args.steps)
    t1 = time.perf_counter()
    wall = t1 - t0
    dev_ms = eng.last_step_ms()
    launches, kms = eng.kernel_timing()
    eng.time_kernel(None)
    st = eng.download() if not args.no_cpu_baseline else None
    eng.close()

    value = n * args.steps / wall
    kname = args.timed_kernel
    roof = None
    if kname in FLOP and launches > 0:
        units = n * VISITS_PER_STEP[kname] * args.steps
        flops = FLOP[kname] * units
        tflops = flops / (kms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": tflops, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": tflops / PEAK_FP64_TFLOPS, "traffic": None,
                "kernel": kname, "launches": launches,
                "avg_launch_ms": kms / launches,
                "algorithmic_flop_per_launch": flops / launches,
                "hbm_gbs_algorithmic": BYTES[kname] * units / (kms * 1e-3) / 1e9,
                "kernel_share_of_step": kms / dev_ms if dev_ms > 0 else None}
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(v, st, at.box, n, args.cpu_steps)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "device_ms_per_step": dev_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"bcc Fe {args.n_unit}^3 cubic cells ({n} atoms), 600 K, Langevin "
                               "lattice+spin, eltemp (Te0 600 K), localcolmot, renormalizeEnergy off, "
                               "dt 1 fs, observables every step",
                   "natom": n, "device": dev, "cu": ncu},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
