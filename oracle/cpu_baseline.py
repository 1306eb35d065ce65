"""CPU baseline of the 2M-atom production step: the compiled SPILADY reference
(oracle/_ref/libspilady_ref.so, OpenMP) on the host's cores.

Test infrastructure only (the checker, timed beside the product): bench.py
starts this as a child process, so that the OpenMP runtime starts with the
binding set here and nothing else in the process (torch, the GPU runtime)
has started one first:

    python -m oracle.cpu_baseline --n-unit 100 --threads 16 --repeats 3

The workload is the bench's: bcc Fe, n_unit^3 cubic cells, 600 K, seed 7,
Langevin lattice and spin thermostats, eltemp and localcolmot, from a state
whose forces and density the reference computes itself. One untimed warm-up
step, then `repeats` timings of `steps` steps each; the median is reported.
Prints one JSON object.
"""
from __future__ import annotations

import argparse
import glob
import json
import os
import statistics
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def host_cpu():
    """CPU model, logical CPUs, the lease's affinity set and the physical
    cores (distinct package/core pairs) it spans."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    aff = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else list(range(os.cpu_count() or 1))
    cores = set()
    for cpu in aff:
        base = f"/sys/devices/system/cpu/cpu{cpu}/topology"
        try:
            with open(os.path.join(base, "physical_package_id")) as fh:
                pkg = fh.read().strip()
            with open(os.path.join(base, "core_id")) as fh:
                core = fh.read().strip()
            cores.add((pkg, core))
        except OSError:
            cores.add(("?", str(cpu)))
    return {"model": model, "logical_cpus": os.cpu_count(), "affinity_cpus": len(aff),
            "physical_cores_in_affinity": len(cores)}


def default_threads() -> int:
    """The physical cores of the lease, no more than OMP_NUM_THREADS if the
    host sets it."""
    info = host_cpu()
    n = info["physical_cores_in_affinity"]
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        n = min(n, int(env))
    return max(1, n)


def run(n_unit: int, threads: int, steps: int, repeats: int, host: dict) -> dict:
    sys.path.insert(0, ROOT)
    import numpy as np  # noqa: F401
    from oracle import refpy
    from spilady_amd import configs
    at = configs.make_bcc_fe(n_unit, 600.0, 7)
    v = configs.fe_variables(600.0, 7, thermostats=True, eltemp=True, localcolmot=True, renormalize=False,
                             print_every=1000000)
    v["OMP_threads"] = threads
    d = tempfile.mkdtemp(prefix="sld_cpu_baseline_")
    ref = refpy.Reference()
    ref.init_memory(d, v, at)
    ref.ref_calculate_rho()
    ref.ref_calculate_force_energy()
    ref.run(0, 1, False)  # warm-up
    times = []
    step0 = 1
    for _ in range(repeats):
        t0 = time.perf_counter()
        ref.run(step0, steps, False)
        times.append(time.perf_counter() - t0)
        step0 += steps
    ref.free()
    for f in glob.glob(os.path.join(d, "*")):
        os.remove(f)
    os.rmdir(d)
    med = statistics.median(times)
    return {"value": at.natom * steps / med, "unit": "atom-steps/s", "cores": threads, "kind": "reference",
            "seconds": times, "steps_per_repeat": steps, "natom": int(at.natom),
            "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES",
                                                   "OMP_SCHEDULE")},
            "host": host,
            "sample": f"compiled reference (OpenMP, {threads} threads bound close to cores), {at.natom} atoms: "
                      f"1 warm-up step, then the median of {repeats} timings of {steps} full SLDHL step(s)"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--n-unit", type=int, default=100)
    ap.add_argument("--threads", type=int, default=0, help="0: the physical cores of the lease")
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--repeats", type=int, default=3)
    a = ap.parse_args()
    # before any OpenMP runtime starts and binds this thread to one core
    host = host_cpu()
    threads = a.threads or default_threads()
    print(json.dumps(run(a.n_unit, threads, a.steps, a.repeats, host)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
