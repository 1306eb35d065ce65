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
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "atom-steps/s at 2M-atom bcc Fe SLD (FP64); batched-vs-serial replica speedup"
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0

# Algorithmic work per unit, counted from the reference formulas with
# transcendentals, square roots and divisions as one operation each, and
# compulsory HBM bytes per unit (state read and written once, gathered
# neighbour data assumed on chip). bcc Fe: 26 neighbours within rcut_pot and
# 14 within rcut_mag per atom.
PAIRS_POT = 26
PAIRS_MAG = 14
NLOOP_DTE = 50          # core_dTe sub-steps per call (core_dTe_CPU.cpp:51)
KERNELS = {
    # one cell update of the diffusion sweep: 3 x (2 add, 1 mul, 1 sub, 1 div)
    # + 3 add + 1 mul, then T(E) = acosh(exp(bE/3))/b (4); reads Te, Ee of the
    # cell and writes both (the six neighbour Te come from cache)
    "k_dte_sweep": {"unit": "cell update", "flop": 23, "bytes": 32},
    # one atom visit of the spin sweep: field W_H + W_J = 64 per magnetic pair,
    # 300 for the Langevin SLDHL update; reads s, rho, the 14 neighbour ids
    # and J values, writes s, H_eff, H_L and sum J|s_j|
    "k_spin_phase": {"unit": "atom visit", "flop": PAIRS_MAG * (31 + 33) + 300,
                     "bytes": 32 + 8 + 4 + PAIRS_MAG * (4 + 8) + 3 * 32 + 8},
    # one atom of the force pass: W_F = 192 per pair + 70; reads r, s, rho,
    # F'(rho) and the neighbour ids, writes f, pe, me, me0, vir, stress[6]
    "k_force": {"unit": "atom", "flop": PAIRS_POT * 192 + 70,
                "bytes": 3 * 32 + 2 * 8 + 4 + PAIRS_POT * 4 + 32 + 4 * 8 + 6 * 8},
}
# whole step: FLOP and bytes per atom-step of the full SLDHL step with
# eltemp (electron cell state spread over 6.09 atoms per cell)
STEP_FLOP = 14.6e3
STEP_BYTES = 381.0


def traffic_of(kname: str, key: str = "bytes_per_launch"):
    """HBM bytes per launch of a kernel from the committed counter summary
    (rocprofv3 --pmc, one pass each for FETCH_SIZE, WRITE_SIZE and the L2-to-
    memory request counts). "bytes_per_launch": 128 B per read request (every
    request moves one 128-byte line on MI355X; FETCH_SIZE counts 64 B for
    each, so this is 2 x FETCH_SIZE) plus WRITE_SIZE; "raw_bytes_per_launch":
    FETCH_SIZE + WRITE_SIZE as rocprofv3 reports them. None when the summary
    does not cover the kernel."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as fh:
            t = json.load(fh)
        return t["kernels"][kname][key]
    except (OSError, KeyError, ValueError):
        return None


def units_per_step(kname: str, natom: int, ncell: int) -> int:
    if kname == "k_dte_sweep":
        return 2 * ncell * 2 * NLOOP_DTE   # two core_dTe calls, forward + backward sweep
    if kname == "k_spin_phase":
        return 4 * natom                   # two core_ds calls, forward + backward sweep
    return natom


def build_workload(n_unit: int):
    from spilady_amd import abi, configs
    at = configs.make_bcc_fe(n_unit, 600.0, 7)
    v = configs.fe_variables(600.0, 7, thermostats=True, eltemp=True, localcolmot=True,
                             renormalize=False, print_every=1, out_body="fe2m")
    p = abi.params_from_variables(v, at.box, at.natom)
    return v, p, at


def nve_drift(nsteps: int, every: int = 100):
    """Energy conservation of the microcanonical step: 50^3 cubic cells =
    250,000 atoms, 300 K, seed 1, thermostats off, dt = 1 fs. E = pe + ke + me
    per atom (the enr- columns) every `every` steps."""
    from spilady_amd import abi, configs
    from spilady_amd.engine import Engine
    at = configs.make_bcc_fe(50, 300.0, 1)
    v = configs.fe_variables(300.0, 1, thermostats=False, print_every=every)
    p = abi.params_from_variables(v, at.box, at.natom)
    eng = Engine(p, at, compute_initial_forces=True)
    o0 = eng.check(-1)
    t0 = time.perf_counter()
    obs = eng.step(0, nsteps)
    wall = time.perf_counter() - t0
    eng.close()
    e = np.array([o0["ave_pe"] + o0["ave_ke"] + o0["ave_me"]] +
                 [o["ave_pe"] + o["ave_ke"] + o["ave_me"] for o in obs])
    t = np.array([0.0] + [o["total_time"] for o in obs]) * 1e12     # ps
    slope = float(np.polyfit(t, e, 1)[0])
    # after the first picosecond, once the initial swing of the splitting
    # error has settled
    late = t >= 1.0 - 1e-9
    fit = np.polyfit(t[late], e[late], 1) if late.sum() >= 3 else None
    return {"metric": "NVE energy drift, 250k-atom bcc Fe SLD (FP64)", "steps": nsteps,
            "dt_fs": p.step * 1e15, "e0_ev_per_atom": float(e[0]), "e_end_ev_per_atom": float(e[-1]),
            "max_abs_dev_ev_per_atom": float(np.max(np.abs(e - e[0]))),
            "drift_ev_per_atom_per_ps": slope,
            "drift_after_1ps_ev_per_atom_per_ps": None if fit is None else float(fit[0]),
            "std_about_fit_after_1ps_ev_per_atom":
                None if fit is None else float(np.std(e[late] - np.polyval(fit, t[late]))),
            "ke_ev_per_atom": float(obs[-1]["ave_ke"]),
            "ms_per_step": wall / nsteps * 1e3, "atom_steps_per_s": at.natom * nsteps / wall,
            "series_every": every, "series": e.tolist()}


def batch_workload(nrep: int = 64, seed: int = 3):
    """The 64-replica batched workload (configs.batch_replicas) as SldParams
    with replica r on noise stream r."""
    from spilady_amd import abi, configs
    return [(abi.params_from_variables(v, at.box, at.natom, replica=r), at)
            for r, (v, at) in enumerate(configs.batch_replicas(nrep, seed))]


def adjoint_leg(reps: int):
    """Loss gradient of the potential-fitting workload: 32 frozen bcc Fe
    configurations of 16^3 cubic cells (262,144 atoms), MSE of energies,
    forces and fields against targets at 2%-perturbed parameters. One
    gradient = sld_model_eval + sld_model_vjp on device buffers (the table
    build of sld_model_set_params runs once per parameter vector and is
    reported apart)."""
    import torch
    from spilady_amd import fit
    confs = fit.frozen_configs()
    natoms = np.array([c.r.shape[0] for c in confs], dtype=np.float64)
    m = fit.PotentialModel(confs)
    th0 = fit.default_params()
    t0 = time.perf_counter()
    m.set_params(fit.perturbed_params(th0, 0.02, 999))
    t_set = time.perf_counter() - t0
    dev = torch.device("cuda")
    targets = tuple(torch.tensor(x, device=dev) for x in m.eval())
    n_t = torch.tensor(natoms, device=dev)
    m.set_params(th0)
    E = torch.empty(m.nconf, dtype=torch.float64, device=dev)
    F = torch.empty((m.ntot, 3), dtype=torch.float64, device=dev)
    H = torch.empty_like(F)
    gE = torch.empty_like(E)
    gF = torch.empty_like(F)
    gH = torch.empty_like(F)
    grad = torch.empty(fit.NPARAM, dtype=torch.float64, device=dev)
    L, h = m.L, m.h

    def fwd():
        m.engine._rc(L.sld_model_eval(h, E.data_ptr(), F.data_ptr(), H.data_ptr()))

    def one():
        fwd()
        Et, Ft, Ht = targets
        gE.copy_(2.0 * (E - Et) / (n_t * n_t) / E.numel())
        gF.copy_(2.0 * (F - Ft) / F.numel())
        gH.copy_(2.0 * (H - Ht) / H.numel())
        torch.cuda.synchronize()
        m.engine._rc(L.sld_model_vjp(h, gE.data_ptr(), gF.data_ptr(), gH.data_ptr(), grad.data_ptr()))

    one()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        one()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / reps
    kern = {}
    # device time of each kernel from HIP events on the library's stream: the
    # timing covers one entry-point call (k_model_rho runs in both)
    for kname, call in (("k_model_rho", fwd), ("k_model_forward", fwd), ("k_model_q", one),
                        ("k_model_adjoint", one)):
        m.engine.time_kernel(kname)
        call()
        launches, kms = m.engine.kernel_timing()
        m.engine.time_kernel(None)
        kern[kname] = kms / max(launches, 1) * 1e3
    m.close()
    ntot = int(natoms.sum())
    return {"workload": "32 frozen bcc Fe configurations of 16^3 cubic cells (262,144 atoms), sigma 0.08 A, "
                        "spin tilts <= 30 deg, seeds 200-231; MSE loss of E, F, H",
            "atoms": ntot, "ms_per_loss_gradient": wall * 1e3, "atoms_per_s": ntot / wall,
            "set_params_s": t_set, "kernel_us": kern}


def batched_vs_serial(steps: int, warmup: int):
    """The same replicas advanced in one batched context and one after the
    other in single-system contexts; steps timed after warmup in both."""
    from spilady_amd.engine import BatchEngine, Engine
    reps = batch_workload()
    natoms = sum(p.natom for p, _ in reps)
    b = BatchEngine([p for p, _ in reps], [a for _, a in reps], compute_initial_forces=True)
    b.step(0, warmup)
    # untuned batched rate, then the autotuner (state restored), then the
    # timed tuned run that the serial one is compared with
    t0 = time.perf_counter()
    b.step(warmup, steps)
    t_default = time.perf_counter() - t0
    default_shapes = b.shapes()
    tuned, table = b.autotune(2)
    t0 = time.perf_counter()
    ob = b.step(warmup + steps, steps)
    tb = time.perf_counter() - t0
    # the same steps replayed from one graph per step
    b.set_graph_mode(2)
    b.step(warmup + 2 * steps, 1)
    t0 = time.perf_counter()
    b.step(warmup + 2 * steps + 1, steps)
    t_graph = time.perf_counter() - t0
    b.close()
    ts = 0.0
    equal = True
    for r, (p, a) in enumerate(reps):
        e = Engine(p, a, compute_initial_forces=True)
        e.step(0, warmup + steps)
        t0 = time.perf_counter()
        os_ = e.step(warmup + steps, steps)
        ts += time.perf_counter() - t0
        e.close()
        equal = equal and all(np.array_equal(ob[k][r][x], os_[k][x]) for k in range(steps) for x in os_[k])
    return {"replicas": len(reps), "atoms": natoms, "steps": steps,
            "batched_atom_steps_per_s": natoms * steps / tb,
            "serial_atom_steps_per_s": natoms * steps / ts,
            "speedup": ts / tb, "observables_bitwise_equal": equal,
            "untuned_batched_atom_steps_per_s": natoms * steps / t_default,
            # like for like: the serial contexts run the default shapes too
            "speedup_default_shapes": ts / t_default,
            "step_graph_batched_atom_steps_per_s": natoms * steps / t_graph,
            "autotune": {"default": default_shapes, "chosen": tuned, "us_per_step": table},
            "workload": "64 bcc Fe replicas, L^3 cubic cells, L uniform in {8..32}, 100-1400 K, "
                        "seed 100+r, Langevin lattice+spin, dt 1 fs"}


def cpu_baseline(v, st_dict, box, natom, steps: int, n_unit: int, repeats: int = 3):
    """The compiled reference (oracle/_ref) on the same workload, timed on the
    host's cores in a child process (oracle/cpu_baseline.py) whose OpenMP
    runtime starts bound one thread per physical core of the lease: one
    warm-up step, then the median of `repeats` timings of `steps` steps.
    Falls back to the single-threaded plain-C oracle in this process when
    the reference library is absent."""
    from oracle import refpy
    if refpy.available():
        from oracle import cpu_baseline as cb
        threads = int(os.environ.get("SLD_CPU_THREADS", "0")) or cb.default_threads()
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores",
                   OMP_SCHEDULE="static")
        out = subprocess.run([sys.executable, "-m", "oracle.cpu_baseline", "--n-unit", str(n_unit),
                              "--threads", str(threads), "--steps", str(steps), "--repeats", str(repeats)],
                             cwd=ROOT, env=env, capture_output=True, text=True, timeout=1800)
        if out.returncode != 0:
            raise RuntimeError(f"cpu baseline failed: {out.stderr[-2000:]}")
        return json.loads(out.stdout.strip().splitlines()[-1])
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
    ap.add_argument("--kernels", default=",".join(KERNELS),
                    help="kernels timed with HIP events in extra steps after the timed region")
    ap.add_argument("--cpu-steps", type=int, default=1, help="steps per timed repeat of the CPU baseline")
    ap.add_argument("--cpu-repeats", type=int, default=3, help="timed repeats of the CPU baseline (median)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--nve-drift", type=int, default=0,
                    help="instead of the benchmark, measure NVE energy drift over this many steps")
    ap.add_argument("--batch-steps", type=int, default=20,
                    help="timed steps of the batched-vs-serial replica comparison (0: skip)")
    ap.add_argument("--graph-steps", type=int, default=5,
                    help="steps timed per graph mode after the timed region")
    ap.add_argument("--autotune-steps", type=int, default=2,
                    help="steps per candidate launch shape in the autotuner before timing (0: default shapes)")
    ap.add_argument("--adjoint-reps", type=int, default=10,
                    help="timed loss gradients of the potential-fitting workload (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: the spin-lattice engine is single-device in this build "
              "(see DESIGN.md for how the path would shard).", file=sys.stderr)
        return 2
    if args.nve_drift > 0:
        print(json.dumps(nve_drift(args.nve_drift)), flush=True)
        return 0

    from spilady_amd.engine import Engine, device_info
    v, p, at = build_workload(args.n_unit)
    n = at.natom
    eng = Engine(p, at, compute_initial_forces=True)
    dev, ncu = device_info()
    eng.step(0, args.warmup)
    tune = None
    if args.autotune_steps > 0:
        default_shapes = eng.shapes()
        chosen, table = eng.autotune(args.autotune_steps)
        tune = {"default": default_shapes, "chosen": chosen, "us_per_step": table}
    t0 = time.perf_counter()
    eng.step(args.warmup, args.steps)
    t1 = time.perf_counter()
    wall = t1 - t0
    dev_ms = eng.last_step_ms()
    # per-kernel durations from HIP events on the engine's stream, one extra
    # step per kernel (outside the timed region)
    kern = {}
    s0 = args.warmup + args.steps
    for k, kname in enumerate(x for x in args.kernels.split(",") if x):
        eng.time_kernel(kname)
        eng.step(s0 + k, 1)
        launches, kms = eng.kernel_timing()
        eng.time_kernel(None)
        step_ms = eng.last_step_ms()
        if launches == 0 or kname not in KERNELS:
            continue
        spec = KERNELS[kname]
        units = units_per_step(kname, n, eng.ncells)
        per_launch = units / launches
        avg_s = kms / launches * 1e-3
        flop_rate = spec["flop"] * per_launch / avg_s / 1e12
        byte_rate = spec["bytes"] * per_launch / avg_s / 1e9
        intensity = spec["flop"] / spec["bytes"]
        hbm_bound = intensity < PEAK_FP64_TFLOPS * 1e3 / PEAK_HBM_GBS
        kern[kname] = {
            "bound": "hbm" if hbm_bound else "fp64",
            "achieved": byte_rate if hbm_bound else flop_rate,
            "peak": PEAK_HBM_GBS if hbm_bound else PEAK_FP64_TFLOPS,
            "unit": "GB/s" if hbm_bound else "TFLOP/s",
            "frac": byte_rate / PEAK_HBM_GBS if hbm_bound else flop_rate / PEAK_FP64_TFLOPS,
            "launches_per_step": launches, "avg_launch_us": kms / launches * 1e3,
            "units_per_launch": per_launch, "unit_of_work": spec["unit"],
            "algorithmic_bytes_per_launch": spec["bytes"] * per_launch,
            "algorithmic_flop_per_launch": spec["flop"] * per_launch,
            "share_of_step": kms / step_ms if step_ms > 0 else None,
        }
    # the same step issued eagerly (the default above), with the replayed
    # core_dTe graph and as one replayed graph per step
    graph_ms = {}
    s1 = s0 + len([x for x in args.kernels.split(",") if x])
    for mode in ((0, 1, 2) if args.graph_steps > 0 else ()):
        eng.set_graph_mode(mode)
        eng.step(s1, 1)
        t0 = time.perf_counter()
        eng.step(s1 + 1, args.graph_steps)
        graph_ms[str(mode)] = (time.perf_counter() - t0) / args.graph_steps * 1e3
        s1 += 1 + args.graph_steps
    eng.set_graph_mode(0)
    st = eng.download() if not args.no_cpu_baseline else None
    eng.close()

    value = n * args.steps / wall
    roof = None
    if kern:
        dom = max(kern, key=lambda k: kern[k]["share_of_step"] or 0.0)
        r = kern[dom]
        roof = {"bound": r["bound"], "achieved": r["achieved"], "peak": r["peak"], "unit": r["unit"],
                "frac": r["frac"], "traffic": traffic_of(dom),
                "traffic_raw": traffic_of(dom, "raw_bytes_per_launch"), "kernel": dom,
                "avg_launch_us": r["avg_launch_us"], "launches_per_step": r["launches_per_step"],
                "share_of_step": r["share_of_step"]}
    step_roof = {"flop_per_atom_step": STEP_FLOP, "bytes_per_atom_step": STEP_BYTES,
                 "fp64_tflops": value * STEP_FLOP / 1e12,
                 "fp64_frac": value * STEP_FLOP / 1e12 / PEAK_FP64_TFLOPS,
                 "hbm_gbs": value * STEP_BYTES / 1e9,
                 "hbm_frac": value * STEP_BYTES / 1e9 / PEAK_HBM_GBS}
    batched = batched_vs_serial(args.batch_steps, 2) if args.batch_steps > 0 else None
    adjoint = adjoint_leg(args.adjoint_reps) if args.adjoint_reps > 0 else None
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(v, st, at.box, n, args.cpu_steps, args.n_unit, args.cpu_repeats)
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
        "kernels": kern,
        "step_roofline": step_roof,
        "autotune": tune,
        "ms_per_step_by_graph_mode": graph_ms,
        "batched": batched,
        "adjoint": adjoint,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
