"""Seeded bcc Fe workloads (the synthetic inputs of the spin-lattice runs).

Lattice: bcc [100] with the atom order and a/4 offset of bcc100.cpp:55-87
(for each z, y, x unit cell: the corner atom, then the body-centre atom, all
shifted by a/4 along each axis), so every atom sits at fractional coordinate
q in [0, 1) at least a/4 from a box face.

Velocities: p ~ N(0, m k_B T) per component, net momentum removed.
Spins: |s| = mu/g along +z (M = -g muB S points along -z), each tilted by
theta ~ U[0, theta_max] with a uniform azimuth.

Random numbers come from numpy's Philox bit generator keyed by the seed, so
every workload is reproducible from its seed alone.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict

import numpy as np

from . import constants as K
from .io import AtomState


def bcc100_positions(nx: int, ny: int, nz: int, a: float) -> np.ndarray:
    """Positions in the order of bcc100.cpp:55-87."""
    # loop order z (outer), y, x (inner), then the 2 basis atoms
    i = np.arange(nz)[:, None, None] * np.ones((1, ny, nx))
    j = np.arange(ny)[None, :, None] * np.ones((nz, 1, nx))
    k = np.arange(nx)[None, None, :] * np.ones((nz, ny, 1))
    base = np.stack([k.ravel(), j.ravel(), i.ravel()], axis=1) * a
    r = np.empty((2 * base.shape[0], 3))
    r[0::2] = base
    r[1::2] = base + a / 2.0
    r += a / 4.0
    return r


@dataclass
class Workload:
    """One run description: variables.in values plus the initial atoms."""
    variables: Dict[str, object]
    atoms: AtomState
    cells_per_side: int


def make_bcc_fe(n_unit: int, temperature_k: float, seed: int, *,
                a: float = K.FE_A_LATTICE, mag_mom: float = 2.2,
                tilt_max_deg: float = 10.0, displace_sigma: float = 0.0,
                mass_u: float = K.FE_MASS_U) -> AtomState:
    """bcc Fe block of n_unit^3 cubic cells (2 n_unit^3 atoms)."""
    r = bcc100_positions(n_unit, n_unit, n_unit, a)
    n = r.shape[0]
    L = n_unit * a
    rng = np.random.Generator(np.random.Philox(seed))
    if displace_sigma > 0.0:
        r = r + rng.normal(0.0, displace_sigma, size=r.shape)
        r = np.mod(r, L)
    m = K.mass_internal(mass_u)
    kT = temperature_k * K.BOLTZ
    p = rng.normal(0.0, np.sqrt(m * kT), size=(n, 3))
    p -= p.mean(axis=0, keepdims=True)
    s0 = mag_mom / K.EL_G
    theta = np.deg2rad(tilt_max_deg) * rng.random(n)
    phi = 2.0 * np.pi * rng.random(n)
    s = np.stack([s0 * np.sin(theta) * np.cos(phi),
                  s0 * np.sin(theta) * np.sin(phi),
                  s0 * np.cos(theta)], axis=1)
    box = np.array([L, 0.0, L, 0.0, 0.0, L])
    z = np.zeros(n)
    return AtomState(box=box, r=r, p=p, f=np.zeros((n, 3)), s=s,
                     stress=np.zeros((n, 6)), rho=z.copy(),
                     ke=np.sum(p * p, axis=1) / 2.0 / m, pe=z.copy(),
                     me=z.copy(), me0=z.copy(), heff_h=np.zeros((n, 3)),
                     heff_l=np.zeros((n, 3)), local_volume=z.copy())


def fe_variables(temperature_k: float, seed: int, *, step: float = 1e-15,
                 thermostats: bool = True, eltemp: bool = False,
                 localcolmot: bool = False, renormalize: bool = False,
                 nsteps: int = 1000, print_every: int = 1,
                 out_body: str = "fe") -> Dict[str, object]:
    """variables.in values for the Fe runs (variables.in:11-22 parameters)."""
    v = {
        "out_body": out_body,
        "no_of_production_steps": nsteps,
        "step": step,
        "interval_of_print_out": print_every,
        "interval_of_config_out": 100000000,
        "in_config": "con-input.dat",
        "kappa_e": 80.0,
        "start_time": 0.0,
        "atmass": K.FE_MASS_U,
        "temperature": float(temperature_k),
        "gamma_L_over_mass": 6e11 if thermostats else 0.0,
        "gamma_S_HL": 5.88e13 if thermostats else 0.0,
        "random_seed": int(seed),
        "rcut_pot": 4.1001,
        "rcut_mag": 3.7501,
        "rcut_max": 4.1001,
        "min_length_link_cell": 4.11,
        "OMP_threads": 1,
        "spinlang": int(thermostats),
        "lattlang": int(thermostats),
        "eltemp": int(eltemp),
        "localcolmot": int(localcolmot),
        "renormalizeEnergy": int(renormalize),
    }
    return v


def workload(name: str) -> Workload:
    """The named workloads: plain small (2,000 atoms, 300 K), plain NVE at
    scale (250,000 atoms), production (2,000,000 atoms, 600 K, all couplings)."""
    if name == "fe2k_langevin":
        at = make_bcc_fe(10, 300.0, 1)
        v = fe_variables(300.0, 1, nsteps=1000)
        return Workload(v, at, 6)
    if name == "fe250k_nve":
        at = make_bcc_fe(50, 300.0, 1)
        v = fe_variables(300.0, 1, thermostats=False, nsteps=500)
        return Workload(v, at, 34)
    if name == "fe2m_full":
        at = make_bcc_fe(100, 600.0, 7)
        v = fe_variables(600.0, 7, eltemp=True, localcolmot=True, renormalize=True,
                         nsteps=1000)
        return Workload(v, at, 69)
    raise KeyError(name)


BATCH_SIZES = (8, 12, 16, 20, 24, 28, 32)


def batch_replicas(nrep: int = 64, draw_seed: int = 3, *, nsteps: int = 2000):
    """The batched-replica workload: nrep independent bcc Fe replicas of L^3
    cubic cells with L drawn uniformly from {8,12,...,32} (a Philox-free
    numpy draw keyed by draw_seed), temperatures evenly spaced over
    100-1,400 K, replica r seeded 100 + r and given noise stream r, spins as
    in the small plain run, Langevin lattice and spin thermostats, dt = 1 fs.
    Returns [(variables, atoms)] in replica order."""
    rng = np.random.default_rng(draw_seed)
    sizes = rng.choice(list(BATCH_SIZES), size=nrep)
    temps = np.linspace(100.0, 1400.0, nrep)
    out = []
    for r in range(nrep):
        at = make_bcc_fe(int(sizes[r]), float(temps[r]), 100 + r)
        v = fe_variables(float(temps[r]), 100 + r, thermostats=True, print_every=1, nsteps=nsteps)
        out.append((v, at))
    return out
