"""Shared set-up for the tests: seeded bcc Fe states made consistent by the
oracle, and comparison helpers with the tolerance written out."""
from __future__ import annotations

import numpy as np

from spilady_amd import abi, configs
from oracle import orcpy

# Normwise relative tolerance for floating-point parity with the oracle:
# max_i |a_i - b_i| <= TOL * max_i |a_i| per array (1e-10, the bound the
# thermostat-off comparison is held to).
TOL = 1e-10


def fe_params(nu: int, T: float, seed: int, *, thermostats: bool, eltemp: bool = False,
              localcolmot: bool = False, renormalize: bool = False, print_every: int = 1,
              step: float = 1e-15, tilt: float = 10.0, displace: float = 0.0, replica: int = 0):
    at = configs.make_bcc_fe(nu, T, seed, tilt_max_deg=tilt, displace_sigma=displace)
    v = configs.fe_variables(T, seed, thermostats=thermostats, eltemp=eltemp,
                             localcolmot=localcolmot, renormalize=renormalize,
                             print_every=print_every, step=step)
    p = abi.params_from_variables(v, at.box, at.natom, replica=replica)
    return v, p, at


def consistent_state(p, at):
    """Initial forces, density and energies from the oracle's own evaluation
    (the non-readconf path of initial.cpp:78-83), returned as a dict."""
    o = orcpy.Oracle(p, at, compute_initial_forces=True, rng=orcpy.RNG_PHILOX)
    st = o.download()
    o.close()
    return st


def relerr(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.max(np.abs(a))
    if scale == 0.0:
        return float(np.max(np.abs(b)))
    return float(np.max(np.abs(a - b)) / scale)


def assert_close(a, b, name: str, tol: float = TOL):
    e = relerr(a, b)
    assert e <= tol, f"{name}: normwise relative error {e:.3e} > {tol:.1e}"
    return e
