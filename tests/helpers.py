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


def ulp_distance(a, b) -> np.ndarray:
    """Distance in units in the last place between float64 arrays: the gap
    between their bit patterns on the monotone integer line of doubles."""
    ia = np.asarray(a, dtype=np.float64).view(np.int64)
    ib = np.asarray(b, dtype=np.float64).view(np.int64)
    # map sign-magnitude bit patterns onto one monotone integer scale
    # (exact integers: the patterns of nearby doubles differ in bits a
    # float64 difference would round away)
    ia = [(-(2 ** 63) - int(v)) if v < 0 else int(v) for v in ia.ravel()]
    ib = [(-(2 ** 63) - int(v)) if v < 0 else int(v) for v in ib.ravel()]
    return np.array([float(abs(x - y)) for x, y in zip(ia, ib)])


def bitwise_diffs(a, b, keys=None):
    """Every array of two dicts that is not bitwise equal, as one line each
    (elements that differ, largest ulp distance); empty when all are equal."""
    out = []
    for k in keys if keys is not None else sorted(a):
        x, y = np.asarray(a[k]), np.asarray(b[k])
        if x.shape != y.shape:
            out.append(f"{k}: shapes {x.shape} and {y.shape}")
            continue
        same = (x == y) | (np.isnan(x) & np.isnan(y)) if x.dtype.kind == "f" else (x == y)
        nd = int(np.size(same) - np.count_nonzero(same))
        if nd and x.dtype.kind == "f":
            u = float(np.max(ulp_distance(x[~same], y[~same])))
            out.append(f"{k}: {nd} of {x.size} differ, largest distance {u:.0f} ulp")
        elif nd:
            u = int(np.max(np.abs(x[~same].astype(np.int64) - y[~same].astype(np.int64))))
            out.append(f"{k}: {nd} of {x.size} differ, largest difference {u}")
    return out
