"""ctypes front end of oracle/liborc.so, the plain-C restatement of the
SPILADY SLDHL hot path (oracle/sld_oracle.c).

Test infrastructure only: the checker for the HIP engine. Only tests, the
smoke check and bench.py's cpu_baseline leg import it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Dict, Optional

import numpy as np

from spilady_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
ORC_LIB = os.path.join(HERE, "liborc.so")

RNG_MWC = 0
RNG_PHILOX = 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build(force: bool = False) -> str:
    """Compile oracle/sld_oracle.c (plain C, no FP contraction, so that the
    arithmetic order of the reference is kept)."""
    srcs = [os.path.join(HERE, f) for f in ("sld_oracle.c", "sld_model_oracle.c", "sld_oracle.h")]
    if force or not os.path.exists(ORC_LIB) or any(os.path.getmtime(ORC_LIB) < os.path.getmtime(f)
                                                   for f in srcs):
        subprocess.check_call(["make", "-s", "-C", HERE, "-f", "Makefile", "liborc.so"])
    return ORC_LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORC_LIB):
            raise FileNotFoundError(f"{ORC_LIB} not built (oracle/Makefile)")
        L = C.CDLL(ORC_LIB)
        P = C.POINTER(abi.SldParams)
        A = C.POINTER(abi.SldAtoms)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [P, A, C.c_int, C.c_int]
        L.orc_destroy.argtypes = [C.c_void_p]
        for nm in ("orc_links", "orc_calculate_rho", "orc_calculate_force_energy",
                   "orc_calculate_temperature"):
            getattr(L, nm).argtypes = [C.c_void_p]
        for nm in ("orc_core_ds", "orc_core_dp", "orc_core_dr", "orc_core_dTe"):
            getattr(L, nm).argtypes = [C.c_void_p, C.c_double]
        L.orc_check.argtypes = [C.c_void_p, C.c_int, C.POINTER(abi.SldObs)]
        L.orc_core.argtypes = [C.c_void_p, C.c_int]
        L.orc_step.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(abi.SldObs)]
        L.orc_step.restype = C.c_int
        L.orc_download.argtypes = [C.c_void_p, A]
        L.orc_upload.argtypes = [C.c_void_p, A]
        L.orc_get_cells.argtypes = [C.c_void_p, C.POINTER(abi.SldCells)]
        L.orc_set_cells_te.argtypes = [C.c_void_p, _dp]
        L.orc_get_geometry.argtypes = [C.c_void_p, _ip, _ip, _ip]
        L.orc_get_groups.argtypes = [C.c_void_p, _ip]
        L.orc_get_groups.restype = C.c_int
        L.orc_get_cell_lists.argtypes = [C.c_void_p, C.c_int, _ip, _ip]
        L.orc_get_cell_lists.restype = C.c_int
        L.orc_total_time.argtypes = [C.c_void_p]
        L.orc_total_time.restype = C.c_double
        L.orc_table.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.orc_table.restype = C.c_double
        L.orc_table_raw.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_table_raw.restype = C.c_double
        L.orc_analytic.argtypes = [C.c_int, C.c_double]
        L.orc_analytic.restype = C.c_double
        L.orc_philox4x32.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.orc_normal3.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_uint64, _dp]
        L.orc_scale_step.argtypes = [C.c_void_p]
        L.orc_set_barostat.argtypes = [C.c_void_p, C.POINTER(abi.SldBarostat)]
        L.orc_scale.argtypes = [C.c_void_p, C.c_int]
        L.orc_get_box.argtypes = [C.c_void_p, _dp]
        L.orc_quantum_H.argtypes = [C.c_void_p, _dp]
        L.orc_set_external.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_get_step.argtypes = [C.c_void_p]
        L.orc_get_step.restype = C.c_double
        L.orc_pair_v.argtypes = [C.c_double, _dp]
        L.orc_pair_v.restype = C.c_double
        L.orc_bridge_coeffs.argtypes = [_dp, _dp]
        L.orc_fields.argtypes = [C.c_void_p, _dp]
        L.orcm_default_params.argtypes = [_dp]
        L.orcm_gen.argtypes = [_dp, C.c_int, C.c_double]
        L.orcm_gen.restype = C.c_double
        L.orcm_dgen.argtypes = [_dp, C.c_int, C.c_int, C.c_double]
        L.orcm_dgen.restype = C.c_double
        L.orcm_build.argtypes = [_dp, C.c_double]
        L.orcm_build.restype = C.c_void_p
        L.orcm_free.argtypes = [C.c_void_p]
        L.orcm_table_raw.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.orcm_table_raw.restype = C.c_double
        L.orcm_eval_grad.argtypes = [C.c_void_p, _dp, C.c_int, _dp, _dp, C.c_double, C.c_double, C.c_double,
                                     C.c_double, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orcm_eval_grad.restype = C.c_int
        _lib = L
    return _lib


class Oracle(abi.EngineLike):
    """One restated SPILADY run (same surface as the HIP engine)."""

    def __init__(self, params: abi.SldParams, atoms, compute_initial_forces: bool = False,
                 rng: int = RNG_PHILOX):
        self.L = lib()
        self.params = params
        self._in = abi.AtomBuffers.from_state(atoms)
        self.h = self.L.orc_create(C.byref(params), C.byref(self._in.view()),
                                   int(compute_initial_forces), int(rng))
        if not self.h:
            raise RuntimeError("orc_create failed")
        self.natom = params.natom
        self._geometry()

    def _geometry(self):
        n3 = np.zeros(3, dtype=np.int32)
        ng = np.zeros(1, dtype=np.int32)
        self.L.orc_get_geometry(self.h, n3.ctypes.data_as(_ip), ng.ctypes.data_as(_ip), None)
        self.ncell3 = tuple(int(x) for x in n3)
        self.ncells = int(np.prod(n3))
        self.ngroups = int(ng[0])

    def close(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    __del__ = close

    # reference entry points
    def links(self): self.L.orc_links(self.h)
    def calculate_rho(self): self.L.orc_calculate_rho(self.h)
    def calculate_force_energy(self): self.L.orc_calculate_force_energy(self.h)
    def calculate_temperature(self): self.L.orc_calculate_temperature(self.h)
    def core_ds(self, dt): self.L.orc_core_ds(self.h, dt)
    def core_dp(self, dt): self.L.orc_core_dp(self.h, dt)
    def core_dr(self, dt): self.L.orc_core_dr(self.h, dt)
    def core_dTe(self, dt): self.L.orc_core_dTe(self.h, dt)
    def core(self, step): self.L.orc_core(self.h, step)

    def check(self, step: int) -> Dict[str, float]:
        ob = abi.SldObs()
        self.L.orc_check(self.h, step, C.byref(ob))
        return abi.obs_dict(ob)

    def step(self, first_step: int, nsteps: int):
        obs = (abi.SldObs * max(nsteps, 1))()
        k = self.L.orc_step(self.h, first_step, nsteps, obs)
        return [abi.obs_dict(obs[i]) for i in range(k)]

    def download(self) -> Dict[str, np.ndarray]:
        out = abi.AtomBuffers.empty(self.natom)
        self.L.orc_download(self.h, C.byref(out.view()))
        return out.as_dict()

    def upload(self, st: Dict[str, np.ndarray]):
        buf = abi.AtomBuffers.from_dict(st, self.natom)
        self.L.orc_upload(self.h, C.byref(buf.view()))

    def cells(self) -> Dict[str, np.ndarray]:
        out = abi.CellBuffers.empty(self.ncells)
        self.L.orc_get_cells(self.h, C.byref(out.view()))
        return out.as_dict()

    def set_cells_te(self, te):
        te = np.ascontiguousarray(te, dtype=np.float64)
        self.L.orc_set_cells_te(self.h, te.ctypes.data_as(_dp))

    def groups(self):
        sizes = np.zeros(self.ngroups, dtype=np.int32)
        self.L.orc_get_geometry(self.h, None, None, sizes.ctypes.data_as(_ip))
        members = np.zeros(self.ncells, dtype=np.int32)
        self.L.orc_get_groups(self.h, members.ctypes.data_as(_ip))
        return sizes, members

    def cell_lists(self, cap: int = 64):
        counts = np.zeros(self.ncells, dtype=np.int32)
        lists = np.full(self.ncells * cap, -1, dtype=np.int32)
        mx = self.L.orc_get_cell_lists(self.h, cap, counts.ctypes.data_as(_ip), lists.ctypes.data_as(_ip))
        if mx > cap:
            raise RuntimeError(f"cell holds {mx} atoms > cap {cap}")
        return counts, lists.reshape(self.ncells, cap)

    @property
    def total_time(self) -> float:
        return self.L.orc_total_time(self.h)

    @property
    def dt(self) -> float:
        return self.L.orc_get_step(self.h)

    def scale_step(self):
        self.L.orc_scale_step(self.h)

    def set_barostat(self, b: abi.SldBarostat):
        self._baro = b
        self.L.orc_set_barostat(self.h, C.byref(b))

    def scale(self, step: int):
        """scale(step): barostats at their intervals, then changestep."""
        self.L.orc_scale(self.h, int(step))

    def box(self) -> np.ndarray:
        b = np.zeros(6)
        self.L.orc_get_box(self.h, b.ctypes.data_as(_dp))
        return b

    def quantum_H(self) -> np.ndarray:
        """The quantum thermostat's filter (empty without quantumnoise)."""
        h = np.zeros(self.L.orc_quantum_H(self.h, None))
        if h.size:
            self.L.orc_quantum_H(self.h, h.ctypes.data_as(_dp))
        return h

    def set_external(self, hext=None, fext=None):
        """Per-atom external field (eV) and force (eV/A), [natom][3]."""
        h = None if hext is None else np.ascontiguousarray(hext, dtype=np.float64)
        f = None if fext is None else np.ascontiguousarray(fext, dtype=np.float64)
        self.L.orc_set_external(self.h, None if h is None else h.ctypes.data_as(_dp),
                                None if f is None else f.ctypes.data_as(_dp))

    def table(self, which: int, x: float) -> float:
        return self.L.orc_table(self.h, which, x)

    def table_raw(self, which: int, n: int) -> float:
        return self.L.orc_table_raw(self.h, which, n)

    def fields(self) -> np.ndarray:
        """Effective field of every atom at the current spins (no update)."""
        h = np.zeros((self.natom, 3))
        self.L.orc_fields(self.h, h.ctypes.data_as(_dp))
        return h


def analytic(which: int, x: float) -> float:
    return lib().orc_analytic(which, x)


def philox4x32(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32(c, k, o)
    return [int(x) for x in o]


def normal3(seed: int, replica: int, atom: int, call: int):
    out = np.zeros(3)
    lib().orc_normal3(seed, replica, atom, call, out.ctypes.data_as(_dp))
    return out


# ---------------------------------------------------------------------------
# parametric potential model (oracle/sld_model_oracle.c)

MODEL_EMBED, MODEL_DENS, MODEL_PAIR, MODEL_EXCH = 0, 1, 2, 3
NPARAM = 22


def _arr(x, n=None):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    if n is not None:
        a = a.reshape(n)
    return a


def model_default_params() -> np.ndarray:
    th = np.zeros(NPARAM)
    lib().orcm_default_params(th.ctypes.data_as(_dp))
    return th


def model_gen(theta, kind: int, x: float) -> float:
    th = _arr(theta, NPARAM)
    return lib().orcm_gen(th.ctypes.data_as(_dp), kind, x)


def model_dgen(theta, kind: int, k: int, x: float) -> float:
    th = _arr(theta, NPARAM)
    return lib().orcm_dgen(th.ctypes.data_as(_dp), kind, k, x)


def pair_v(r: float, V) -> float:
    v = _arr(V, 8)
    return lib().orc_pair_v(r, v.ctypes.data_as(_dp))


def bridge_coeffs(ends) -> np.ndarray:
    e = _arr(ends, 6)
    a = np.zeros(6)
    lib().orc_bridge_coeffs(e.ctypes.data_as(_dp), a.ctypes.data_as(_dp))
    return a


class ModelTables:
    """Tables of the model at theta and of their derivatives in theta."""

    def __init__(self, theta, rcut_max: float):
        self.theta = _arr(theta, NPARAM).copy()
        self.L = lib()
        self.h = self.L.orcm_build(self.theta.ctypes.data_as(_dp), float(rcut_max))

    def raw(self, kind: int, deriv: int, n: int) -> float:
        """kind >= 0: value table of that generator kind; kind = -1 - k: the
        derivative table of parameter k."""
        return self.L.orcm_table_raw(self.h, kind, deriv, n)

    def eval_grad(self, atoms, rcut_max: float, rcut_pot: float, rcut_mag: float,
                  gE: float = 0.0, gF=None, gH=None):
        """(E, F[n,3], H[n,3], grad[22]) of one configuration (cubic box)."""
        n = atoms.r.shape[0]
        Lb = np.array([atoms.box[0], atoms.box[2], atoms.box[5]], dtype=np.float64)
        r = _arr(atoms.r)
        s = _arr(atoms.s)
        F = np.zeros((n, 3))
        H = np.zeros((n, 3))
        E = np.zeros(1)
        g = np.zeros(NPARAM)
        pF = None if gF is None else _arr(gF, (n, 3))
        pH = None if gH is None else _arr(gH, (n, 3))
        rc = self.L.orcm_eval_grad(self.h, Lb.ctypes.data_as(_dp), n, r.ctypes.data_as(_dp),
                                   s.ctypes.data_as(_dp), rcut_max, rcut_pot, rcut_mag, float(gE),
                                   None if pF is None else pF.ctypes.data_as(_dp),
                                   None if pH is None else pH.ctypes.data_as(_dp),
                                   E.ctypes.data_as(_dp), F.ctypes.data_as(_dp), H.ctypes.data_as(_dp),
                                   g.ctypes.data_as(_dp))
        if rc != 0:
            raise ValueError("box shorter than twice rcut_max")
        return float(E[0]), F, H, g

    def close(self):
        if getattr(self, "h", None):
            self.L.orcm_free(self.h)
            self.h = None

    __del__ = close
