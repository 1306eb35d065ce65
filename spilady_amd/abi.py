"""ctypes mirror of include/spilady_hip.h (plain-data structs only) and
helpers that move SPILADY state between numpy arrays and the C ABI."""
from __future__ import annotations

import ctypes as C
from typing import Dict

import numpy as np

from . import constants as K

_dp = C.POINTER(C.c_double)


# SLD_ABI_VERSION of include/spilady_hip.h, whose struct layouts this module mirrors
ABI_VERSION = 2

class SldParams(C.Structure):
    _fields_ = [
        ("natom", C.c_int32),
        ("interval_of_print_out", C.c_int32),
        ("box", C.c_double * 6),
        ("step", C.c_double),
        ("temperature", C.c_double),
        ("atmass", C.c_double),
        ("gamma_L_over_mass", C.c_double),
        ("gamma_S_HL", C.c_double),
        ("kappa_e", C.c_double),
        ("rcut_pot", C.c_double),
        ("rcut_mag", C.c_double),
        ("rcut_max", C.c_double),
        ("min_length_link_cell", C.c_double),
        ("start_time", C.c_double),
        ("random_seed", C.c_uint64),
        ("replica", C.c_int32),
        ("spinlang", C.c_int32),
        ("lattlang", C.c_int32),
        ("eltemp", C.c_int32),
        ("localcolmot", C.c_int32),
        ("renormalize_energy", C.c_int32),
        ("changestep", C.c_int32),
        ("localvol", C.c_int32),
        ("displace_limit", C.c_double),
        ("phi_limit", C.c_double),
        ("extfield", C.c_int32),
        ("extforce", C.c_int32),
        ("quantumnoise", C.c_int32),
        ("msteps_quantum", C.c_int32),
        ("nfrequency_quantum", C.c_int32),
        ("magmom", C.c_int32),
        ("rcut_vol", C.c_double),
    ]


class SldAtoms(C.Structure):
    _fields_ = [(k, _dp) for k in ("r", "p", "f", "s", "heff_h", "heff_l", "rho", "pe", "ke",
                                   "me", "me0", "stress", "vir", "sum_jij_sj", "local_volume")]


class SldObs(C.Structure):
    _fields_ = [
        ("step", C.c_int32), ("pad", C.c_int32), ("total_time", C.c_double),
        ("ave_pe", C.c_double), ("ave_ke", C.c_double), ("ave_me", C.c_double),
        ("ave_me0", C.c_double), ("ave_ee", C.c_double),
        ("tl", C.c_double), ("ts_r", C.c_double), ("ts_l", C.c_double), ("te", C.c_double),
        ("ave_s", C.c_double * 3), ("ave_s0", C.c_double), ("ave_m", C.c_double * 3),
        ("ave_m0", C.c_double), ("virial", C.c_double),
        ("ave_stress", C.c_double * 6), ("pressure", C.c_double),
    ]


class SldCells(C.Structure):
    _fields_ = [(k, _dp) for k in ("te", "ee", "tl", "ts_r", "ts_l", "gel", "ges", "sum_r_up",
                                   "sum_r_dn", "sum_l_up", "sum_l_dn")]


class SldShapes(C.Structure):
    _fields_ = [(k, C.c_int32) for k in ("spin_block", "sweep_block", "force_block", "nbuild_cells",
                                         "temperature_block", "sweep_tail", "spin_lanes")]


SHAPE_CANDIDATES = {"spin_block": (32, 64, 128), "sweep_block": (64, 128, 256),
                    "force_block": (64, 128, 256), "nbuild_cells": (1, 2, 4),
                    "temperature_block": (64, 128, 256), "sweep_tail": (0, 2048, 8192),
                    "spin_lanes": (1, 2, 4)}


class SldBarostat(C.Structure):
    _fields_ = [("stress_on", C.c_int32), ("pressure_on", C.c_int32),
                ("interval_of_scale_stress", C.c_int32), ("interval_of_scale_pressure", C.c_int32),
                ("stress", C.c_double * 6), ("pressure", C.c_double), ("baro_damping_time", C.c_double)]


STRESS_KEYS = ("stress_xx", "stress_yx", "stress_yy", "stress_zx", "stress_zy", "stress_zz")


def barostat_from_variables(v: Dict[str, object]) -> SldBarostat:
    """The STRESS / PRESSURE switches and their variables.in values
    (read_variables.cpp:196-214; GPa and seconds, no conversion)."""
    b = SldBarostat()
    b.stress_on = int(v.get("STRESS", 0))
    b.pressure_on = int(v.get("PRESSURE", 0))
    b.interval_of_scale_stress = int(v.get("interval_of_scale_stress", 1))
    b.interval_of_scale_pressure = int(v.get("interval_of_scale_pressure", 1))
    for k, key in enumerate(STRESS_KEYS):
        b.stress[k] = float(v.get(key, 0.0))
    b.pressure = float(v.get("pressure", 0.0))
    b.baro_damping_time = float(v.get("baro_damping_time", 1e-13))
    return b


VEC_FIELDS = ("r", "p", "f", "s", "heff_h", "heff_l")
SCALAR_FIELDS = ("rho", "pe", "ke", "me", "me0", "vir", "sum_jij_sj", "local_volume")
CELL_FIELDS = ("te", "ee", "tl", "ts_r", "ts_l", "gel", "ges", "sum_r_up", "sum_r_dn",
               "sum_l_up", "sum_l_dn")


def obs_dict(ob: SldObs) -> Dict[str, float]:
    d = {k: getattr(ob, k) for k, _ in SldObs._fields_ if k not in ("pad", "ave_s", "ave_m", "ave_stress")}
    d["ave_s"] = np.array(ob.ave_s[:])
    d["ave_m"] = np.array(ob.ave_m[:])
    d["ave_stress"] = np.array(ob.ave_stress[:])
    return d


class AtomBuffers:
    """Owned contiguous float64 arrays plus an SldAtoms view of them."""

    def __init__(self, arrays: Dict[str, np.ndarray]):
        self.a = arrays

    @classmethod
    def empty(cls, n: int) -> "AtomBuffers":
        a = {k: np.zeros((n, 3)) for k in VEC_FIELDS}
        a.update({k: np.zeros(n) for k in SCALAR_FIELDS})
        a["stress"] = np.zeros((n, 6))
        return cls(a)

    @classmethod
    def from_dict(cls, d: Dict[str, np.ndarray], n: int) -> "AtomBuffers":
        a = {}
        for k, v in d.items():
            if v is None or k not in VEC_FIELDS + SCALAR_FIELDS + ("stress",):
                continue
            a[k] = np.ascontiguousarray(v, dtype=np.float64)
        return cls(a)

    @classmethod
    def from_state(cls, st) -> "AtomBuffers":
        """From an io.AtomState or a dict of arrays."""
        if isinstance(st, dict):
            return cls.from_dict(st, len(st["r"]))
        d = {k: getattr(st, k, None) for k in VEC_FIELDS + ("rho", "pe", "ke", "me", "me0", "stress")}
        return cls.from_dict(d, st.natom)

    def view(self) -> SldAtoms:
        v = SldAtoms()
        for k, _ in SldAtoms._fields_:
            arr = self.a.get(k)
            setattr(v, k, arr.ctypes.data_as(_dp) if arr is not None else None)
        self._v = v
        return v

    def as_dict(self) -> Dict[str, np.ndarray]:
        return dict(self.a)


class CellBuffers:
    def __init__(self, arrays):
        self.a = arrays

    @classmethod
    def empty(cls, c: int) -> "CellBuffers":
        return cls({k: np.zeros(c) for k in CELL_FIELDS})

    def view(self) -> SldCells:
        v = SldCells()
        for k in CELL_FIELDS:
            setattr(v, k, self.a[k].ctypes.data_as(_dp))
        self._v = v
        return v

    def as_dict(self):
        return dict(self.a)


def params_from_variables(v: Dict[str, object], box, natom: int, replica: int = 0) -> SldParams:
    """variables.in values (input units) -> SldParams (internal units), with
    the conversions of read_variables.cpp:155-159,258-276."""
    p = SldParams()
    p.natom = int(natom)
    p.interval_of_print_out = int(v.get("interval_of_print_out", 1))
    for i in range(6):
        p.box[i] = float(box[i])
    p.step = float(v.get("step", 1e-15))
    p.temperature = float(v.get("temperature", 300.0)) * K.BOLTZ
    p.atmass = K.mass_internal(float(v.get("atmass", K.FE_MASS_U)))
    p.gamma_L_over_mass = float(v.get("gamma_L_over_mass", 6e11))
    p.gamma_S_HL = float(v.get("gamma_S_HL", 5.88e13))
    p.kappa_e = float(v.get("kappa_e", 80.0))
    p.rcut_pot = float(v.get("rcut_pot", 4.1))
    p.rcut_mag = float(v.get("rcut_mag", 3.75))
    p.rcut_max = float(v.get("rcut_max", 4.1))
    mlc = float(v.get("min_length_link_cell", 4.11))
    if p.rcut_max > mlc:
        mlc = p.rcut_max + 0.001
    p.min_length_link_cell = mlc
    p.start_time = float(v.get("start_time", 0.0))
    p.random_seed = int(v.get("random_seed", 1234)) & 0xFFFFFFFFFFFFFFFF
    p.replica = int(replica)
    p.spinlang = int(v.get("spinlang", 1))
    p.lattlang = int(v.get("lattlang", 1))
    p.eltemp = int(v.get("eltemp", 1))
    p.localcolmot = int(v.get("localcolmot", 1)) if p.eltemp else 0
    p.renormalize_energy = int(v.get("renormalizeEnergy", 1)) if p.eltemp else 0
    # changestep (control.h) and its limits (default.h:218-219)
    p.changestep = int(v.get("changestep", 0))
    p.displace_limit = float(v.get("displace_limit", 0.01))
    p.phi_limit = float(v.get("phi_limit", 2.0 * np.pi / 10.0))
    p.extfield = int(v.get("extfield", 0))
    p.extforce = int(v.get("extforce", 0))
    # quantumnoise (control.h) and its sizes (default.h:222-225, read_variables.cpp:245-250)
    p.quantumnoise = int(v.get("quantumnoise", 0))
    p.msteps_quantum = int(v.get("Msteps_quantum", 50))
    p.nfrequency_quantum = int(v.get("Nfrequency_quantum", 150))
    # localvol and magmom (control.h); rcut_vol defaults to 1.2 a_lattice (default.h:96,208-211)
    p.localvol = int(v.get("localvol", 0))
    p.rcut_vol = float(v.get("rcut_vol", 2.83 * 1.2))
    p.magmom = int(v.get("magmom", 0))
    return p


class EngineLike:
    """Common surface of the HIP engine and the oracle (for tests)."""

    def state_for_compare(self) -> Dict[str, np.ndarray]:
        return self.download()
