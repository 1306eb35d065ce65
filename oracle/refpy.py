"""ctypes front end of oracle/_ref/libspilady_ref.so (the real SPILADY CPU
reference, built by oracle/Makefile.ref from the reference's own sources).

Test infrastructure only: tests, the smoke check and bench.py's cpu_baseline
leg use it as the checker. The reference keeps its state in process globals,
so one Reference is live per process; `run_isolated` runs a job in a child
process when a fresh reference state is needed.
"""
from __future__ import annotations

import ctypes as C
import os
import tempfile
from typing import Dict, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_LIB = os.path.join(HERE, "_ref", "libspilady_ref.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

# The reference's variables.in keys (read_variables.cpp:51-242) that these
# runs set; the control.h switches are compile-time in the reference build.
REF_KEYS = ["OMP_threads", "out_body", "step", "interval_of_print_out",
            "interval_of_config_out", "in_config", "kappa_e", "start_time",
            "atmass", "temperature", "gamma_L_over_mass", "gamma_S_HL",
            "random_seed", "rcut_pot", "rcut_mag", "rcut_max",
            "min_length_link_cell", "displace_limit", "phi_limit",
            "interval_of_scale_stress", "stress_xx", "stress_yx", "stress_yy",
            "stress_zx", "stress_zy", "stress_zz", "interval_of_scale_pressure",
            "pressure", "baro_damping_time", "Msteps_quantum", "Nfrequency_quantum", "rcut_vol",
            "interval_of_vsim", "vsim_prec"]

# Fixed switches of the reference build (shipped control.h).
REF_SWITCHES = {"spinlang": 1, "lattlang": 1, "eltemp": 1, "localcolmot": 1,
                "renormalizeEnergy": 1}


# the same sources built with -Dextfield -Dextforce (Makefile.ref OUT=_ref/ext)
REF_LIB_EXT = os.path.join(HERE, "_ref", "ext", "libspilady_ref.so")


def available() -> bool:
    return os.path.exists(REF_LIB)


def available_ext() -> bool:
    return os.path.exists(REF_LIB_EXT)


# the same sources built with -DSTRESS -DPRESSURE (Makefile.ref OUT=_ref/baro)
REF_LIB_BARO = os.path.join(HERE, "_ref", "baro", "libspilady_ref.so")


def available_baro() -> bool:
    return os.path.exists(REF_LIB_BARO)


# the same sources built with -Dquantumnoise (Makefile.ref OUT=_ref/qn)
REF_LIB_QN = os.path.join(HERE, "_ref", "qn", "libspilady_ref.so")


def available_qn() -> bool:
    return os.path.exists(REF_LIB_QN)


# the same sources built with -Dlocalvol (Makefile.ref OUT=_ref/lv)
REF_LIB_LV = os.path.join(HERE, "_ref", "lv", "libspilady_ref.so")


def available_lv() -> bool:
    return os.path.exists(REF_LIB_LV)


# the same sources built with -Dmagmom (Makefile.ref OUT=_ref/mm)
REF_LIB_MM = os.path.join(HERE, "_ref", "mm", "libspilady_ref.so")


def available_mm() -> bool:
    return os.path.exists(REF_LIB_MM)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


class Reference:
    def __init__(self, lib: Optional[str] = None):
        path = lib or REF_LIB
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (oracle/Makefile.ref)")
        self.lib = C.CDLL(path)
        L = self.lib
        L.ref_init_files.argtypes = [C.c_char_p]
        L.ref_init_memory.argtypes = [C.c_char_p, C.c_int] + [_dp] * 11 + [C.c_double]
        L.ref_run.argtypes = [C.c_int, C.c_int, C.c_int]
        for nm in ("ref_core_ds", "ref_core_dp", "ref_core_dr", "ref_core_dTe"):
            getattr(L, nm).argtypes = [C.c_double]
        L.ref_check.argtypes = [C.c_int]
        L.ref_scale.argtypes = [C.c_int]
        L.ref_get_box.argtypes = [_dp]
        L.ref_write_vsim.argtypes = [C.c_int]
        L.ref_quantum_H.argtypes = [_dp]
        for nm in ("ref_step", "ref_total_time", "ref_virial", "ref_initial_ave_energy"):
            getattr(L, nm).restype = C.c_double
        L.ref_scalar.restype = C.c_double
        L.ref_scalar.argtypes = [C.c_char_p]
        L.ref_table.restype = C.c_double
        L.ref_table.argtypes = [C.c_int, C.c_double]
        L.ref_table_raw.restype = C.c_double
        L.ref_table_raw.argtypes = [C.c_int, C.c_int]
        L.ref_get_atoms.argtypes = [_dp] * 14 + [_ip]
        L.ref_set_atoms.argtypes = [_dp] * 5
        L.ref_get_local_volume.argtypes = [_dp]
        L.ref_get_cell_lists.argtypes = [C.c_int, _ip, _ip]
        L.ref_get_groups.argtypes = [_ip, _ip]
        L.ref_get_cells.argtypes = [_dp] * 11
        L.ref_set_cells_te.argtypes = [_dp]
        L.ref_cell_dims.argtypes = [_ip]
        self.dir = None

    # --- set-up -----------------------------------------------------------
    @staticmethod
    def write_inputs(d: str, variables: Dict[str, object], atoms=None) -> None:
        from spilady_amd.io import write_variables, write_con
        v = dict(variables)
        write_variables(os.path.join(d, "variables.in"), v,
                        keys=[k for k in REF_KEYS if k in v])
        if atoms is not None:
            write_con(os.path.join(d, v["in_config"]), atoms)

    def init_files(self, d: str) -> int:
        self.dir = d
        return self.lib.ref_init_files(d.encode())

    def init_memory(self, d: str, variables: Dict[str, object], atoms) -> int:
        self.write_inputs(d, variables, None)
        self.dir = d
        a = atoms
        arr = [np.ascontiguousarray(x, dtype=np.float64) for x in
               (a.box, a.r, a.p, a.f, a.stress, a.rho, a.ke, a.pe, a.s, a.me, a.me0)]
        self._keep = arr
        return self.lib.ref_init_memory(d.encode(), a.natom, *[_ptr(x) for x in arr],
                                        C.c_double(float(variables.get("start_time", 0.0))))

    # --- stepping ---------------------------------------------------------
    def run(self, first_step: int, nsteps: int, check: bool = True) -> None:
        self.lib.ref_run(first_step, nsteps, int(check))

    def __getattr__(self, name):
        if name.startswith("ref_"):
            return getattr(self.lib, name)
        raise AttributeError(name)

    # --- state ------------------------------------------------------------
    @property
    def natom(self) -> int:
        return self.lib.ref_natom()

    @property
    def ncells(self) -> int:
        return self.lib.ref_ncells()

    def cell_dims(self):
        n3 = np.zeros(3, dtype=np.int32)
        self.lib.ref_cell_dims(n3.ctypes.data_as(_ip))
        return tuple(int(x) for x in n3)

    def atoms(self) -> Dict[str, np.ndarray]:
        n = self.natom
        out = {k: np.zeros((n, 3)) for k in ("r", "p", "f", "s", "heff_h", "heff_l")}
        for k in ("rho", "pe", "ke", "me", "me0", "vir", "sum_jij_sj"):
            out[k] = np.zeros(n)
        out["stress"] = np.zeros((n, 6))
        out["cell"] = np.zeros(n, dtype=np.int32)
        self.lib.ref_get_atoms(*[_ptr(out[k]) for k in
                                 ("r", "p", "f", "s", "heff_h", "heff_l", "rho", "pe",
                                  "ke", "me", "me0", "stress", "vir", "sum_jij_sj")],
                               out["cell"].ctypes.data_as(_ip))
        return out

    def local_volume(self) -> np.ndarray:
        lv = np.zeros(self.natom)
        self.lib.ref_get_local_volume(_ptr(lv))
        return lv

    def moments(self) -> np.ndarray:
        m = np.zeros((self.natom, 3))
        self.lib.ref_get_m(_ptr(m))
        return m

    def set_atoms(self, r, p, f, s, rho) -> None:
        arr = [np.ascontiguousarray(x, dtype=np.float64) for x in (r, p, f, s, rho)]
        self.lib.ref_set_atoms(*[_ptr(x) for x in arr])

    def cells(self) -> Dict[str, np.ndarray]:
        c = self.ncells
        keys = ("te", "ee", "tl", "ts_r", "ts_l", "gel", "ges", "sum_r_up", "sum_r_dn",
                "sum_l_up", "sum_l_dn")
        out = {k: np.zeros(c) for k in keys}
        self.lib.ref_get_cells(*[_ptr(out[k]) for k in keys])
        return out

    def cell_lists(self, cap: int = 64):
        c = self.ncells
        counts = np.zeros(c, dtype=np.int32)
        lists = np.full(c * cap, -1, dtype=np.int32)
        mx = self.lib.ref_get_cell_lists(cap, counts.ctypes.data_as(_ip), lists.ctypes.data_as(_ip))
        if mx > cap:
            raise RuntimeError(f"cell holds {mx} atoms > cap {cap}")
        return counts, lists.reshape(c, cap)

    def groups(self):
        sizes = np.zeros(self.ncells, dtype=np.int32)
        members = np.zeros(self.ncells, dtype=np.int32)
        self.lib.ref_get_groups(sizes.ctypes.data_as(_ip), members.ctypes.data_as(_ip))
        g = self.lib.ref_ngroups()
        return sizes[:g].copy(), members

    def box(self) -> np.ndarray:
        b = np.zeros(6)
        self.lib.ref_get_box(_ptr(b))
        return b

    def quantum_H(self) -> np.ndarray:
        h = np.zeros(self.lib.ref_quantum_taps())
        if h.size:
            self.lib.ref_quantum_H(_ptr(h))
        return h

    def scalar(self, name: str) -> float:
        return self.lib.ref_scalar(name.encode())

    def free(self) -> None:
        self.lib.ref_free()


def consistent_start(variables: Dict[str, object], atoms, workdir: Optional[str] = None,
                     lib: Optional[str] = None):
    """Fill f, rho, pe, me, me0 and stress of `atoms` with the reference's own
    evaluation (calculate_rho + calculate_force_energy, as initialize() does
    without readconf, initial.cpp:78-83), so that a readconf start is
    consistent. `lib` picks a variant build (the magmom one reads the spin
    column as moments). Returns a new AtomState."""
    own = workdir is None
    d = workdir or tempfile.mkdtemp(prefix="sldref_")
    ref = Reference(lib)
    ref.init_memory(d, variables, atoms)
    ref.ref_calculate_rho()
    ref.ref_calculate_force_energy()
    a = ref.atoms()
    ref.free()
    out = atoms.copy()
    for k in ("f", "rho", "pe", "me", "me0", "stress"):
        setattr(out, k, a[k].copy())
    out.local_volume = np.full(out.natom, np.prod(out.box[[0, 2, 5]]) / out.natom)
    if own:
        import shutil
        shutil.rmtree(d, ignore_errors=True)
    return out
