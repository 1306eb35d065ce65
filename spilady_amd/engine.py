"""Python binding of libspilady_hip.so (include/spilady_hip.h) over ctypes.

The engine runs only on a gfx950 device through the HIP library: there is no
CPU fallback. Importing this module does not touch the GPU; creating an
Engine does, and raises if the library is missing or no device is usable.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import sys
import weakref
from typing import Dict, List, Optional

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libspilady_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "spilady_hip.h")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lib = None
# every context not yet destroyed: the exit hook destroys them while the HIP
# runtime is still whole (before interpreter teardown and the runtime's own
# exit handlers), so no device call is left to a finaliser at shutdown
_LIVE: "weakref.WeakSet[Engine]" = weakref.WeakSet()


@atexit.register
def _destroy_live_contexts():
    for e in list(_LIVE):
        try:
            e.close()
        except Exception:
            pass

ERRORS = {1: "SLD_ERR_ARG", 2: "SLD_ERR_HIP", 3: "SLD_ERR_CAPACITY", 4: "SLD_ERR_STATE",
          5: "SLD_ERR_NOGPU"}


class SldError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def exported_symbols() -> List[str]:
    """Every function declared in include/spilady_hip.h."""
    import re
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(sld_[a-z_0-9]+)\s*\(", txt)))


def _one_hip_runtime():
    """torch's ROCm wheel carries its own libamdhip64, which it loads under a
    file name that does not match the system runtime's soname. A process that
    loads this library first and torch afterwards holds two HIP runtimes, and
    the second finds no GPU. Importing torch first (when it is installed)
    makes this library bind to the runtime torch loaded, so device tensors
    and the library share one runtime in every import order."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load(path: Optional[str] = None):
    """Load the HIP library (no device call)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    _one_hip_runtime()
    if not os.path.exists(p):
        raise FileNotFoundError(f"{p} is not built: run __graft_entry__.build() "
                                "(make -C spilady_amd/csrc)")
    L = C.CDLL(p)
    P = C.POINTER(abi.SldParams)
    A = C.POINTER(abi.SldAtoms)
    V = C.c_void_p
    L.sld_version.restype = C.c_int
    if L.sld_version() != abi.ABI_VERSION:
        # the structs below would be read with another layout
        raise RuntimeError(f"{p} has ABI version {L.sld_version()}, this package expects {abi.ABI_VERSION}: "
                           "rebuild it (make -C spilady_amd/csrc)")
    L.sld_plan.argtypes = [_dp, C.c_double, _ip, _ip, _ip, _ip, _ip, _ip]
    L.sld_device_info.argtypes = [C.c_char_p, C.c_int, _ip]
    L.sld_create.argtypes = [P, A, C.c_int, C.POINTER(V)]
    L.sld_create_batch.argtypes = [C.c_int, P, A, C.c_int, C.POINTER(V)]
    L.sld_nreplicas.argtypes = [V]
    L.sld_select_replica.argtypes = [V, C.c_int]
    L.sld_destroy.argtypes = [V]
    L.sld_last_error.argtypes = [V]
    L.sld_last_error.restype = C.c_char_p
    for nm in ("sld_links", "sld_calculate_rho", "sld_calculate_force_energy",
               "sld_calculate_temperature"):
        getattr(L, nm).argtypes = [V]
    for nm in ("sld_core_ds", "sld_core_dp", "sld_core_dr", "sld_core_dTe"):
        getattr(L, nm).argtypes = [V, C.c_double]
    L.sld_check.argtypes = [V, C.c_int, C.POINTER(abi.SldObs)]
    L.sld_step.argtypes = [V, C.c_int, C.c_int, C.POINTER(abi.SldObs), _ip]
    L.sld_download.argtypes = [V, A]
    L.sld_upload.argtypes = [V, A]
    L.sld_get_cells.argtypes = [V, C.POINTER(abi.SldCells)]
    L.sld_set_cells_te.argtypes = [V, _dp]
    L.sld_get_geometry.argtypes = [V, _ip, _ip, _ip]
    L.sld_get_cell_lists.argtypes = [V, C.c_int, _ip, _ip]
    L.sld_total_time.argtypes = [V]
    L.sld_total_time.restype = C.c_double
    L.sld_scale_step.argtypes = [V]
    L.sld_set_barostat.argtypes = [V, C.POINTER(abi.SldBarostat)]
    L.sld_scale.argtypes = [V, C.c_int]
    L.sld_get_box.argtypes = [V, _dp]
    L.sld_set_external.argtypes = [V, _dp, _dp]
    L.sld_get_step.argtypes = [V]
    L.sld_get_step.restype = C.c_double
    L.sld_last_step_timing.argtypes = [V, _dp]
    L.sld_time_kernel.argtypes = [V, C.c_char_p]
    L.sld_kernel_timing.argtypes = [V, _ip, _dp]
    Sh = C.POINTER(abi.SldShapes)
    L.sld_get_shapes.argtypes = [V, Sh]
    L.sld_set_shapes.argtypes = [V, Sh]
    L.sld_autotune.argtypes = [V, C.c_int, Sh, _dp]
    L.sld_set_graph_mode.argtypes = [V, C.c_int]
    L.sld_model_nparams.argtypes = []
    L.sld_model_default_params.argtypes = [_dp]
    L.sld_model_set_params.argtypes = [V, _dp]
    L.sld_model_get_params.argtypes = [V, _dp]
    L.sld_model_natoms.argtypes = [V]
    # buffers are raw addresses: host arrays or device tensors
    L.sld_model_eval.argtypes = [V, C.c_void_p, C.c_void_p, C.c_void_p]
    L.sld_model_vjp.argtypes = [V, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.sld_model_eval_noalloc.argtypes = [V, C.c_void_p, C.c_void_p, C.c_void_p]
    if path is None:
        _lib = L
    return L


class Engine(abi.EngineLike):
    """One spin-lattice run on the GPU (the reference's global state, as a
    context). Method names follow the reference entry points."""

    def __init__(self, params: abi.SldParams, atoms, compute_initial_forces: bool = False):
        self._create([params], [atoms], compute_initial_forces)

    def _create(self, plist, alist, compute_initial_forces):
        self.L = load()
        self.nrep = len(plist)
        self.params = plist[0]
        self._natoms = [int(p.natom) for p in plist]
        self._bufs = [abi.AtomBuffers.from_state(a) for a in alist]
        self._pv = (abi.SldParams * self.nrep)(*plist)
        self._av = (abi.SldAtoms * self.nrep)(*[b.view() for b in self._bufs])
        h = C.c_void_p()
        if self.nrep == 1:
            rc = self.L.sld_create(self._pv, self._av, int(compute_initial_forces), C.byref(h))
        else:
            rc = self.L.sld_create_batch(self.nrep, self._pv, self._av, int(compute_initial_forces),
                                         C.byref(h))
        self.h = h
        if h:
            _LIVE.add(self)
        if rc != 0:
            msg = self.L.sld_last_error(h).decode() if h else "sld_create failed"
            self.close()
            raise SldError(rc, msg)
        self.select(0)

    def select(self, replica: int):
        """Replica that download/upload/cells/groups/cell_lists act on."""
        self._rc(self.L.sld_select_replica(self.h, replica))
        self.sel = replica
        self.natom = self._natoms[replica]
        n3 = np.zeros(3, dtype=np.int32)
        ng = np.zeros(1, dtype=np.int32)
        self._rc(self.L.sld_get_geometry(self.h, n3.ctypes.data_as(_ip), ng.ctypes.data_as(_ip), None))
        self.ncell3 = tuple(int(x) for x in n3)
        self.ncells = int(np.prod(n3))
        self.ngroups = int(ng[0])

    def _rc(self, rc: int):
        if rc != 0:
            raise SldError(rc, self.L.sld_last_error(self.h).decode())

    def close(self):
        """sld_destroy (free_memory.cpp:31-52): every device allocation,
        event, graph and the stream of the context."""
        if getattr(self, "h", None):
            h, self.h = self.h, None
            _LIVE.discard(self)
            self.L.sld_destroy(h)

    def __del__(self):
        # a context still open at interpreter shutdown was destroyed by the
        # exit hook; a finaliser that runs later makes no device call
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    # reference entry points
    def links(self): self._rc(self.L.sld_links(self.h))
    def calculate_rho(self): self._rc(self.L.sld_calculate_rho(self.h))
    def calculate_force_energy(self): self._rc(self.L.sld_calculate_force_energy(self.h))
    def calculate_temperature(self): self._rc(self.L.sld_calculate_temperature(self.h))
    def core_ds(self, dt): self._rc(self.L.sld_core_ds(self.h, dt))
    def core_dp(self, dt): self._rc(self.L.sld_core_dp(self.h, dt))
    def core_dr(self, dt): self._rc(self.L.sld_core_dr(self.h, dt))
    def core_dTe(self, dt): self._rc(self.L.sld_core_dTe(self.h, dt))

    def _check_all(self, step: int):
        ob = (abi.SldObs * self.nrep)()
        self._rc(self.L.sld_check(self.h, step, ob))
        return [abi.obs_dict(ob[r]) for r in range(self.nrep)]

    def _step_all(self, first_step: int, nsteps: int):
        obs = (abi.SldObs * (max(nsteps, 1) * self.nrep))()
        k = C.c_int(0)
        self._rc(self.L.sld_step(self.h, first_step, nsteps, obs, C.byref(k)))
        return [[abi.obs_dict(obs[i * self.nrep + r]) for r in range(self.nrep)] for i in range(k.value)]

    def check(self, step: int) -> Dict[str, float]:
        return self._check_all(step)[0]

    def step(self, first_step: int, nsteps: int):
        """nsteps of core() + check(); one observables dict per printed step."""
        return [x[0] for x in self._step_all(first_step, nsteps)]

    def last_step_ms(self) -> float:
        ms = C.c_double(0.0)
        self._rc(self.L.sld_last_step_timing(self.h, C.byref(ms)))
        return ms.value

    def time_kernel(self, name: Optional[str]):
        """Time every launch of one kernel with HIP events (None: off)."""
        self._rc(self.L.sld_time_kernel(self.h, (name or "").encode()))

    def kernel_timing(self):
        """(launches, total device ms) of the timed kernel over the last step or
        core_dTe call."""
        n = C.c_int(0)
        ms = C.c_double(0.0)
        self._rc(self.L.sld_kernel_timing(self.h, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def set_graph_mode(self, mode: int):
        """0 eager launches, 1 core_dTe from a graph, 2 every step from a graph."""
        self._rc(self.L.sld_set_graph_mode(self.h, int(mode)))

    def shapes(self) -> Dict[str, int]:
        """Launch shapes of the shape-independent kernels."""
        sh = abi.SldShapes()
        self._rc(self.L.sld_get_shapes(self.h, C.byref(sh)))
        return {k: int(getattr(sh, k)) for k, _ in abi.SldShapes._fields_}

    def set_shapes(self, shapes: Dict[str, int]):
        sh = abi.SldShapes(**{k: int(v) for k, v in {**self.shapes(), **shapes}.items()})
        self._rc(self.L.sld_set_shapes(self.h, C.byref(sh)))

    def autotune(self, steps: int = 2):
        """Pick the fastest launch shape of each tunable kernel by timing
        `steps` steps per candidate; the state is restored afterwards. Returns
        (chosen shapes, {kernel field: {candidate: device us per step}})."""
        sh = abi.SldShapes()
        us = np.zeros(3 * len(abi.SldShapes._fields_))
        self._rc(self.L.sld_autotune(self.h, int(steps), C.byref(sh), us.ctypes.data_as(_dp)))
        table = {}
        for k, (name, _) in enumerate(abi.SldShapes._fields_):
            cands = abi.SHAPE_CANDIDATES[name]
            table[name] = {c: float(us[3 * k + q]) for q, c in enumerate(cands)}
        return {k: int(getattr(sh, k)) for k, _ in abi.SldShapes._fields_}, table

    def download(self) -> Dict[str, np.ndarray]:
        out = abi.AtomBuffers.empty(self.natom)
        self._rc(self.L.sld_download(self.h, C.byref(out.view())))
        return out.as_dict()

    def upload(self, st: Dict[str, np.ndarray]):
        buf = abi.AtomBuffers.from_dict(st, self.natom)
        self._rc(self.L.sld_upload(self.h, C.byref(buf.view())))

    def cells(self) -> Dict[str, np.ndarray]:
        out = abi.CellBuffers.empty(self.ncells)
        self._rc(self.L.sld_get_cells(self.h, C.byref(out.view())))
        return out.as_dict()

    def set_cells_te(self, te):
        te = np.ascontiguousarray(te, dtype=np.float64)
        self._rc(self.L.sld_set_cells_te(self.h, te.ctypes.data_as(_dp)))

    def groups(self):
        sizes = np.zeros(self.ngroups, dtype=np.int32)
        self._rc(self.L.sld_get_geometry(self.h, None, None, sizes.ctypes.data_as(_ip)))
        return sizes

    def cell_lists(self, cap: int = 64):
        counts = np.zeros(self.ncells, dtype=np.int32)
        lists = np.full(self.ncells * cap, -1, dtype=np.int32)
        self._rc(self.L.sld_get_cell_lists(self.h, cap, counts.ctypes.data_as(_ip),
                                           lists.ctypes.data_as(_ip)))
        return counts, lists.reshape(self.ncells, cap)

    @property
    def total_time(self) -> float:
        return self.L.sld_total_time(self.h)

    @property
    def dt(self) -> float:
        """The current time step (changes under changestep)."""
        return self.L.sld_get_step(self.h)

    def scale_step(self):
        """scale_step (scale_step_CPU.cpp:33-97)."""
        self._rc(self.L.sld_scale_step(self.h))

    def set_barostat(self, b: abi.SldBarostat):
        """The STRESS / PRESSURE Berendsen barostats (scale_stress_CPU.cpp,
        scale_pressure_CPU.cpp); step() applies them after every step."""
        self._baro = b
        self._rc(self.L.sld_set_barostat(self.h, C.byref(b)))

    def scale(self, step: int):
        """scale(step) (scale.cpp:31-44): barostats at their intervals, then
        changestep; step() calls it itself."""
        self._rc(self.L.sld_scale(self.h, int(step)))

    def box(self) -> np.ndarray:
        """The current box matrix (xx yx yy zx zy zz) of the selected replica."""
        b = np.zeros(6)
        self._rc(self.L.sld_get_box(self.h, b.ctypes.data_as(_dp)))
        return b

    def set_external(self, hext=None, fext=None):
        """Per-atom external field (eV) and force (eV/A) of the selected
        replica, [natom][3] (needs extfield / extforce in the params)."""
        h = None if hext is None else np.ascontiguousarray(hext, dtype=np.float64)
        f = None if fext is None else np.ascontiguousarray(fext, dtype=np.float64)
        self._rc(self.L.sld_set_external(self.h, None if h is None else h.ctypes.data_as(_dp),
                                         None if f is None else f.ctypes.data_as(_dp)))


class BatchEngine(Engine):
    """Independent replicas of any size, temperature and seed advanced
    together: every kernel launch covers all of them (sld_create_batch)."""

    def __init__(self, params_list, atoms_list, compute_initial_forces: bool = False):
        if len(params_list) != len(atoms_list) or not params_list:
            raise ValueError("one params and one atoms entry per replica")
        self._create(list(params_list), list(atoms_list), compute_initial_forces)

    def check(self, step: int):
        """Observables of every replica (list of dicts)."""
        return self._check_all(step)

    def step(self, first_step: int, nsteps: int):
        """nsteps of core() + check(); per printed step a list of per-replica dicts."""
        return self._step_all(first_step, nsteps)

    def download_replica(self, replica: int):
        self.select(replica)
        return self.download()


def plan(box, min_length_link_cell: float):
    """Host-only link-cell plan (no device call): cell counts, greedy groups
    and the per-cell levels of the four level-scheduled sweeps."""
    L = load()
    b = np.ascontiguousarray(box, dtype=np.float64)
    n3 = np.zeros(3, dtype=np.int32)
    rc = L.sld_plan(b.ctypes.data_as(_dp), min_length_link_cell, n3.ctypes.data_as(_ip),
                    None, None, None, None, None)
    if rc != 0:
        raise SldError(rc, "sld_plan")
    C_ = int(np.prod(n3))
    ng = np.zeros(1, dtype=np.int32)
    gs = np.zeros(C_, dtype=np.int32)
    gc = np.zeros(C_, dtype=np.int32)
    nl = np.zeros(4, dtype=np.int32)
    lv = np.zeros(4 * C_, dtype=np.int32)
    rc = L.sld_plan(b.ctypes.data_as(_dp), min_length_link_cell, n3.ctypes.data_as(_ip),
                    ng.ctypes.data_as(_ip), gs.ctypes.data_as(_ip), gc.ctypes.data_as(_ip),
                    nl.ctypes.data_as(_ip), lv.ctypes.data_as(_ip))
    if rc != 0:
        raise SldError(rc, "sld_plan")
    g = int(ng[0])
    return {"ncell3": tuple(int(x) for x in n3), "group_sizes": gs[:g].copy(), "group_cells": gc,
            "nlevels": nl, "levels": lv.reshape(4, C_)}


def device_info():
    L = load()
    name = C.create_string_buffer(64)
    ncu = C.c_int(0)
    rc = L.sld_device_info(name, 64, C.byref(ncu))
    if rc != 0:
        raise SldError(rc, "no HIP device")
    return name.value.decode(), ncu.value
