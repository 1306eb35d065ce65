"""SPILADY file formats: variables.in, con- and cel- files, observable logs.

The formats are the reference's observable contract and are kept as is:
  * variables.in: whitespace "key value" pairs (read_variables.cpp:51-242),
    with the reference's defaults (default.h:44-226);
  * con-<body>_NNNN.dat: header "natom total_time TYPE", two box lines, then
    one line per atom (write_config.cpp:84-131). The Heff_H and Heff_L
    columns hold x, x, z: the reference writes (and reads) the x component
    twice (write_config.cpp:117,120; read_config.cpp:103,106);
  * cel-<body>_NNNN.dat: per link cell centre, Te (K), Tl, Ts_R, Ts_L
    (write_config.cpp:135-204);
  * enr-/tmp-/spn- logs (check_energy_CPU.cpp:113-137,
    check_temperature_CPU.cpp:136-161, check_spin_CPU.cpp:58-79).
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field
from typing import Dict

import numpy as np

from . import constants as K

# Default values of the runtime keys this engine reads (default.h:44-226).
# Switches that the reference fixes at compile time in control.h:8-43 are
# runtime keys here (1 = on), so one binary covers every configuration.
VARIABLE_DEFAULTS: Dict[str, object] = {
    "out_body": "spilady",
    "no_of_production_steps": 1000,
    "step": 1e-15,
    "interval_of_print_out": 1,
    "interval_of_config_out": 1000,
    "in_config": "con-input.dat",
    "kappa_e": 80e0,
    "start_time": 0e0,
    "atmass": 55.847,
    "temperature": 300e0,
    "gamma_L_over_mass": 6e11,
    "gamma_S_HL": 5.88e13,
    "random_seed": 1234,
    "rcut_pot": 4.1e0,
    "rcut_mag": 3.75e0,
    "rcut_max": 4.1e0,
    "min_length_link_cell": 4.11e0,
    "OMP_threads": 2,
    # control.h switches (runtime here)
    "spinlang": 1,
    "lattlang": 1,
    "eltemp": 1,
    "localcolmot": 1,
    "renormalizeEnergy": 1,
    "runstep": 1,              # 0: run while total_time <= total_production_time
    "changestep": 0,           # adapt dt after every step (scale_step_CPU.cpp)
    # time loop and changestep limits (default.h:218-219, read_variables.cpp:79,238-241)
    "total_production_time": 1e-12,
    "displace_limit": 0.01,
    "phi_limit": 0.628319,
    "extfield": 0,             # read extfield.in (tesla) at the start
    "extforce": 0,             # read extforce.in (eV/A) at the start
    # Berendsen barostats (control.h STRESS / PRESSURE; default.h:177-193)
    "STRESS": 0,
    "PRESSURE": 0,
    "interval_of_scale_stress": 1,
    "stress_xx": 0e0, "stress_yx": 0e0, "stress_yy": 0e0,   # GPa
    "stress_zx": 0e0, "stress_zy": 0e0, "stress_zz": 0e0,
    "interval_of_scale_pressure": 1,
    "pressure": 0e0,                                          # GPa
    "baro_damping_time": 1e-13,                               # s
    # quantum thermostat (control.h quantumnoise; default.h:222-225)
    "quantumnoise": 0,
    "Msteps_quantum": 50,
    "Nfrequency_quantum": 150,
    # control.h readTe: cell electron temperatures from a cel- file (default.h:83-85)
    "readTe": 0,
    "in_eltemp": "cel-input.dat",
    # control.h writevsim: V_sim files vsm-<body>_NNNN.ascii/.spin/.dat (default.h:67-69)
    "writevsim": 0,
    "interval_of_vsim": 1000,
    "vsim_prec": 4,
}


def read_variables(path: str) -> Dict[str, object]:
    """Parse variables.in the way read_variables.cpp:51-242 does: pairs of
    whitespace-separated tokens, later keys overriding earlier ones; unknown
    keys are ignored."""
    out = dict(VARIABLE_DEFAULTS)
    with open(path) as fh:
        toks = fh.read().split()
    for k, v in zip(toks[0::2], toks[1::2]):
        if k not in VARIABLE_DEFAULTS:
            continue
        d = VARIABLE_DEFAULTS[k]
        if isinstance(d, int) and not isinstance(d, bool):
            out[k] = int(float(v)) if k != "random_seed" else int(v)
        elif isinstance(d, float):
            out[k] = float(v)
        else:
            out[k] = v
    # derived quantity (read_variables.cpp:262)
    if out["rcut_max"] > out["min_length_link_cell"]:
        out["min_length_link_cell"] = out["rcut_max"] + 0.001
    return out


def write_variables(path: str, values: Dict[str, object], keys=None) -> None:
    keys = list(values.keys()) if keys is None else keys
    with open(path, "w") as fh:
        for k in keys:
            v = values[k]
            if isinstance(v, float):
                fh.write(f"{k} {v!r}\n")
            else:
                fh.write(f"{k} {v}\n")


@dataclass
class AtomState:
    """Per-atom state in atom-index order, vectors as [N, 3] float64."""
    box: np.ndarray                       # [6]: xx yx yy zx zy zz
    r: np.ndarray
    p: np.ndarray
    f: np.ndarray
    s: np.ndarray
    stress: np.ndarray                    # [N, 6]: 11 22 33 12 23 31
    rho: np.ndarray
    ke: np.ndarray
    pe: np.ndarray
    me: np.ndarray
    me0: np.ndarray
    heff_h: np.ndarray = None
    heff_l: np.ndarray = None
    local_volume: np.ndarray = None
    total_time: float = 0.0
    extra: dict = field(default_factory=dict)

    @property
    def natom(self) -> int:
        return int(self.r.shape[0])

    def copy(self) -> "AtomState":
        kw = {}
        for k, v in self.__dict__.items():
            kw[k] = v.copy() if isinstance(v, np.ndarray) else (dict(v) if isinstance(v, dict) else v)
        return AtomState(**kw)


def _fmt(x: np.ndarray) -> np.ndarray:
    return np.char.mod("%.15e", x)


def write_con(path: str, st: AtomState, kind: str = "SLDHL") -> None:
    """Write a con- file in the column layout of write_config.cpp:84-131."""
    n = st.natom
    d = st.box
    s0 = np.sqrt(np.sum(st.s * st.s, axis=1))
    hh = st.heff_h if st.heff_h is not None else np.zeros((n, 3))
    hl = st.heff_l if st.heff_l is not None else np.zeros((n, 3))
    lv = st.local_volume if st.local_volume is not None else np.zeros(n)
    cols = np.column_stack([
        st.r, st.p, st.f, st.stress, st.rho, st.ke, st.pe,
        st.s, s0, hh[:, 0], hh[:, 0], hh[:, 2], hl[:, 0], hl[:, 0], hl[:, 2],
        st.me, st.me0, lv])
    body = np.char.add(np.char.add(np.arange(n).astype(str), " Fe "),
                       np.array([" ".join(row) for row in _fmt(cols)]))
    with open(path, "w") as fh:
        fh.write(f"{n} {st.total_time:.15e} {kind}\n")
        fh.write(f"{d[0]:.15e} {d[1]:.15e} {d[2]:.15e}\n")
        fh.write(f"{d[3]:.15e} {d[4]:.15e} {d[5]:.15e}\n")
        fh.write("\n".join(body.tolist()))
        fh.write("\n")


def read_con(path: str, kind: str = "SLDHL") -> AtomState:
    """Read a con- file (read_config.cpp:39-108). The Heff y columns are the
    duplicated x columns and are returned as written."""
    with open(path) as fh:
        head = fh.readline().split()
        n = int(head[0])
        total_time = float(head[1])
        if head[2] != kind:
            raise ValueError(f"the input configuration file is for {head[2]}, not for {kind}")
        b1 = [float(x) for x in fh.readline().split()]
        b2 = [float(x) for x in fh.readline().split()]
        data = np.loadtxt(fh, usecols=range(2, 33), ndmin=2)
    if data.shape[0] != n:
        raise ValueError(f"con- file holds {data.shape[0]} atoms, header says {n}")
    c = 0

    def take(k):
        nonlocal c
        v = data[:, c:c + k]
        c += k
        return v.copy()
    r, p, f, stress = take(3), take(3), take(3), take(6)
    rho, ke, pe = take(1)[:, 0], take(1)[:, 0], take(1)[:, 0]
    s = take(3)
    take(1)  # s0
    hh, hl = take(3), take(3)
    me, me0, lv = take(1)[:, 0], take(1)[:, 0], take(1)[:, 0]
    return AtomState(box=np.array(b1 + b2), r=r, p=p, f=f, s=s, stress=stress,
                     rho=rho, ke=ke, pe=pe, me=me, me0=me0, heff_h=hh, heff_l=hl,
                     local_volume=lv, total_time=total_time)


def con_name(out_body: str, n_output: int) -> str:
    """File name rule of write_config.cpp:42-76."""
    return f"con-{out_body}_{n_output:04d}.dat"


def write_cel(path: str, box, ncell3, te_ev, tl, ts_r, ts_l, total_time: float) -> None:
    """cel- file of write_config.cpp:160-200 (Te written in K)."""
    nx, ny, nz = ncell3
    d = box
    lines = [f"{nx*ny*nz} {total_time:.15e}",
             f"{d[0]:.15e} {d[1]:.15e} {d[2]:.15e}",
             f"{d[3]:.15e} {d[4]:.15e} {d[5]:.15e}"]
    for ix in range(nx):
        for iy in range(ny):
            for iz in range(nz):
                c = ix + iy * nx + iz * nx * ny
                q = ((ix + 0.5) / nx, (iy + 0.5) / ny, (iz + 0.5) / nz)
                x = d[0] * q[0] + d[1] * q[1] + d[3] * q[2]
                y = d[2] * q[1] + d[4] * q[2]
                z = d[5] * q[2]
                lines.append(f"{x:.15e} {y:.15e} {z:.15e} {te_ev[c]/K.BOLTZ:.15e} "
                             f"{tl[c]:.15e} {ts_r[c]:.15e} {ts_l[c]:.15e} ")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def read_cel(path: str):
    """cel- file as the readTe start reads it (map_cells.cpp:33-60,125-160):
    returns (total_time, box[6], columns[C][7]) with the cells in file order
    (x outermost, z innermost) and Te in K."""
    with open(path) as fh:
        toks = fh.read().split()
    nc, t = int(toks[0]), float(toks[1])
    box = np.array([float(x) for x in toks[2:8]])
    cols = np.array([float(x) for x in toks[8:8 + 7 * nc]]).reshape(nc, 7)
    return t, box, cols


def read_log(path: str) -> np.ndarray:
    """enr-/tmp-/spn- logs: whitespace columns, one line per printed step."""
    if not os.path.exists(path):
        return np.zeros((0, 0))
    return np.loadtxt(path, ndmin=2)
