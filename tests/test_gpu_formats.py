"""The C++ driver's logs against the compiled reference's own writers
(check_energy_CPU.cpp:113-137, check_temperature_CPU.cpp:136-161,
check_stress_CPU.cpp:63-83, check_pressure_CPU.cpp:40-58,
check_spin_CPU.cpp:58-79): one run with the switches the reference was built
with (spinlang, lattlang, eltemp, localcolmot, renormalizeEnergy) and both
friction constants at zero, so that the Langevin noise vanishes and both sides
follow one trajectory, written by the driver (spilady_amd/spilady_hip) and by
the reference (oracle/_ref), compared token by token:
- the same files with the same lines and the same number of columns;
- the step and time columns character for character;
- every value in the reference's notation (scientific, 15 digits after the
  point) and equal to it within 1e-9 relative.
The con- and cel- writers are compared byte for byte in test_gpu_dropin.py,
where SPILADY's own write_config writes from this library's state."""
import os
import re
import subprocess

import numpy as np
import pytest

from helpers import consistent_state, fe_params
from oracle import refpy
from spilady_amd import abi, io

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "spilady_amd", "spilady_hip")
NUM = re.compile(r"^-?\d\.\d{15}e[+-]\d{2,3}$")


@pytest.mark.skipif(not refpy.available(), reason="the compiled reference (oracle/_ref) is not built")
def test_driver_logs_match_the_reference_writers(tmp_path):
    v, p, at = fe_params(8, 300.0, 23, thermostats=True, eltemp=True, localcolmot=True, renormalize=True)
    assert all(v[k] == on for k, on in refpy.REF_SWITCHES.items())
    v = dict(v, gamma_L_over_mass=0.0, gamma_S_HL=0.0)
    p = abi.params_from_variables(v, at.box, at.natom)
    st = consistent_state(p, at)
    state = io.AtomState(box=np.asarray(at.box), r=st["r"], p=st["p"], f=st["f"], s=st["s"],
                         stress=st["stress"], rho=st["rho"], ke=st["ke"], pe=st["pe"], me=st["me"],
                         me0=st["me0"])
    v = dict(v, in_config="con-start.dat", out_body="fmt", runstep=1, changestep=0, no_of_production_steps=6,
             interval_of_print_out=2, interval_of_config_out=1000)
    drv, ref = tmp_path / "drv", tmp_path / "ref"
    drv.mkdir()
    ref.mkdir()
    io.write_con(str(drv / "con-start.dat"), state)
    io.write_variables(str(drv / "variables.in"), v)
    res = subprocess.run([DRIVER, "variables.in"], cwd=str(drv), capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    r = refpy.Reference()
    r.init_memory(str(ref), v, state)   # initialize() and check(-1)
    r.run(0, 6, True)                   # core() and check() of steps 0..5
    r.free()
    names = [f"{k}-fmt.dat" for k in ("enr", "tmp", "spn", "str", "prs")]
    for name in names:
        a = (ref / name).read_text().splitlines()
        b = (drv / name).read_text().splitlines()
        assert len(a) == len(b) == 4, (name, len(a), len(b))   # steps -1, 1, 3, 5
        for la, lb in zip(a, b):
            ta, tb = la.split(" "), lb.split(" ")
            assert len(ta) == len(tb), (name, la, lb)
            assert ta[:2] == tb[:2], (name, la, lb)
            for x, y in zip(ta[2:], tb[2:]):
                assert NUM.match(x) and NUM.match(y), (name, x, y)
                fx, fy = float(x), float(y)
                assert abs(fx - fy) <= 1e-9 * max(abs(fx), 1e-300) or abs(fx - fy) <= 1e-15, (name, x, y)
