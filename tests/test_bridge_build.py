"""The drop-in boundary from the reference's side: SPILADY's own sources,
compiled where they lie with -DHIPLIB and spilady_amd/csrc/spilady_hip_bridge.cpp
in place of its CPU/CUDA backends, link against libspilady_hip.so
(oracle/build_hiplib.sh). The bridge is compiled with
-Werror=missing-declarations, so each function it defines must match a
prototype SPILADY declares (prototype_CPU.h:120-289, or the GPU lifecycle
block of prototype_CPU.h:275-289 that spilady_hip_bridge.h declares for
HIPLIB). Every optional switch the library supports is built: the shipped
control.h, the external field and force, the two barostats, the quantum
thermostat, readTe, and the local volumes with the moments as the spin
variable. Skipped where the reference's sources are absent."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
LIB = os.path.join(ROOT, "spilady_amd", "libspilady_hip.so")
SCRIPT = os.path.join(ROOT, "oracle", "build_hiplib.sh")

# the backend functions a HIPLIB build of SPILADY leaves to the bridge, per
# switch set (what the linker reports undefined without it)
BASE = ["initial_GPU()", "potential_table_GPU()", "free_potential_ptr_GPU()", "copy_CPU_to_GPU()",
        "copy_atoms_from_GPU_to_CPU()", "copy_cells_from_GPU_to_CPU()", "free_copy_CPU_to_GPU()",
        "initial_links()", "links()", "allocate_cells()", "free_allocate_memory()", "calculate_rho()",
        "calculate_force_energy()", "calculate_temperature()", "core_ds(double)", "core_dp(double)",
        "core_dr(double)", "core_dTe(double)", "check_energy(int)", "check_temperature(int)",
        "check_stress(int)", "check_pressure(int)", "check_spin(int)", "scale_step()",
        "initialize_random_number()", "free_random_number()"]
VARIANTS = {
    "shipped": ([], []),
    "external": (["-Dextfield", "-Dextforce"], ["external_field(int)", "external_force(int)"]),
    "barostats": (["-DSTRESS", "-DPRESSURE"], ["scale_stress()", "scale_pressure()"]),
    "quantum": (["-Dquantumnoise"], ["initial_quantum_noise()", "free_quantum_noise()"]),
    "readTe": (["-DreadTe"], []),
    "localvol_magmom": (["-Dlocalvol", "-Dmagmom"], []),
}


@pytest.mark.skipif(not os.path.isfile(os.path.join(REF, "spilady.cpp")), reason="reference sources absent")
@pytest.mark.skipif(not os.path.exists(LIB), reason="libspilady_hip.so not built")
@pytest.mark.parametrize("variant", sorted(VARIANTS))
def test_spilady_builds_and_links_on_the_library(tmp_path, variant):
    flags, extra = VARIANTS[variant]
    out = str(tmp_path / "hiplib")
    res = subprocess.run([SCRIPT, REF] + flags, env=dict(os.environ, SLD_HIPLIB_OUT=out),
                         capture_output=True, text=True, timeout=600)
    assert res.returncode == 0, res.stdout[-3000:] + res.stderr[-3000:]
    exe = os.path.join(out, "spilady")
    assert os.access(exe, os.X_OK)
    # the bridge defines each backend function, with SPILADY's signature
    nm = subprocess.run(["nm", "-C", "--defined-only", os.path.join(out, "obj", "zz_spilady_hip_bridge.o")],
                        capture_output=True, text=True, check=True).stdout
    defined = {ln.split(" ", 2)[2] for ln in nm.splitlines() if " T " in ln}
    missing = [f for f in BASE + extra if f not in defined]
    assert not missing, missing
    # and the program takes the C ABI from the library
    dyn = subprocess.run(["nm", "-D", "--undefined-only", exe], capture_output=True, text=True, check=True).stdout
    used = {ln.split()[-1] for ln in dyn.splitlines()}
    for sym in ("sld_create", "sld_links", "sld_core_ds", "sld_core_dTe", "sld_check", "sld_download",
                "sld_destroy"):
        assert sym in used, sym
