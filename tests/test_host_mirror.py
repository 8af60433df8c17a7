"""The C++ call surface (ccake_amd/host/ccake.hpp: SystemState,
SPHWorkstation, Evolver) driving the gfx950 library, through its Gubser
driver program."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROG = os.path.join(ROOT, "ccake_amd", "bin", "gubser_run")


def test_driver_built_and_fails_cleanly_without_device():
    assert os.path.exists(PROG), "build() has not produced ccake_amd/bin/gubser_run"
    if os.path.exists("/dev/kfd"):
        pytest.skip("a HIP device is present")
    r = subprocess.run([PROG, "--steps", "1"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "ccake_dev_create" in r.stderr


@pytest.mark.gpu
def test_gubser_through_cpp_call_surface():
    """120 RK2 steps of ideal Gubser flow (40,401 particles) through
    SPHWorkstation::advance_timestep: energy density within the SPH
    smoothing error of the analytic solution, entropy unchanged, E + E_z
    conserved to 1 percent; 5 steps through Evolver::execute_timestep with
    get_time_derivatives as the functional agree with the fused device step."""
    r = subprocess.run([PROG, "--steps", "120", "--evolver-steps", "5"], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout.strip().splitlines()[-1])
    assert d["particles"] == 40401
    assert abs(d["t"] - 1.6) < 1e-9
    assert d["max_rel_err_e_r_lt_2.5"] < 0.07
    assert abs(d["entropy_rel_change"]) < 1e-14
    assert abs(d["energy_rel_change"]) < 0.01
    assert 0.0 <= d["evolver_vs_fused_max_abs_diff"] < 1e-12
