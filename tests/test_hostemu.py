"""The library's HIP sources built for the CPU (tests/hostemu) and run under
AddressSanitizer: every device-buffer and private-array access of every
kernel on the initialize and time-step paths is checked, and the results
are compared with the oracle particle by particle.

This is how the cause of the GPU memory faults seen at 251,001 particles
was found: an overlap in the private column enumeration made k_init write
one whole column past the end of the field buffer (2 MB past it at that
size).  test_private_columns_disjoint and the static_asserts in
ccake_internal.h keep the columns disjoint.
"""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EMU = os.path.join(HERE, "hostemu")


@pytest.fixture(scope="module")
def emu_lib():
    sys.path.insert(0, EMU)
    import build_emu
    return build_emu.build(asan=True)


def run_case(*args, timeout=600):
    env = dict(os.environ)
    r = subprocess.run(["bash", os.path.join(EMU, "emu.sh"), os.path.join(EMU, "run_case.py"), *args],
                       capture_output=True, text=True, timeout=timeout, env=env)
    print(r.stdout[-4000:], r.stderr[-8000:])
    return r


def test_gubser_on_host_emulation_matches_oracle(emu_lib):
    r = run_case("gubser", "0.2", "2")
    assert r.returncode == 0 and "AddressSanitizer" not in r.stderr
    assert "OK" in r.stdout


def test_lumpy_table_event_on_host_emulation(emu_lib):
    """10,201-particle lumpy event on the table EoS with its fallback chain:
    initialize and 3 steps, no invalid access, fields within 1e-10 per
    particle of the oracle (1e-12 for positive fields), integer fields equal."""
    r = run_case("table", "0.1", "3")
    assert r.returncode == 0 and "AddressSanitizer" not in r.stderr
    assert "OK" in r.stdout


def test_mixed_precision_on_host_emulation(emu_lib):
    """The FP32 pair kernels (pair_precision 1) under AddressSanitizer: no
    invalid access, and after one derivative evaluation and 2 steps of a
    lumpy viscous BSQ event (2,042 particles) every compared field is within
    1e-6 (whole-field relative) of the FP64 path, with equal btrack."""
    env = dict(os.environ)
    r = subprocess.run(["bash", os.path.join(EMU, "emu.sh"), os.path.join(EMU, "run_mixed.py"), "0.2", "2", "1e-6"],
                       capture_output=True, text=True, timeout=600, env=env)
    print(r.stdout[-4000:], r.stderr[-8000:])
    assert r.returncode == 0 and "AddressSanitizer" not in r.stderr
    assert "OK" in r.stdout


@pytest.mark.parametrize("case", ["milne", "milne_fact"])
def test_milne_on_host_emulation(emu_lib, case):
    """The 3+1D path (1,331 particles of a Milne lumpy event on the table EoS,
    the isotropic 3D kernel or the factorized one) under AddressSanitizer:
    no invalid access in the (x, y, eta) grid, the three-pass key sort, the
    3D pair walks and the D = 3 equations of motion; after initialization
    and 2 steps every field within 1e-12 of the oracle (whole field; per
    particle 1e-12 for positive fields, 1e-10 for signed ones)."""
    r = run_case(case, "0.2", "2")
    assert r.returncode == 0 and "AddressSanitizer" not in r.stderr
    assert "OK" in r.stdout


def test_batch_on_host_emulation(emu_lib):
    """A batch of three lumpy events (2,042, 864 and 1,465 particles) under
    AddressSanitizer: no invalid access in the batch's grid setup, key sort
    and pair walks, and after initialization and 2 steps each event's fields
    are within 1e-12 of the same event run alone, with equal btrack."""
    env = dict(os.environ)
    r = subprocess.run(["bash", os.path.join(EMU, "emu.sh"), os.path.join(EMU, "run_batch.py"), "0.2", "2", "1e-12"],
                       capture_output=True, text=True, timeout=600, env=env)
    print(r.stdout[-4000:], r.stderr[-8000:])
    assert r.returncode == 0 and "AddressSanitizer" not in r.stderr
    assert "OK" in r.stdout


def test_private_columns_disjoint():
    """Every multi-column private member owns its columns (the overlap that
    wrote past the field buffer is gone): parse the enumeration and check
    each member's first column against the previous member's width."""
    import re
    src = open(os.path.join(HERE, "..", "ccake_amd", "csrc", "ccake_internal.h")).read()
    body = src[src.index("enum PrivateField"):]
    body = body[body.index("{") + 1: body.index("};")]
    body = re.sub(r"//[^\n]*", "", body)
    widths = {"P_C_SHV0": 5, "P_FB_E0": 4, "P_FB_G0": 4, "P_C_Q0": 9}
    names, val, cols = {}, None, []
    for item in [x.strip() for x in body.split(",") if x.strip()]:
        if "=" in item:
            nm, expr = [x.strip() for x in item.split("=", 1)]
            for k, v in sorted(names.items(), key=lambda kv: -len(kv[0])):
                expr = re.sub(r"\b%s\b" % k, str(v), expr)
            val = eval(expr.replace("CCAKE_NUM_FIELDS", "0"))
        else:
            nm, val = item, val + 1
        names[nm] = val
        cols.append(nm)
    used = set()
    for nm in cols:
        if nm == "NUM_COLUMNS":
            assert all(c < names[nm] for c in used)
            continue
        span = set(range(names[nm], names[nm] + widths.get(nm, 1)))
        assert not (span & used), f"{nm} overlaps another private column"
        used |= span


def test_bsqhydro_sequence_on_host_emulation(emu_lib, tmp_path):
    """The reference's driver sequence on the C++ call surface
    (ccake_amd/bin/bsqhydro_run) against the host emulation build: the
    phase-by-phase initialisation and 2 steps equal the fused one bit for
    bit, with no invalid memory access."""
    sys.path.insert(0, HERE)
    from test_host_mirror import BSQ, write_bsqhydro_inputs
    write_bsqhydro_inputs(str(tmp_path), 0.1)
    env = dict(os.environ, LD_LIBRARY_PATH=os.path.join(EMU, "_build", "lib"),
               LD_PRELOAD=subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True,
                                         text=True).stdout.strip(),
               ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([BSQ, "--ic", str(tmp_path / "ic.dat"), "--table", str(tmp_path / "table.bin"),
                        "--max-steps", "2", "--output", str(tmp_path)], capture_output=True, text=True,
                       timeout=600, env=env)
    print(r.stdout, r.stderr[-4000:])
    assert r.returncode == 0 and "AddressSanitizer" not in r.stderr
    import json
    d = json.loads(r.stdout.strip().splitlines()[-1])
    assert d["particles"] == 8077 and d["steps_sequence"] == d["steps_fused"] == 2
    assert d["fields_differing"] == 0 and d["int_fields_differing"] == 0 and d["rows_equal"]
