"""Initial-condition readers (BSQHydro::read_ccake, src/BSQHydro.cpp:160-425;
read_ICCING, 83-139) and the system-state text writer
(Output::print_system_state_to_txt, src/output.cpp:102-172).

The C++ readers (ccake_amd/host/ccake_io.hpp, through ccake_amd/bin/ic_dump)
and their Python twins (ccake_amd.ic.read_ccake / read_iccing) are checked
against hand-worked answers on constructed files and against each other, to
the bit, on every flag combination; the writer on the GPU against the
oracle's fields."""
import json
import os
import subprocess

import numpy as np
import pytest

from ccake_amd import config, ic

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DUMP = os.path.join(ROOT, "ccake_amd", "bin", "ic_dump")
RUN = os.path.join(ROOT, "ccake_amd", "bin", "gubser_run")
HC = 0.1973269804
T_FO_DEFAULT = 150.0 / 197.3269804      # the C++ Settings default, fm^-1

CCAKE_TEXT = """#0 0.1 0.2 0.3 0 -5 -6 -0.5
0.5 -0.5 0 1.973269804 0.1 0.2 0.3 0.5 0.6 0 0.01 0.02 0.03 0.04 0.05 0.06 0.07

   # a comment line after the header
1 1 0 0.15 1 1 1 0 0 0 0 0 0 0 0 0 0
2 2 0 0.15000001 1 1 1 0.1 0.2 0 0.5 1 2 3 4 5 6
3 3 0 5.0 7
4 4 0 abc 9 9 9
5.5 6.5 0 2e1 -1 -2 -3 0.25 -0.25 0 0 0 0 0 0 0 0
"""


def dump(path, *flags, kind="ccake"):
    r = subprocess.run([DUMP, "--type", kind, "--file", str(path), *flags], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    grid = tuple(float(v) for v in lines[0].split())
    n = int(lines[1])
    rows = np.array([[float(v) for v in ln.split()] for ln in lines[2:]]).reshape(n, 37)
    return grid, rows


def as_rows(d):
    """The Python reader's output in ic_dump's column order."""
    n = len(d["x"])
    cols = [d["x"], d["y"], d["e"], d["rhoB"], d["rhoS"], d["rhoQ"], d["ux"], d["uy"], d["bulk"]]
    return np.column_stack(cols + [d["shv"][c] for c in range(16)] + [d["diffusion"].reshape(n, 12)[:, k]
                                                                       for k in range(12)])


def test_ccake_reader_known_answers(tmp_path):
    """Hand-worked: e = 1.973269804 GeV/fm^3 is 10 fm^-4; pi^{ij} and Pi are
    divided by hbar c and pi is symmetric; blank and '#' lines are skipped; a
    point at exactly e_cutoff is dropped and one just above it kept; the short
    line "3 3 0 5.0 7" keeps rhoS, rhoQ, u and pi from the line before it (and
    divides pi and Pi by hbar c a second time); the line with "abc" in the
    energy column reads e = 0 and is dropped."""
    f = tmp_path / "ic.dat"
    f.write_text(CCAKE_TEXT)
    d = ic.read_ccake(str(f), e_cutoff_gev=0.15)
    assert d["grid"] == (0.1, 0.2, 0.3, -5.0, -6.0, -0.5)
    assert list(d["x"]) == [0.5, 2.0, 3.0, 5.5] and list(d["y"]) == [-0.5, 2.0, 3.0, 6.5]
    assert abs(d["e"][0] - 10.0) < 1e-14
    assert d["e"][1] == 0.15000001 / HC and d["e"][2] == 5.0 / HC and d["e"][3] == 2e1 / HC
    assert list(d["rhoB"]) == [0.1, 1.0, 7.0, -1.0]
    assert list(d["rhoS"]) == [0.2, 1.0, 1.0, -2.0] and list(d["rhoQ"]) == [0.3, 1.0, 1.0, -3.0]
    assert list(d["ux"]) == [0.5, 0.1, 0.1, 0.25] and list(d["uy"]) == [0.6, 0.2, 0.2, -0.25]
    assert d["bulk"][0] == 0.01 / HC and d["bulk"][1] == 0.5 / HC and d["bulk"][2] == 0.5 / HC / HC
    shv = d["shv"]
    want = {(1, 1): 0.02, (1, 2): 0.03, (1, 3): 0.04, (2, 2): 0.05, (2, 3): 0.06, (3, 3): 0.07}
    for (a, b), v in want.items():
        assert shv[4 * a + b][0] == v / HC and shv[4 * b + a][0] == v / HC
    assert shv[0][0] == 0.0 and shv[1][0] == 0.0
    assert shv[5][2] == 1.0 / HC / HC and shv[15][2] == 6.0 / HC / HC    # re-divided on the short line
    assert np.all(d["diffusion"] == 0)
    # charges switched off; no cutoff keeps the point at 0.15 (the "abc" line has e = 0: still dropped)
    d2 = ic.read_ccake(str(f), e_cutoff_gev=0.0, baryon=False, electric=False)
    assert len(d2["x"]) == 5 and np.all(d2["rhoB"] == 0) and np.all(d2["rhoQ"] == 0)
    assert list(d2["rhoS"]) == [0.2, 1.0, 1.0, 1.0, -2.0]


@pytest.mark.parametrize("flags", [(), ("--e-cutoff", "0"), ("--no-baryon", "--no-strange"),
                                   ("--no-electric", "--e-cutoff", "1.0")])
def test_cpp_and_python_ccake_readers_agree(tmp_path, flags):
    f = tmp_path / "ic.dat"
    f.write_text(CCAKE_TEXT)
    grid, rows = dump(f, *flags)
    kw = dict(e_cutoff_gev=0.15)
    if "--e-cutoff" in flags:
        kw["e_cutoff_gev"] = float(flags[flags.index("--e-cutoff") + 1])
    kw.update(baryon="--no-baryon" not in flags, strange="--no-strange" not in flags,
              electric="--no-electric" not in flags)
    d = ic.read_ccake(str(f), **kw)
    assert grid == d["grid"]
    assert np.array_equal(rows, as_rows(d))


def test_cpp_and_python_readers_agree_with_diffusion_columns(tmp_path):
    """29-column lines (input_initial_diffusion), a strange charge switched
    off, and a malformed number ("1e+") that ends its line."""
    rng = np.random.default_rng(3)
    lines = ["#0 0.05 0.05 0.1 0 -1 -1 -0.1"]
    for k in range(40):
        v = rng.normal(size=29)
        v[3] = abs(v[3]) + 0.2
        lines.append(" ".join("%.17g" % x for x in v))
    lines.append("0.1 0.2 0 3.0 1e+ 5 5")
    lines.append("0.3 0.4 0 4.0 1 2 3 0.1 0.2 0 0.3")
    f = tmp_path / "icd.dat"
    f.write_text("\n".join(lines) + "\n")
    grid, rows = dump(f, "--diffusion", "--no-strange")
    d = ic.read_ccake(str(f), initial_diffusion=True, strange=False)
    assert len(d["x"]) == 42
    assert np.array_equal(rows, as_rows(d))
    assert d["rhoB"][40] == 0.0                        # "1e+" reads as 0 ...
    assert d["rhoQ"][40] == d["rhoQ"][39]              # ... and the rest of the line keeps its values


def test_cpp_and_python_iccing_readers_agree(tmp_path):
    """ICCING: numeric header, fluid at rest, no comment skipping (a blank
    line repeats the previous point with its energy divided by hbar c again)."""
    text = "0 0.1 0.1 0 -3 -3\n0 0 5.0 0.1 0.2 0.3\n1 0 0.1 0 0 0\n\n2 0 2.5 1 1 1\n"
    f = tmp_path / "iccing.dat"
    f.write_text(text)
    grid, rows = dump(f, "--e-cutoff", "0.0", kind="ICCING")
    d = ic.read_iccing(str(f), e_cutoff_gev=0.0)
    assert np.array_equal(rows, as_rows(d))
    assert grid == (0.1, 0.1, 0.0, -3.0, -3.0, 0.0)
    assert list(d["x"]) == [0.0, 1.0, 1.0, 2.0]
    assert d["e"][2] == 0.1 / HC / HC and np.all(d["ux"] == 0)


def test_write_read_roundtrip(tmp_path):
    d = ic.gubser_ideal(spacing=0.25)
    f = tmp_path / "g.dat"
    ic.write_ccake(str(f), d, grid=(0.25, 0.25, 0.1, -5.0, -5.0, -0.1))
    r = ic.read_ccake(str(f), e_cutoff_gev=0.0)
    assert np.array_equal(r["x"], d["x"]) and np.array_equal(r["y"], d["y"])
    assert np.array_equal(r["ux"], d["ux"]) and np.array_equal(r["uy"], d["uy"])
    assert np.max(np.abs(r["e"] / d["e"] - 1)) < 4.5e-16
    assert r["grid"] == (0.25, 0.25, 0.1, -5.0, -5.0, -0.1)
    grid, rows = dump(f, "--e-cutoff", "0")
    assert np.array_equal(rows, as_rows(r))


def test_missing_file_is_reported(tmp_path):
    r = subprocess.run([DUMP, "--type", "ccake", "--file", str(tmp_path / "nope.dat")], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2 and "could not be opened" in r.stderr
    r = subprocess.run([DUMP, "--type", "freezein", "--file", str(tmp_path / "nope.dat")], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2


# ------------------------------------------------------------------ GPU: a file through the C++ call surface
def _oracle_from(orc, d, cfg, steps_at):
    oeos = orc.Eos(0, T_FO_DEFAULT)
    # SPHWorkstation::initialize sets e_freeze from the EoS at T_FO whether or
    # not particlization is on (src/sph_workstation.cpp:29-110)
    cfg = dict(cfg, e_freeze=oeos.efreeze(T_FO_DEFAULT))
    o = orc.Event(cfg, oeos, len(d["x"]))
    for k, fld in [("x", "X"), ("y", "Y"), ("e", "IN_E"), ("ux", "UX"), ("uy", "UY"),
                   ("rhoB", "IN_B"), ("rhoS", "IN_SS"), ("rhoQ", "IN_Q")]:
        o.set("ORC_F_" + fld, d[k])
    assert o.initialize() == 0
    snaps, done = {}, 0
    for s in steps_at:
        while done < s:
            assert o.advance_timestep(cfg["dt"], 2) == 0
            done += 1
        snaps[s] = {f: o.field("ORC_F_" + f).copy() for f in ("X", "Y", "P", "T", "E", "S", "UX", "UY", "GAMMA")}
        snaps[s]["t"] = o.t
        snaps[s]["freeze"] = o.ifield("ORC_I_FREEZE").copy()
    return snaps


@pytest.mark.gpu
def test_ccake_file_through_cpp_call_surface_and_system_state_files(tmp_path, orc):
    """The ideal Gubser lattice (0.1 fm) written as a ccake file in the
    format of the reference's IC generator, read by read_ccake inside the
    C++ driver, evolved 20 steps with system_state_<k>.dat every 10 steps:
    the physics matches the built-in lattice run; each file holds the time
    line and one 46-column line per particle (time in the default format on
    the first particle line, scientific after it; EoS name "default"), and
    its columns equal the oracle's fields at that step to the 6 significant
    digits written."""
    d = ic.gubser_ideal(spacing=0.1)
    f = tmp_path / "gubser.dat"
    ic.write_ccake(str(f), d, grid=(0.1, 0.1, 0.1, -5.0, -5.0, -0.1))
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run([RUN, "--ic", str(f), "--e-cutoff", "0", "--steps", "20", "--txt-evolution", "10",
                        "--output", str(out)], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    got = json.loads(r.stdout.strip().splitlines()[-1])
    r2 = subprocess.run([RUN, "--spacing", "0.1", "--steps", "20"], capture_output=True, text=True, timeout=300)
    assert r2.returncode == 0, r2.stderr
    ref = json.loads(r2.stdout.strip().splitlines()[-1])
    assert got["particles"] == ref["particles"] == len(d["x"])
    for k in ("t", "max_rel_err_e_r_lt_2.5", "entropy_rel_change", "energy_rel_change"):
        assert abs(got[k] - ref[k]) <= 1e-9 * max(1.0, abs(ref[k])), k

    cfg = config.gubser_ideal_params()
    dd = ic.read_ccake(str(f), e_cutoff_gev=0.0)
    snaps = _oracle_from(orc, dd, cfg, [0, 10, 20])
    n = len(d["x"])
    for k, step in enumerate([0, 10, 20]):
        lines = (out / f"system_state_{k}.dat").read_text().split("\n")
        assert lines[-1] == "" and len(lines) == n + 2
        t = float(lines[0])
        assert abs(t - snaps[step]["t"]) < 1e-5
        toks = [ln.split(" ") for ln in lines[1:-1]]
        assert all(len(tk) == 47 and tk[-1] == "" for tk in toks)        # 46 columns, trailing space
        assert "e+" not in toks[0][1] and "e+" in toks[1][1]             # default format, then scientific
        assert all(tk[38] == "default" for tk in toks)
        a = np.array([[float(v) for i, v in enumerate(tk[:-1]) if i != 38] for tk in toks])
        s = snaps[step]
        want = {2: s["X"], 3: s["Y"], 5: s["P"] * HC, 6: s["T"] * HC, 10: s["E"] * HC, 14: s["S"],
                33: s["UX"], 34: s["UY"], 36: s["GAMMA"]}
        for col, w in want.items():
            assert np.all(np.abs(a[:, col] - w) <= 6e-6 * np.abs(w) + 1e-12 * np.max(np.abs(w))), (k, col)
        assert np.array_equal(a[:, 37].astype(int), s["freeze"])
        assert np.all(a[:, [4, 35]] == 0) and np.all(a[:, 20:25] == 0)   # eta components; no shear (ideal)
