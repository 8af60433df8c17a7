"""HDF5 input and output of the C++ call surface (libccake_hdf5.so).

1. The flat (s, nB) pre-inverted table in the reference's HDF5 layout
   (EoS_Interpolator::load_flat_table, src/eos_interpolator.cpp:291-527):
   a file written here with the HDF5 C library reads back bit for bit, and
   malformed files are refused with a message.
2. The system-state file of Output::print_system_state_to_HDF
   (src/output.cpp:213-266, include/interface_to_HDF5.h): the Gubser driver
   writes system_state.h5 and the text files system_state_<k>.dat in the
   same run; the file has the reference's groups, attributes, datasets,
   units and labels, and every dataset equals the matching text column (to
   the text's 7 significant digits; T, mu and e are in MeV there and in GeV
   in the text file, as in the reference) -- on the host emulation build of
   the library (no GPU) and on the GPU.
"""
import os
import subprocess

import numpy as np
import pytest

import h5util
from ccake_amd import eos_tables

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
PROG = os.path.join(ROOT, "ccake_amd", "bin", "gubser_run")
SHIM = os.path.join(ROOT, "ccake_amd", "libccake_hdf5.so")
EMU = os.path.join(HERE, "hostemu")

pytestmark = pytest.mark.skipif(not h5util.available(), reason="no HDF5 library in this image")


def flat_arrays(Ns=7, NB=5, seed=1):
    rng = np.random.default_rng(seed)
    s_ax = 10.0 ** np.linspace(0.0, 3.0, Ns)          # MeV^3, uniform in log10
    nb_ax = 10.0 ** np.linspace(-2.0, 1.0, NB)
    S, B = np.meshgrid(s_ax, nb_ax, indexing="ij")
    arrays = [S, B] + [rng.uniform(1.0, 100.0, size=(Ns, NB)) for _ in range(10)]
    return arrays


def test_flat_table_reads_back_bit_for_bit(tmp_path):
    assert os.path.exists(SHIM), "build() has not produced ccake_amd/libccake_hdf5.so"
    arrays = flat_arrays()
    path = str(tmp_path / "flat.h5")
    h5util.write_flat_table(path, 7, 5, arrays)
    table, (Ns, NB) = eos_tables.load_flat_table_h5(path)
    assert (Ns, NB) == (7, 5)
    want = np.concatenate([a.ravel() for a in arrays])
    assert np.array_equal(table, want)


def test_flat_table_refusals(tmp_path):
    bad_rank = str(tmp_path / "rank.h5")
    h5util.write_flat_table(bad_rank, 7, 5, flat_arrays(), dims_shape=3)
    with pytest.raises(ValueError, match="dimensions"):
        eos_tables.load_flat_table_h5(bad_rank)
    with pytest.raises(ValueError, match="open"):
        eos_tables.load_flat_table_h5(str(tmp_path / "missing.h5"))
    missing = str(tmp_path / "missing_chi.h5")
    h5util.write_flat_table(missing, 7, 5, flat_arrays()[:11])       # no /data/chiBB
    with pytest.raises(ValueError, match="chiBB"):
        eos_tables.load_flat_table_h5(missing)


def read_text_state(path):
    lines = open(path).read().splitlines()
    t = float(lines[0])
    rows = np.array([[float(v) for v in ln.split()[:15]] for ln in lines[1:]])
    return t, rows


def check_state_file(out):
    f = h5util.File(str(out / "system_state.h5"))
    try:
        assert f.double_attribute("/Event", "h") == 0.3
        assert f.double_attribute("/Event", "e_cutoff") == 0.15
        frames = sorted(f.members("/Event"))
        # max_tau / dt = 50 / 0.005: four digits
        assert frames == ["Frame_0000", "Frame_0001", "Frame_0002"], frames
        names = ["x", "y", "T", "muB", "muS", "muQ", "e", "s", "B", "S", "Q"]
        units = ["fm", "fm", "MeV", "MeV", "MeV", "MeV", "MeV/fm^3", "1/fm^3", "1/fm^3", "1/fm^3", "1/fm^3"]
        text_col = {"x": 2, "y": 3, "T": 6, "muB": 7, "muS": 8, "muQ": 9, "e": 10, "B": 11, "S": 12, "Q": 13, "s": 14}
        for k, fr in enumerate(frames):
            g = "/Event/" + fr
            assert sorted(f.members(g)) == sorted(names + ["labels"])
            t_txt, rows = read_text_state(str(out / f"system_state_{k}.dat"))
            assert f.double_attribute(g, "Time") == pytest.approx(t_txt, rel=1e-6)
            assert f.double_attribute(g, "e_2_X") == -10.0 and f.double_attribute(g, "e_2_P") == -10.0
            for name, unit in zip(names, units):
                assert f.units(f"{g}/{name}") == unit
                a = f.dataset(f"{g}/{name}")
                # the text file prints T, mu and e in GeV units (hbar c in GeV fm),
                # the HDF5 file in MeV (hbar c in MeV fm), as the reference does
                b = rows[:, text_col[name]] * (1000.0 if unit.startswith("MeV") else 1.0)
                assert a.shape == b.shape
                scale = max(float(np.max(np.abs(b))), 1e-300)
                assert np.all(np.abs(a - b) <= 1e-6 * np.maximum(np.abs(b), 1e-6 * scale)), name
            labels = f.dataset(f"{g}/labels", dtype=np.int32)
            assert labels.shape == (rows.shape[0],) and np.all(labels == 0)   # the host name "default" maps to 0
    finally:
        f.close()


def run_driver(out, env=None):
    r = subprocess.run([PROG, "--spacing", "0.2", "--steps", "4", "--txt-evolution", "2", "--hdf-evolution", "2",
                        "--output", str(out)], capture_output=True, text=True, timeout=600, env=env)
    print(r.stdout[-2000:], r.stderr[-4000:])
    return r


def test_system_state_h5_on_host_emulation(tmp_path):
    """The driver against the host emulation build of the library (CPU)."""
    import sys
    sys.path.insert(0, EMU)
    import build_emu
    build_emu.build(asan=False)
    env = dict(os.environ, LD_LIBRARY_PATH=os.path.join(EMU, "_build", "lib_noasan"))
    r = run_driver(tmp_path, env)
    assert r.returncode == 0, r.stderr
    check_state_file(tmp_path)


@pytest.mark.gpu
def test_system_state_h5_on_gpu(tmp_path):
    r = run_driver(tmp_path)
    assert r.returncode == 0, r.stderr
    check_state_file(tmp_path)
