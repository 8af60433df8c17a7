"""Parity of the HIP library against the oracle on the same inputs.

Neighbour sets and counts are compared bit-exactly; FP64 fields to
1e-12 relative, measured as max|gpu - oracle| / max|oracle| over the
particles of a field (summation orders differ between the two, so the
fields agree to rounding, not bit for bit).  Every test calls the HIP path
through the C-ABI (ccake_amd -> libccake_amd.so).
"""
import json
import os

import numpy as np
import pytest

import ccake_amd
from ccake_amd import config, eos_tables, ic
from parity_metrics import DERIVS, STATE, TRANSPORT, field_err, pp_tol, whole_field

pytestmark = pytest.mark.gpu

TOL = 1e-12
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))
T_FO = 150.0 / config.HBARC_MEVFM

# ------------------------------------------------------------------ helpers
_FIDS = sorted((v, k[len("CCAKE_F_"):]) for k, v in ccake_amd.FIELDS.items() if k != "CCAKE_NUM_FIELDS")


def field_names():
    """(fid, label) for every field column, multi-component ones expanded."""
    out = []
    for i, (v, k) in enumerate(_FIDS):
        nxt = _FIDS[i + 1][0] if i + 1 < len(_FIDS) else ccake_amd.FIELDS["CCAKE_NUM_FIELDS"]
        for c in range(nxt - v):
            out.append((v + c, k if nxt - v == 1 else f"{k}+{c}"))
    return out


def fid(label):
    if "+" in label:
        b, c = label.split("+")
        return ccake_amd.FIELDS["CCAKE_F_" + b] + int(c)
    return ccake_amd.FIELDS["CCAKE_F_" + label]


def rel(a, b):
    """max|a - b| / max|b| over a field (the coarse whole-field measure)."""
    s = np.max(np.abs(b)) if b.size else 0.0
    d = np.max(np.abs(a - b)) if b.size else 0.0
    if s == 0.0:
        return float(d)
    return float(d / s)


def compare(gev, oev, labels, tol=TOL, what=""):
    """Assert every labelled field agrees with the oracle twice over: the
    whole-field error max|a-b| / max|b| within tol, and the per-particle
    error (parity_metrics.field_err) within pp_tol; print the worst first."""
    errs, pps = {}, {}
    for lab in labels:
        a, b = gev.download(fid(lab)), oev.field(fid(lab))
        pps[lab] = field_err(lab, a, b, tol)
        fin = np.isfinite(b)
        errs[lab] = whole_field(a[fin], b[fin]) if np.array_equal(np.isfinite(a), fin) else np.inf
    bad = {k: v for k, v in errs.items() if not v <= tol}
    bad_pp = {k: v for k, v in pps.items() if not v <= pp_tol(k, tol)}
    print(what, "worst whole-field:", max(errs.items(), key=lambda kv: kv[1]) if errs else None,
          "worst per-particle:", max(pps.items(), key=lambda kv: kv[1]) if pps else None, flush=True)
    assert not bad, f"{what}: fields beyond {tol} (whole field): {bad}"
    assert not bad_pp, f"{what}: fields beyond their per-particle tolerance: {bad_pp}"
    return errs


def compare_lumpy(p, labels, what=""):
    return compare(p.g, p.o, labels, what=what)


def report_all(gev, oev, what):
    """Print the per-particle error of every field beyond TOL (diagnostic only)."""
    lines = []
    for f, lab in field_names():
        a = gev.download(f)
        b = oev.field(f)
        with np.errstate(all="ignore"):
            e = field_err(lab, a, b, TOL)
        if e > pp_tol(lab, TOL):
            lines.append(f"{lab}={e:.2e}")
    print(f"[{what}] fields beyond {TOL}: {' '.join(lines) if lines else 'none'}", flush=True)


def ints_equal(gev, oev, name):
    a = gev.download_int("CCAKE_I_" + name)
    b = oev.ifield("ORC_I_" + name)
    assert np.array_equal(a, b), f"{name}: {int(np.sum(a != b))} particles differ"


def oracle_load(oev, d):
    for k, f in [("x", "X"), ("y", "Y"), ("e", "IN_E"), ("ux", "UX"), ("uy", "UY"),
                 ("rhoB", "IN_B"), ("rhoS", "IN_SS"), ("rhoQ", "IN_Q")]:
        oev.set("ORC_F_" + f, d[k])


def pair_key(pi, pj):
    return pi.astype(np.int64) * (1 << 32) + pj.astype(np.int64)


class Pair:
    """The same system on the GPU and on the oracle."""

    def __init__(self, dev, orc, cfg, d, eos_kind=0, table=None):
        n = len(d["x"])
        if eos_kind == 0:
            self.geos = ccake_amd.Eos(dev, 0, T_FO)
            self.oeos = orc.Eos(0, T_FO)
        else:
            tab, dims, mins, maxs = table
            self.geos = ccake_amd.Eos(dev, 1, T_FO, dims, mins, maxs, tab)
            self.oeos = orc.Eos(1, T_FO, dims, mins, maxs, tab)
        if cfg["particlization"]:
            eg, eo = self.geos.efreeze(T_FO), self.oeos.efreeze(T_FO)
            assert abs(eg - eo) <= 1e-13 * abs(eo)
            cfg = dict(cfg, e_freeze=eo)
        self.cfg = cfg
        self.g = ccake_amd.Event(dev, self.geos, cfg, n)
        ccake_amd.load_event(self.g, d)
        self.o = orc.Event(cfg, self.oeos, n)
        oracle_load(self.o, d)

    def initialize(self):
        self.g.initialize()
        assert self.o.initialize() == 0

    def step(self, k=1):
        for _ in range(k):
            self.g.advance_timestep(self.cfg["dt"], 2)
            assert self.o.advance_timestep(self.cfg["dt"], 2) == 0

    def close(self):
        self.g.close()
        self.geos.close()


def lumpy_small(spacing=0.1):
    return ic.lumpy(spacing=spacing)


@pytest.fixture(scope="module")
def table():
    return eos_tables.tanh_conformal_table()


# ------------------------------------------------------------------ KATs
def test_smoothing_kat_on_gpu(dev):
    k = GOLD["smoothing"]
    eos = ccake_amd.Eos(dev, 0)
    ev = ccake_amd.Event(dev, eos, config.base_params(hT=k["hT"]), 4)
    ev.upload("CCAKE_F_X", k["x"]); ev.upload("CCAKE_F_Y", k["y"])
    for X, f in (("s", "S"), ("rhoB", "B"), ("rhoQ", "Q"), ("rhoS", "SS")):
        ev.upload("CCAKE_F_MASS_" + f, k["mass"][X]); ev.upload("CCAKE_F_EXT_" + f, k["extensive"][X])
    ev.t = 1.0
    ev.build_neighbors()
    ev.smooth_fields()
    e = k["expected_p0"]
    assert abs(ev.download("CCAKE_F_SM_S")[0] - e["s"]) < 1e-14
    assert abs(ev.download("CCAKE_F_SM_B")[0] - e["rhoB"]) < 1e-14
    assert abs(ev.download("CCAKE_F_SM_Q")[0] - e["rhoQ"]) < 1e-14
    assert abs(ev.download("CCAKE_F_SM_SS")[0] - e["rhoS"]) < 1e-13
    assert abs(ev.download("CCAKE_F_SIGMA_LAB")[0] - e["sigma_lab"]) < 1e-14
    ev.close(); eos.close()


def test_gradient_kat_on_gpu(dev):
    k = GOLD["grad_smoothing"]
    eos = ccake_amd.Eos(dev, 0)
    ev = ccake_amd.Event(dev, eos, config.base_params(hT=k["hT"], using_shear=0), 5)
    ev.upload("CCAKE_F_X", k["x"]); ev.upload("CCAKE_F_Y", k["y"])
    ev.upload("CCAKE_F_SIGMA_LAB", 1.0); ev.upload("CCAKE_F_P", k["p"])
    ev.build_neighbors()
    ev.smooth_gradients()
    g = ev.download("CCAKE_F_GRADP0"), ev.download(fid("GRADP0+1"))
    assert abs(g[0][0] - k["expected_gradP_p0"][0]) < 1e-14
    assert abs(g[1][0] - k["expected_gradP_p0"][1]) < 1e-14
    ev.close(); eos.close()


# ------------------------------------------------------------------ neighbours
@pytest.mark.parametrize("spacing", [0.1, 0.05])
def test_pairs_bit_exact_on_lattice(dev, orc, spacing):
    d = ic.gubser_ideal(spacing=spacing)
    n = len(d["x"])
    cfg = config.gubser_ideal_params()
    eos = ccake_amd.Eos(dev, 0)
    ev = ccake_amd.Event(dev, eos, cfg, n)
    ev.upload("CCAKE_F_X", d["x"]); ev.upload("CCAKE_F_Y", d["y"])
    ev.build_neighbors()
    gi, gj = ev.pairs()
    oeos = orc.Eos(0)
    o = orc.Event(cfg, oeos, n)
    o.set("ORC_F_X", d["x"]); o.set("ORC_F_Y", d["y"])
    o.build_neighbors(True)                       # brute force
    oi, oj = o.pairs()
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj)
    assert np.array_equal(ev.download_int("CCAKE_I_BTRACK"), o.ifield("ORC_I_BTRACK"))
    assert ev.last_pair_count() == len(oi)
    if spacing == 0.05:
        centre = int(np.argmin(d["x"] ** 2 + d["y"] ** 2))
        assert ev.download_int("CCAKE_I_BTRACK")[centre] == 440
    ev.close(); eos.close()


def test_pairs_bit_exact_random_and_boundary(dev, orc):
    rng = np.random.default_rng(3)
    n = 3000
    x = rng.uniform(-3, 3, n); y = rng.uniform(-3, 3, n)
    x[:10] = np.arange(10) * 0.6; y[:10] = 5.0       # exact 2h separations
    x[10:20] = 1.0; y[10:20] = 1.0                   # coincident particles
    cfg = config.base_params(hT=0.3)
    eos = ccake_amd.Eos(dev, 0)
    ev = ccake_amd.Event(dev, eos, cfg, n)
    ev.upload("CCAKE_F_X", x); ev.upload("CCAKE_F_Y", y)
    ev.build_neighbors()
    gi, gj = ev.pairs()
    oeos = orc.Eos(0)
    o = orc.Event(cfg, oeos, n)
    o.set("ORC_F_X", x); o.set("ORC_F_Y", y)
    o.build_neighbors(True)
    oi, oj = o.pairs()
    assert np.array_equal(pair_key(gi, gj), pair_key(oi, oj))
    ev.close(); eos.close()


@pytest.mark.parametrize("cells_per_rc,x_subcells", [(1, 1), (2, 1), (3, 1), (3, 8), (4, 3), (4, 8)])
def test_pairs_independent_of_cell_size(dev, orc, cells_per_rc, x_subcells):
    """The pair set does not depend on the cell size along y (key rows) or
    on how finely the rows are cut along x."""
    d = lumpy_small(0.05)
    n = len(d["x"])
    rng = np.random.default_rng(11)
    x = d["x"] + rng.uniform(-0.01, 0.01, n); y = d["y"] + rng.uniform(-0.01, 0.01, n)
    cfg = config.base_params(hT=0.3, cells_per_rc=cells_per_rc, x_subcells=x_subcells)
    eos = ccake_amd.Eos(dev, 0)
    ev = ccake_amd.Event(dev, eos, cfg, n)
    ev.upload("CCAKE_F_X", x); ev.upload("CCAKE_F_Y", y)
    ev.build_neighbors()
    gi, gj = ev.pairs()
    oeos = orc.Eos(0)
    o = orc.Event(cfg, oeos, n)
    o.set("ORC_F_X", x); o.set("ORC_F_Y", y)
    o.build_neighbors(False)
    oi, oj = o.pairs()
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj)
    ev.close(); eos.close()


def test_pairs_full_size_lattice(dev, orc):
    """251,001 particles on the 0.02 fm lattice: every row bit-exact."""
    d = ic.lumpy()
    n = len(d["x"])
    assert n == 251001
    cfg = config.lumpy_params()
    eos = ccake_amd.Eos(dev, 0)
    ev = ccake_amd.Event(dev, eos, cfg, n)
    ev.upload("CCAKE_F_X", d["x"]); ev.upload("CCAKE_F_Y", d["y"])
    ev.build_neighbors()
    gi, gj = ev.pairs()
    oeos = orc.Eos(0)
    o = orc.Event(cfg, oeos, n)
    o.set("ORC_F_X", d["x"]); o.set("ORC_F_Y", d["y"])
    o.build_neighbors(False)
    oi, oj = o.pairs()
    assert len(gi) == len(oi)
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj)
    ev.close(); eos.close()


@pytest.mark.parametrize("skin", ["0.04", "0.5"])
def test_pairs_bit_exact_on_a_kept_sort(dev, orc, monkeypatch, skin):
    """Time steps keep the last cell sort while the displacement bound
    allows, widening the search radius by it.  After 8 steps of Gubser flow
    (0.1 fm lattice; at skin 0.5 rc the sort from initialize is never
    renewed) the exported pair set equals a brute-force build on the current
    positions, bit for bit.  (btrack holds the counts of the last stage's
    positions, as in the reference, so it is not compared here.)"""
    monkeypatch.setenv("CCAKE_SORT_SKIN", skin)
    cfg = config.gubser_ideal_params()
    d = ic.gubser_ideal(spacing=0.1)
    n = len(d["x"])
    eos = ccake_amd.Eos(dev, 0)
    ev = ccake_amd.Event(dev, eos, cfg, n)
    ccake_amd.load_event(ev, d)
    ev.initialize()
    assert ev.run_steps(cfg["dt"], 8) == 8
    x, y = ev.download("CCAKE_F_X"), ev.download("CCAKE_F_Y")
    assert np.max(np.hypot(x - d["x"], y - d["y"])) > 0.02      # particles left their sorted cells' positions
    gi, gj = ev.pairs()
    oeos = orc.Eos(0)
    o = orc.Event(cfg, oeos, n)
    o.set("ORC_F_X", x); o.set("ORC_F_Y", y)
    o.build_neighbors(True)                       # brute force
    oi, oj = o.pairs()
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj)
    ev.close(); eos.close()


# ------------------------------------------------------------------ ideal Gubser
@pytest.fixture(scope="module")
def gubser(dev, orc):
    d = ic.gubser_ideal(spacing=0.05)
    p = Pair(dev, orc, config.gubser_ideal_params(), d)
    yield p
    p.close()


def test_gubser_initialize_parity(gubser):
    gubser.initialize()
    report_all(gubser.g, gubser.o, "gubser init")
    compare(gubser.g, gubser.o, STATE + ["IN_E", "IN_S", "MASS_S"], what="gubser init")
    ints_equal(gubser.g, gubser.o, "BTRACK")
    ints_equal(gubser.g, gubser.o, "FREEZE")


def test_gubser_time_derivatives_parity(gubser):
    gubser.g.get_time_derivatives()
    assert gubser.o.get_time_derivatives() == 0
    report_all(gubser.g, gubser.o, "gubser derivatives")
    compare(gubser.g, gubser.o, STATE + TRANSPORT + DERIVS, what="gubser derivatives")


def test_gubser_steps_parity(gubser):
    gubser.step(5)
    assert abs(gubser.g.t - gubser.o.t) < 1e-14
    report_all(gubser.g, gubser.o, "gubser 5 steps")
    compare(gubser.g, gubser.o, STATE + DERIVS + ["EZ"], what="gubser 5 steps")
    ints_equal(gubser.g, gubser.o, "BTRACK")


def test_run_steps_equals_single_steps_and_the_oracle(dev, orc):
    """ccake_run_steps begins each kept-sort stage inside the previous
    stage's EoM kernel (gamma, v, smoothing records); single calls to
    ccake_advance_timestep begin only the second stage that way.  Both give
    the same state bit for bit after 12 steps of ideal Gubser flow (sorts and
    kept sorts both occur), and the run equals the oracle to 1e-12."""
    cfg = config.gubser_ideal_params()
    d = ic.gubser_ideal(spacing=0.1)
    n = len(d["x"])
    eos = ccake_amd.Eos(dev, 0)
    a = ccake_amd.Event(dev, eos, cfg, n)
    b = ccake_amd.Event(dev, eos, cfg, n)
    for ev in (a, b):
        ccake_amd.load_event(ev, d)
        ev.initialize()
    assert a.run_steps(cfg["dt"], 12) == 12
    for _ in range(12):
        b.advance_timestep(cfg["dt"], 2)
    for f, lab in field_names():
        assert np.array_equal(a.download(f), b.download(f)), lab
    assert np.array_equal(a.download_int("CCAKE_I_BTRACK"), b.download_int("CCAKE_I_BTRACK"))
    oeos = orc.Eos(0, T_FO)
    o = orc.Event(cfg, oeos, n)
    oracle_load(o, d)
    assert o.initialize() == 0
    for _ in range(12):
        assert o.advance_timestep(cfg["dt"], 2) == 0
    compare(a, o, STATE + DERIVS + ["EZ"], what="run_steps 12")
    a.close(); b.close(); eos.close()


def test_gubser_physics_and_conservation(dev):
    """120 RK2 steps (tau 1.0 -> 1.6): the energy density follows the
    analytic ideal Gubser solution within the SPH smoothing error present
    at tau0, extensive entropy is unchanged to the bit, and E + E_z drifts
    by less than 1 percent."""
    cfg = config.gubser_ideal_params()
    d = ic.gubser_ideal(spacing=0.05)
    eos = ccake_amd.Eos(dev, 0)
    ev = ccake_amd.Event(dev, eos, cfg, len(d["x"]))
    ccake_amd.load_event(ev, d)
    ev.initialize()
    s0 = ev.download("CCAKE_F_EXT_S")
    E0 = ev.reduce("CCAKE_R_ENERGY") + ev.reduce("CCAKE_R_EZ")
    assert ev.run_steps(cfg["dt"], 120) == 120
    assert abs(ev.t - 1.6) < 1e-9
    assert np.array_equal(ev.download("CCAKE_F_EXT_S"), s0)
    E1 = ev.reduce("CCAKE_R_ENERGY") + ev.reduce("CCAKE_R_EZ")
    assert abs(E1 - E0) / E0 < 0.01
    r = np.hypot(ev.download("CCAKE_F_X"), ev.download("CCAKE_F_Y"))
    ea = ic.gubser_energy(ev.t, r)
    m = r < 2.5
    e = ev.download("CCAKE_F_E")
    assert np.max(np.abs(e[m] - ea[m]) / ea[m]) < 0.07
    ev.close(); eos.close()


# ------------------------------------------------------------------ viscous BSQ
def test_viscous_conformal_parity(dev, orc):
    """Lumpy BSQ event (0.1 fm lattice, 10,201 particles), conformal EoS
    with charges, shear and bulk on: initialize, derivatives, 3 steps."""
    d = lumpy_small(0.1)
    p = Pair(dev, orc, config.lumpy_params(), d)
    p.initialize()
    report_all(p.g, p.o, "viscous conformal init")
    compare_lumpy(p, STATE + ["IN_E", "IN_S", "MASS_S", "MASS_B", "MASS_SS", "MASS_Q"], what="viscous init")
    ints_equal(p.g, p.o, "FREEZE")
    p.g.get_time_derivatives()
    assert p.o.get_time_derivatives() == 0
    report_all(p.g, p.o, "viscous conformal derivatives")
    compare_lumpy(p, STATE + TRANSPORT + DERIVS, what="viscous derivatives")
    p.step(3)
    report_all(p.g, p.o, "viscous conformal 3 steps")
    compare_lumpy(p, STATE + DERIVS, what="viscous 3 steps")
    ints_equal(p.g, p.o, "FREEZE")
    p.close()


def oracle_twin(p, orc, d, scale):
    """A second oracle run of the Pair's event with its initial energy density
    scaled by `scale`: the distance between the two oracle runs measures how
    much the flow amplifies a perturbation of rounding size."""
    o2 = orc.Event(p.cfg, p.oeos, len(d["x"]))
    oracle_load(o2, dict(d, e=d["e"] * scale))
    assert o2.initialize() == 0
    return o2


def test_table_eos_parity(dev, orc, table):
    """Lumpy BSQ event on the 64x32x32x32 table with its fallback chain,
    every lattice point kept (0.1 fm lattice, 10,201 particles; the corner
    particles sit on the analytic members of the chain): initialize and the
    first derivatives within 1e-12; then 20 steps, compared every 5 steps.
    Over the steps the near-vacuum corners amplify rounding differences, so
    each field is held to ten times the distance between two oracle runs
    whose initial energy densities differ by a factor 1 + 1e-15 (a
    perturbation that moves the initialized fields as much as the GPU's
    summation order does, ~4e-14), and never looser than 1e-12."""
    d = lumpy_small(0.1)
    p = Pair(dev, orc, config.lumpy_params(), d, eos_kind=1, table=table)
    p.initialize()
    o2 = oracle_twin(p, orc, d, 1.0 + 1e-15)
    report_all(p.g, p.o, "table init")
    ints_equal(p.g, p.o, "EOS_ID")
    compare_lumpy(p, STATE + ["IN_E", "IN_S", "MASS_S", "MASS_B", "MASS_SS", "MASS_Q"], what="table init")
    p.g.get_time_derivatives()
    assert p.o.get_time_derivatives() == 0
    report_all(p.g, p.o, "table derivatives")
    compare_lumpy(p, STATE + TRANSPORT + DERIVS, what="table derivatives")
    members = [np.bincount(p.o.ifield("ORC_I_EOS_ID"), minlength=4)]
    worst = []
    for k in range(4):
        p.step(5)
        for _ in range(5):
            assert o2.advance_timestep(p.cfg["dt"], 2) == 0
        errs = {}
        for lab in STATE + DERIVS:
            b = p.o.field(fid(lab))
            env = whole_field(o2.field(fid(lab)), b)
            err = whole_field(p.g.download(fid(lab)), b)
            errs[lab] = (err, max(TOL, 10.0 * env))
        bad = {k2: v for k2, v in errs.items() if not v[0] <= v[1]}
        top = max(errs.items(), key=lambda kv: kv[1][0])
        worst.append((5 * (k + 1), top[0], top[1][0], top[1][1]))
        print(f"table {5 * (k + 1)} steps: worst {top[0]} {top[1][0]:.2e} (bound {top[1][1]:.2e})", flush=True)
        assert not bad, f"table {5 * (k + 1)} steps: beyond ten times the oracle's own spread: {bad}"
        ints_equal(p.g, p.o, "FREEZE")
        ints_equal(p.g, p.o, "EOS_ID")
        members.append(np.bincount(p.o.ifield("ORC_I_EOS_ID"), minlength=4))
    print("particles per chain member (table, tanh, conformal, diagonal) every 5 steps:",
          [m.tolist() for m in members], flush=True)
    p.close()


def test_rk4_parity(dev, orc):
    """The reference's fourth-order step (src/evolver.cpp:519-561) on the
    viscous lumpy event (conformal EoS with charges, shear and bulk; 10,201
    particles): 3 steps on the GPU equal the oracle's, and the extensive
    bulk pressure is in the bulk field after each step, as update_rk4 leaves
    it."""
    d = lumpy_small(0.1)
    p = Pair(dev, orc, config.lumpy_params(), d)
    p.initialize()
    for _ in range(3):
        p.g.advance_timestep(p.cfg["dt"], 4)
        assert p.o.advance_timestep(p.cfg["dt"], 4) == 0
    assert abs(p.g.t - p.o.t) < 1e-15
    report_all(p.g, p.o, "rk4 3 steps")
    compare_lumpy(p, STATE + DERIVS, what="rk4 3 steps")
    assert np.array_equal(p.g.download("CCAKE_F_BULK"), p.g.download("CCAKE_F_EXT_BULK"))
    ints_equal(p.g, p.o, "FREEZE")
    p.close()


def test_table_eos_long_run_parity(dev, orc, table):
    """The lumpy event on the table EoS with the reference's default energy
    cutoff (0.15 GeV/fm^3; 8,077 particles on the 0.1 fm lattice): 60 steps,
    compared every 20 steps within 1e-12 (whole field) and per particle
    (the oracle moves by 1e-14 under a 1e-14 perturbation here: no
    amplification without the near-vacuum corners), with equal freeze-out
    status and chain members."""
    d = ic.lumpy(spacing=0.1, e_cutoff_gev=0.15)
    p = Pair(dev, orc, config.lumpy_params(), d, eos_kind=1, table=table)
    p.initialize()
    compare_lumpy(p, STATE, what="cut table init")
    for k in range(3):
        p.step(20)
        report_all(p.g, p.o, f"cut table {20 * (k + 1)} steps")
        compare_lumpy(p, STATE + DERIVS, what=f"cut table {20 * (k + 1)} steps")
        ints_equal(p.g, p.o, "FREEZE")
        ints_equal(p.g, p.o, "EOS_ID")
    p.close()


def test_full_size_table_event_properties(dev, orc, table):
    """The 251,001-particle lumpy event on the table EoS: initialization
    matches the oracle; then 5 steps are deterministic (two runs agree to
    the bit) and conserve the extensive charges exactly (no diffusion)."""
    d = ic.lumpy()
    p = Pair(dev, orc, config.lumpy_params(), d, eos_kind=1, table=table)
    p.initialize()
    report_all(p.g, p.o, "full-size table init")
    compare_lumpy(p, STATE, what="full-size init")
    ints_equal(p.g, p.o, "EOS_ID")
    B0 = p.g.download("CCAKE_F_EXT_B")
    g2 = ccake_amd.Event(dev, p.geos, p.cfg, len(d["x"]))
    ccake_amd.load_event(g2, d)
    g2.initialize()
    p.g.run_steps(p.cfg["dt"], 5)
    g2.run_steps(p.cfg["dt"], 5)
    for f in ("X", "UX", "EXT_S", "E", "T", "SHV0+5", "BULK"):
        assert np.array_equal(p.g.download(fid(f)), g2.download(fid(f))), f
    assert np.array_equal(p.g.download("CCAKE_F_EXT_B"), B0)
    g2.close()
    p.close()


# ------------------------------------------------------------------ edge cases
def test_invalid_inputs_rejected(dev):
    """Empty events and non-finite positions are refused on the host."""
    eos = ccake_amd.Eos(dev, 0)
    with pytest.raises(ccake_amd.CcakeError):
        ccake_amd.Event(dev, eos, config.base_params(), 0)
    ev = ccake_amd.Event(dev, eos, config.base_params(), 10)
    x = np.zeros(10); x[3] = np.nan
    with pytest.raises(ccake_amd.CcakeError):
        ev.upload("CCAKE_F_X", x)
    ev.close(); eos.close()


def test_pairs_bit_exact_on_coarsened_grid(dev, orc):
    """A domain too wide for the 65,536-cell cap (cells are doubled until it
    fits): clusters far apart, pairs still bit-exact."""
    rng = np.random.default_rng(21)
    c = rng.uniform(-400, 400, size=(40, 2))
    pts = np.concatenate([cc + rng.uniform(-1.5, 1.5, size=(60, 2)) for cc in c])
    x, y = pts[:, 0].copy(), pts[:, 1].copy()
    n = len(x)
    cfg = config.base_params(hT=0.3)
    eos = ccake_amd.Eos(dev, 0)
    ev = ccake_amd.Event(dev, eos, cfg, n)
    ev.upload("CCAKE_F_X", x); ev.upload("CCAKE_F_Y", y)
    ev.build_neighbors()
    gi, gj = ev.pairs()
    oeos = orc.Eos(0)
    o = orc.Event(cfg, oeos, n)
    o.set("ORC_F_X", x); o.set("ORC_F_Y", y)
    o.build_neighbors(True)
    oi, oj = o.pairs()
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj)
    ev.close(); eos.close()


def test_sparse_event_with_isolated_particles(dev, orc):
    """A sparse viscous event where some particles have no neighbours: the
    freeze-out machine marks them (Freeze 4, btrack -1) identically, and the
    fields agree after 3 steps."""
    rng = np.random.default_rng(5)
    n = 600
    x = rng.uniform(-6, 6, n); y = rng.uniform(-6, 6, n)
    e = (5.0 + 20.0 * np.exp(-(x * x + y * y) / 8.0)) / config.HBARC_GEVFM
    z = np.zeros(n)
    d = dict(x=x, y=y, e=e, ux=z, uy=z.copy(), rhoB=z.copy(), rhoS=z.copy(), rhoQ=z.copy())
    cfg = config.base_params(t0=0.6, dt=0.01, particlization=1, eta_mode=config.ETA_CONSTANT,
                             constant_eta_over_s=0.08)
    p = Pair(dev, orc, cfg, d)
    p.initialize()
    ints_equal(p.g, p.o, "BTRACK")
    ints_equal(p.g, p.o, "FREEZE")
    assert np.sum(p.o.ifield("ORC_I_BTRACK") == 0) + np.sum(p.o.ifield("ORC_I_BTRACK") == -1) > 0
    p.step(3)
    ints_equal(p.g, p.o, "FREEZE")
    ints_equal(p.g, p.o, "BTRACK")
    assert np.any(p.o.ifield("ORC_I_FREEZE") == 4)
    compare(p.g, p.o, STATE + DERIVS, what="sparse 3 steps")
    p.close()


def test_reductions_match_oracle(dev, orc):
    """Conservation sums and the frozen count (SystemState::conservation_*)
    on the viscous lumpy event after 3 steps."""
    d = lumpy_small(0.1)
    p = Pair(dev, orc, config.lumpy_params(), d)
    p.initialize()
    p.step(3)
    for name in ("ENTROPY", "BARYON", "STRANGE", "CHARGE", "ENERGY", "EZ", "FROZEN"):
        a = p.g.reduce("CCAKE_R_" + name)
        b = p.o.reduce(ccake_amd.REDUCTIONS["CCAKE_R_" + name])
        assert abs(a - b) <= 1e-12 * max(abs(b), 1e-300) or a == b, (name, a, b)
    p.close()


def test_freeze_out_run_to_completion_on_gpu(dev, orc):
    """The lumpy event on the table EoS (0.1 fm lattice, particles below the
    reference's default e_cutoff of 0.15 GeV/fm^3 dropped: 8,077 particles)
    evolved until every particle has frozen out
    (SPHWorkstation::continue_evolution): the run ends before max_tau with
    every particle at Freeze >= 4, two GPU runs agree to the bit, and the
    oracle needs the same number of steps (within 3) with the same frozen
    counts along the way (within 0.5 %)."""
    tab, dims, mins, maxs = eos_tables.tanh_conformal_table()
    eos = ccake_amd.Eos(dev, 1, T_FO, dims, mins, maxs, tab)
    cfg = dict(config.lumpy_params(), e_freeze=eos.efreeze(T_FO))
    d = ic.lumpy(spacing=0.1, e_cutoff_gev=0.15)
    n = len(d["x"])
    outs = []
    for _ in range(2):
        ev = ccake_amd.Event(dev, eos, cfg, n)
        ccake_amd.load_event(ev, d)
        ev.initialize()
        steps, frozen = 0, []
        while ev.continue_evolution(max_tau=30.0):
            ev.advance_timestep(cfg["dt"], 2)
            steps += 1
            if steps % 50 == 0:
                frozen.append(ev.reduce("CCAKE_R_FROZEN"))
            assert steps < 3000
        fz = ev.download_int("CCAKE_I_FREEZE")
        outs.append((steps, ev.t, fz, ev.download("CCAKE_F_T"), ev.download("CCAKE_F_UX"), frozen))
        ev.close()
    (s1, t1, f1, T1, u1, fr1), (s2, t2, f2, T2, u2, fr2) = outs
    print("steps", s1, "t", t1, "frozen every 50 steps", fr1)
    assert t1 < 30.0 and np.all(f1 >= 4)
    assert s1 == s2 and t1 == t2 and np.array_equal(f1, f2) and fr1 == fr2
    assert np.array_equal(T1, T2) and np.array_equal(u1, u2)
    eos.close()

    oeos = orc.Eos(1, T_FO, dims, mins, maxs, tab)
    o = orc.Event(cfg, oeos, n)
    oracle_load(o, d)
    assert o.initialize() == 0
    steps, frozen = 0, []
    while not np.all(o.ifield("ORC_I_FREEZE") >= 4) and o.t <= 30.0:
        assert o.advance_timestep(cfg["dt"], 2) == 0
        steps += 1
        if steps % 50 == 0:
            frozen.append(o.reduce(ccake_amd.REDUCTIONS["CCAKE_R_FROZEN"]))
    print("oracle steps", steps, "frozen every 50 steps", frozen)
    assert abs(steps - s1) <= 3
    for a, b in zip(fr1, frozen):
        assert abs(a - b) <= 0.005 * n


def test_gubser_full_size_lattice_on_gpu(dev):
    """Ideal Gubser flow on the 0.02 fm lattice (251,001 particles, the C1
    lattice with the C0 physics): 40 RK2 steps are deterministic (two runs
    agree to the bit), leave the extensive entropy unchanged, and follow the
    analytic solution within the SPH smoothing error."""
    cfg = config.gubser_ideal_params()
    d = ic.gubser_ideal(spacing=0.02)
    n = len(d["x"])
    assert n == 251001
    eos = ccake_amd.Eos(dev, 0)
    runs = []
    for _ in range(2):
        ev = ccake_amd.Event(dev, eos, cfg, n)
        ccake_amd.load_event(ev, d)
        ev.initialize()
        s0 = ev.download("CCAKE_F_EXT_S")
        assert ev.run_steps(cfg["dt"], 40) == 40
        runs.append((ev.download("CCAKE_F_E"), ev.download("CCAKE_F_UX"), ev.download("CCAKE_F_X"),
                     ev.download("CCAKE_F_Y"), ev.download("CCAKE_F_EXT_S"), s0, ev.t))
        ev.close()
    (e1, u1, x1, y1, s1, s01, t1), (e2, u2, _, _, _, _, _) = runs
    assert np.array_equal(e1, e2) and np.array_equal(u1, u2)
    assert np.array_equal(s1, s01)
    r = np.hypot(x1, y1)
    ea = ic.gubser_energy(t1, r)
    m = r < 2.5
    assert np.max(np.abs(e1[m] - ea[m]) / ea[m]) < 0.07
    eos.close()


STATE_FIELDS = ["X", "Y", "UX", "UY", "SIGMA_LAB", "SIGMA", "S", "E", "T", "MUB", "RHOB", "RHOS", "RHOQ", "BULK",
                "EXT_S", "GRADP0", "GRADE0", "GRADV0", "DUDT0", "DEXT_S"]


def run_lumpy_table(dev, table, steps=5):
    tab, dims, mins, maxs = table
    eos = ccake_amd.Eos(dev, 1, T_FO, dims, mins, maxs, tab)
    d = lumpy_small(0.05)
    cfg = config.lumpy_params(particlization=0)
    ev = ccake_amd.Event(dev, eos, cfg, len(d["x"]))
    ccake_amd.load_event(ev, d)
    ev.initialize()
    ev.run_steps(cfg["dt"], steps)
    out = {f: ev.download(ccake_amd.FIELDS["CCAKE_F_" + f]) for f in STATE_FIELDS}
    ns = ev.nslices()
    ev.close()
    eos.close()
    return out, ns


def test_tile_workgroups_and_split_eos_match_the_slice_path(dev, table, monkeypatch):
    """The tile-workgroup pair kernels (a tile's slices added in LDS) against
    single-wave slices with their sums through HBM (CCAKE_LDS_SLICES=0), and
    the EoS stage in two kernels against one (CCAKE_EOS_SPLIT=0), on a
    40,401-particle lumpy table event after initialization and 5 steps: the
    split EoS stage bit for bit (the same operations), the slice paths within
    1e-12 of each other (whole field; only the grouping of the pair sums
    differs)."""
    base, ns = run_lumpy_table(dev, table)
    assert ns >= 4            # the tile workgroups run
    monkeypatch.setenv("CCAKE_EOS_SPLIT", "0")
    one, _ = run_lumpy_table(dev, table)
    monkeypatch.delenv("CCAKE_EOS_SPLIT")
    monkeypatch.setenv("CCAKE_LDS_SLICES", "0")
    hbm, _ = run_lumpy_table(dev, table)
    for f in STATE_FIELDS:
        assert np.array_equal(base[f], one[f]), f
        assert rel(base[f], hbm[f]) <= 1e-12, (f, rel(base[f], hbm[f]))
