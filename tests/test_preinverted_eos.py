"""Pre-inverted EoS, flat (s, nB) format (online_inverter_enabled = false;
EoS_Interpolator, src/eos_interpolator.cpp:291-892).

The oracle's restatement is pinned by what the mathematics fixes: the
bilinear lookup in (log10 s, log10 nB) is exact on data that are bilinear in
those coordinates and on the nodes; the host-side query equals the per-step
lookup inside the table and clamps outside it; on the conformal EoS the
lookups, the e -> s inversion and efreeze_at_T reproduce the closed forms
within the interpolation error, and dw/ds from the susceptibilities is 4T/3;
an ideal Gubser run on the table follows the analytic solution and the
Newton path.  The GPU then agrees with the oracle to 1e-12."""
import math

import numpy as np
import pytest

import ccake_amd
from ccake_amd import config, eos_tables, ic

HC = config.HBARC_MEVFM
C = eos_tables.C_CONF
T_FO = 150.0 / HC


def bilinear_table(rng, Ns=7, NB=5):
    """Fields that are bilinear in (Ls, Lb) = (log10 s, log10 nB)."""
    s = np.logspace(-2, 3, Ns)
    nB = np.logspace(-3, 0, NB)
    Ls, Lb = np.meshgrid(np.log10(s), np.log10(nB), indexing="ij")
    coef = {}
    fields = {}
    for k in ("T", "muB", "muS", "muQ", "e", "p", "cs2"):
        a, b, c, d = rng.uniform(0.5, 2.0, 4)
        coef[k] = (a, b, c, d)
        fields[k] = a + b * Ls + c * Lb + d * Ls * Lb
    fields["chiTT"] = np.full_like(Ls, 3.0)
    fields["chiTB"] = np.full_like(Ls, 0.5)
    fields["chiBB"] = np.full_like(Ls, 2.0)
    return s, nB, coef, eos_tables.preinverted_flat(fields, s, nB)


def bil(coef, s, nB):
    a, b, c, d = coef
    Ls, Lb = math.log10(s), math.log10(nB)
    return a + b * Ls + c * Lb + d * Ls * Lb


def oracle_eos(orc, dims, table):
    return orc.Eos(2, T_FO, dims, [0] * 4, [0] * 4, table)


# ------------------------------------------------------------------ loader
def test_loader_rejects_bad_tables(orc):
    rng = np.random.default_rng(1)
    s, nB, coef, (dims, tab) = bilinear_table(rng)
    oracle_eos(orc, dims, tab)                                 # accepted
    bad = tab.copy()
    bad[0, 3, :] *= 1.01                                       # s axis not uniform in log10
    with pytest.raises(ValueError):
        oracle_eos(orc, dims, bad)
    bad = tab.copy()
    bad[1, :, 0] = -1.0                                        # non-positive nB
    with pytest.raises(ValueError):
        oracle_eos(orc, dims, bad)
    with pytest.raises(ValueError):
        oracle_eos(orc, (1, 5, 1, 1), tab[:, :1, :])


def test_lookup_exact_on_bilinear_data_and_nodes(orc):
    rng = np.random.default_rng(2)
    s, nB, coef, (dims, tab) = bilinear_table(rng)
    eos = oracle_eos(orc, dims, tab)
    names = ["T", "muB", "muS", "muQ", "e", "p", "cs2"]
    for _ in range(200):
        sq = 10 ** rng.uniform(-2, 3)
        bq = 10 ** rng.uniform(-3, 0) * rng.choice([-1, 1])    # the lookup uses |rhoB|
        o = eos.preinv_point(0, sq, bq)
        for j, k in enumerate(names):
            want = bil(coef[k], sq, abs(bq))
            assert abs(o[j] - want) <= 1e-12 * max(1.0, abs(want)), k
        assert o[11] == o[4] + o[5]                             # w = e + p
        q = eos.preinv_point(1, sq, bq)                         # host query: same inside the table
        assert np.allclose(q[:7], o[:7], rtol=1e-13, atol=1e-13) and np.allclose(q[7:9], o[7:9], rtol=1e-13)
    for i in range(len(s)):
        for j in range(len(nB)):
            o = eos.preinv_point(0, s[i], nB[j])
            assert abs(o[0] - tab[2, i, j] / HC) <= 1e-13 * abs(tab[2, i, j] / HC)


def test_query_clamps_and_lookup_extrapolates_outside(orc):
    rng = np.random.default_rng(3)
    s, nB, coef, (dims, tab) = bilinear_table(rng)
    eos = oracle_eos(orc, dims, tab)
    # below the table in s: the host query clamps to s_min, the lookup extrapolates the first cell
    q = eos.preinv_point(1, 1e-3 * s[0], nB[2])
    assert abs(q[0] - bil(coef["T"], s[0], nB[2])) < 1e-12
    o = eos.preinv_point(0, 0.5 * s[0], nB[2])
    assert abs(o[0] - bil(coef["T"], 0.5 * s[0], nB[2])) < 1e-12
    # above: query clamps to s_max; the lookup extrapolates the last cell (bilinear data: exact)
    q = eos.preinv_point(1, 10 * s[-1], nB[2])
    assert abs(q[0] - bil(coef["T"], s[-1], nB[2])) < 1e-12
    o = eos.preinv_point(0, 2 * s[-1], nB[2])
    assert abs(o[0] - bil(coef["T"], 2 * s[-1], nB[2])) < 1e-11
    # s and |rhoB| below 0.001 are read at 0.001 by the lookup
    o1 = eos.preinv_point(0, 0.01, 1e-9)
    o2 = eos.preinv_point(0, 0.01, 1e-3)
    assert np.array_equal(o1, o2)
    assert np.array_equal(eos.preinv_point(0, 1e-6, 0.5), eos.preinv_point(0, 1e-3, 0.5))


def test_e_to_s_exact_when_e_is_linear_in_log_s(orc):
    """invert_e_to_s_host interpolates log10 s linearly in e inside the
    bracketing cell, so on a table with e = a + b log10 s (and any nB
    dependence of the other fields) it returns s = 10^((e - a)/b) exactly;
    outside the table it returns the end nodes."""
    s = np.logspace(-2, 3, 11)                      # coarse: a wrong interpolation shows
    nB = np.logspace(-3, 0, 4)
    Ls, Lb = np.meshgrid(np.log10(s), np.log10(nB), indexing="ij")
    a, b = 2.0, 3.0
    one = np.ones_like(Ls)
    fields = dict(T=one + Lb * 0.1, muB=0 * one, muS=0 * one, muQ=0 * one, e=a + b * Ls, p=one, cs2=one / 3,
                  chiTT=one, chiTB=0 * one, chiBB=one)
    dims, tab = eos_tables.preinverted_flat(fields, s, nB)
    eos = oracle_eos(orc, dims, tab)
    rng = np.random.default_rng(5)
    for e in rng.uniform(a - 2 * b + 1e-6, a + 3 * b - 1e-6, 50):
        got = eos.preinv_point(2, e, 10 ** rng.uniform(-3, 0))
        assert abs(got / 10 ** ((e - a) / b) - 1) < 1e-12
    assert abs(eos.preinv_point(2, a - 10 * b, 0.1) / s[0] - 1) < 1e-15
    assert abs(eos.preinv_point(2, a + 10 * b, 0.1) / s[-1] - 1) < 1e-15


# ------------------------------------------------------------------ conformal closed forms
def test_conformal_table_reproduces_closed_forms(orc):
    dims, tab = eos_tables.preinverted_conformal()
    eos = oracle_eos(orc, dims, tab)
    rng = np.random.default_rng(4)
    for s in 10 ** rng.uniform(-2, 3.5, 100):
        T = (s / (4 * C)) ** (1 / 3)
        o = eos.preinv_point(0, s, 0.0)
        assert abs(o[0] / T - 1) < 1e-4
        assert abs(o[4] / (3 * C * T ** 4) - 1) < 4e-4 and abs(o[6] - 1 / 3) < 1e-15
        assert abs(o[7] / (4 * T / 3) - 1) < 1e-4                  # dw/ds = T + s dT/ds = 4T/3
    nodes = np.logspace(-3, 4, 600)
    for s in nodes[::37]:
        o = eos.preinv_point(0, s, 0.01)
        T = (s / (4 * C)) ** (1 / 3)
        assert abs(o[7] / (4 * T / 3) - 1) < 1e-12                 # at the nodes: the Schur formula exactly
    for e in 10 ** rng.uniform(-1, 3, 100):                        # e -> s: s = 4c (e/3c)^(3/4)
        s = eos.preinv_point(2, e, 0.0)
        assert abs(s / (4 * C * (e / (3 * C)) ** 0.75) - 1) < 2e-4
    efo = eos.efreeze(T_FO)                                        # efreeze_at_T
    assert abs(efo / (3 * C * T_FO ** 4) - 1) < 2e-4


def gubser_oracle(orc, eos, steps, spacing=0.1):
    cfg = dict(config.gubser_ideal_params(), e_freeze=eos.efreeze(T_FO))
    d = ic.gubser_ideal(spacing=spacing)
    ev = orc.Event(cfg, eos, len(d["x"]))
    ev._eos_keep = eos
    for k, f in [("x", "X"), ("y", "Y"), ("e", "IN_E"), ("ux", "UX"), ("uy", "UY"),
                 ("rhoB", "IN_B"), ("rhoS", "IN_SS"), ("rhoQ", "IN_Q")]:
        ev.set("ORC_F_" + f, d[k])
    assert ev.initialize() == 0
    for _ in range(steps):
        assert ev.advance_timestep(cfg["dt"], 2) == 0
    return ev


def test_gubser_on_preinverted_table_follows_newton_and_analytic(orc):
    """20 RK2 steps of ideal Gubser flow (0.1 fm lattice): the pre-inverted
    conformal table gives the same flow as the Newton inversion of the
    conformal EoS to the interpolation error, and the analytic solution to
    the SPH smoothing error."""
    dims, tab = eos_tables.preinverted_conformal()
    pre = gubser_oracle(orc, oracle_eos(orc, dims, tab), 20)
    newton = gubser_oracle(orc, orc.Eos(0, T_FO), 20)
    e1, e0 = pre.field("ORC_F_E"), newton.field("ORC_F_E")
    assert np.max(np.abs(e1 / e0 - 1)) < 1e-3
    assert np.max(np.abs(pre.field("ORC_F_UX") - newton.field("ORC_F_UX"))) < 1e-4
    r = np.hypot(pre.field("ORC_F_X"), pre.field("ORC_F_Y"))
    m = r < 2.5
    assert np.max(np.abs(e1[m] / ic.gubser_energy(pre.t, r[m]) - 1)) < 0.07


# ------------------------------------------------------------------ GPU
@pytest.mark.gpu
def test_preinverted_gubser_parity_on_gpu(dev, orc):
    """Ideal Gubser flow (0.05 fm lattice, 40,401 particles) on the
    pre-inverted conformal table: initialisation, one derivative evaluation
    and 5 steps equal the oracle's to 1e-12; efreeze agrees."""
    dims, tab = eos_tables.preinverted_conformal()
    geos = ccake_amd.Eos(dev, 2, T_FO, dims, [0] * 4, [0] * 4, tab)
    oeos = oracle_eos(orc, dims, tab)
    assert abs(geos.efreeze(T_FO) - oeos.efreeze(T_FO)) <= 1e-14 * oeos.efreeze(T_FO)
    cfg = dict(config.gubser_ideal_params(), e_freeze=oeos.efreeze(T_FO))
    d = ic.gubser_ideal(spacing=0.05)
    n = len(d["x"])
    g = ccake_amd.Event(dev, geos, cfg, n)
    ccake_amd.load_event(g, d)
    o = orc.Event(cfg, oeos, n)
    for k, f in [("x", "X"), ("y", "Y"), ("e", "IN_E"), ("ux", "UX"), ("uy", "UY"),
                 ("rhoB", "IN_B"), ("rhoS", "IN_SS"), ("rhoQ", "IN_Q")]:
        o.set("ORC_F_" + f, d[k])
    fields = ["T", "E", "S", "P", "CS2", "W", "DWDS", "DWDB", "IN_S", "SIGMA", "EXT_S", "X", "Y", "UX", "UY"]

    def compare(what):
        worst = 0.0
        for f in fields:
            a, b = g.download("CCAKE_F_" + f), o.field("ORC_F_" + f)
            err = np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)
            worst = max(worst, err)
            assert err <= 1e-12, (what, f, err)
        print(what, "worst", worst, flush=True)

    g.initialize()
    assert o.initialize() == 0
    compare("init")
    g.get_time_derivatives()
    assert o.get_time_derivatives() == 0
    compare("derivatives")
    for _ in range(5):
        g.advance_timestep(cfg["dt"], 2)
        assert o.advance_timestep(cfg["dt"], 2) == 0
    compare("5 steps")
    g.close()
    geos.close()


@pytest.mark.gpu
def test_preinverted_gubser_physics_on_gpu(dev):
    """120 steps (tau 1 -> 1.6) on the pre-inverted conformal table: the
    analytic solution within the SPH error and entropy unchanged."""
    dims, tab = eos_tables.preinverted_conformal()
    eos = ccake_amd.Eos(dev, 2, T_FO, dims, [0] * 4, [0] * 4, tab)
    cfg = dict(config.gubser_ideal_params(), e_freeze=eos.efreeze(T_FO))
    d = ic.gubser_ideal(spacing=0.05)
    ev = ccake_amd.Event(dev, eos, cfg, len(d["x"]))
    ccake_amd.load_event(ev, d)
    ev.initialize()
    s0 = ev.download("CCAKE_F_EXT_S")
    assert ev.run_steps(cfg["dt"], 120) == 120
    assert np.array_equal(ev.download("CCAKE_F_EXT_S"), s0)
    r = np.hypot(ev.download("CCAKE_F_X"), ev.download("CCAKE_F_Y"))
    m = r < 2.5
    e = ev.download("CCAKE_F_E")
    assert np.max(np.abs(e[m] / ic.gubser_energy(ev.t, r[m]) - 1)) < 0.07
    ev.close()
    eos.close()
