"""CPU tests that pin the oracle (no GPU).

The oracle is the parity checker for the HIP path, so it is checked here
against the reference's own known-answer tests (tests/golden/reference_kats.json,
from unit_tests/*.cpp), brute force, closed forms, finite differences and
exact invariants -- never against a restatement of its own formulas.
"""
import json
import math
import os

import numpy as np
import pytest

from ccake_amd import config, ic

GOLD = os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")
KATS = json.load(open(GOLD))
HBARC = config.HBARC_MEVFM
C_CONF = math.pi ** 2 * (2 * (3 * 3 - 1) + 3.5 * 3 * 2.5) / 90.0


def test_kernel_kats(orc):
    L = orc.lib()
    for D, r, h, want in KATS["kernel"]:
        assert abs(L.orc_kernel(D, r, h) - want) < KATS["tolerance"]


def test_grad_kernel_kats(orc):
    import ctypes as C
    L = orc.lib()
    for pa, pb, h, want in KATS["grad_kernel_2d"]:
        rel = (C.c_double * 2)(pa[0] - pb[0], pa[1] - pb[1])
        r = math.hypot(pa[0] - pb[0], pa[1] - pb[1])
        g = (C.c_double * 2)()
        L.orc_grad_kernel(2, rel, r, h, g)
        assert abs(g[0] - want[0]) < KATS["tolerance"] and abs(g[1] - want[1]) < KATS["tolerance"]


def _event(orc, cfg, n, eos=None):
    eos = eos or orc.Eos(0)
    ev = orc.Event(cfg, eos, n)
    ev._eos_keep = eos
    return ev


def test_smoothing_kat(orc):
    k = KATS["smoothing"]
    ev = _event(orc, config.base_params(hT=k["hT"]), 4)
    ev.set("ORC_F_X", k["x"]); ev.set("ORC_F_Y", k["y"])
    for X, f in (("s", "S"), ("rhoB", "B"), ("rhoQ", "Q"), ("rhoS", "SS")):
        ev.set("ORC_F_MASS_" + f, k["mass"][X]); ev.set("ORC_F_EXT_" + f, k["extensive"][X])
    ev.t = 1.0
    ev.build_neighbors(False)
    ev.smooth_fields()
    e = k["expected_p0"]
    assert abs(ev.field("ORC_F_SM_S")[0] - e["s"]) < 1e-14
    assert abs(ev.field("ORC_F_SM_B")[0] - e["rhoB"]) < 1e-14
    assert abs(ev.field("ORC_F_SM_Q")[0] - e["rhoQ"]) < 1e-14
    assert abs(ev.field("ORC_F_SM_SS")[0] - e["rhoS"]) < 1e-13
    assert abs(ev.field("ORC_F_SIGMA_LAB")[0] - e["sigma_lab"]) < 1e-14


def test_grad_smoothing_kat(orc):
    k = KATS["grad_smoothing"]
    ev = _event(orc, config.base_params(hT=k["hT"], using_shear=0), 5)
    ev.set("ORC_F_X", k["x"]); ev.set("ORC_F_Y", k["y"])
    ev.set("ORC_F_SIGMA_LAB", 1.0); ev.set("ORC_F_P", k["p"])
    ev.build_neighbors(False)
    ev.smooth_gradients()
    assert abs(ev.field("ORC_F_GRADP0")[0] - k["expected_gradP_p0"][0]) < 1e-14
    assert abs(ev.field(orc.FIELDS["ORC_F_GRADP0"] + 1)[0] - k["expected_gradP_p0"][1]) < 1e-14


def _pairs_set(pi, pj):
    return np.unique(pi.astype(np.int64) * (1 << 32) + pj.astype(np.int64))


@pytest.mark.parametrize("spacing", [0.1, 0.05])
def test_pairs_cell_list_equals_brute_force_on_lattice(orc, spacing):
    d = ic.gubser_ideal(spacing=spacing)
    n = len(d["x"])
    ev = _event(orc, config.gubser_ideal_params(), n)
    ev.set("ORC_F_X", d["x"]); ev.set("ORC_F_Y", d["y"])
    ev.build_neighbors(True)
    a = _pairs_set(*ev.pairs())
    bt = ev.ifield("ORC_I_BTRACK").copy()
    ev.build_neighbors(False)
    b = _pairs_set(*ev.pairs())
    assert np.array_equal(a, b)
    assert np.array_equal(bt, ev.ifield("ORC_I_BTRACK"))
    # interior count (Gauss circle): lattice points with i^2 + j^2 <= (rc/a)^2,
    # self excluded.  With decimal-exact coordinates the four axis points at
    # exactly rc are inside (<=): 440 at 0.05 fm (436 would be the strict-< count).
    R = int(round(0.6 / spacing))
    want = sum(1 for i in range(-R, R + 1) for j in range(-R, R + 1) if i * i + j * j <= R * R) - 1
    centre = np.argmin(np.asarray(d["x"]) ** 2 + np.asarray(d["y"]) ** 2)
    assert bt[centre] == want
    if spacing == 0.05:
        assert want == 440
    # symmetry j in N(i) <=> i in N(j)
    pi, pj = ev.pairs()
    assert np.array_equal(_pairs_set(pi, pj), _pairs_set(pj, pi))


def test_pairs_boundary_and_random(orc):
    rng = np.random.default_rng(3)
    n = 3000
    x = rng.uniform(-3, 3, n); y = rng.uniform(-3, 3, n)
    # plant exact-boundary pairs at distance 2 hT = 0.6 along the axes
    x[:10] = np.arange(10) * 0.6; y[:10] = 5.0
    ev = _event(orc, config.base_params(hT=0.3), n)
    ev.set("ORC_F_X", x); ev.set("ORC_F_Y", y)
    ev.build_neighbors(True)
    a = _pairs_set(*ev.pairs())
    ev.build_neighbors(False)
    b = _pairs_set(*ev.pairs())
    assert np.array_equal(a, b)
    # numpy brute force with the same rounding (no FMA): dx*dx + dy*dy <= rc*rc
    rc2 = 0.6 * 0.6
    dx = x[:, None] - x[None, :]; dy = y[:, None] - y[None, :]
    m = (dx * dx + dy * dy) <= rc2
    np.fill_diagonal(m, False)
    ii, jj = np.nonzero(m)
    assert np.array_equal(a, _pairs_set(ii.astype(np.int32), jj.astype(np.int32)))


# ---------------------------------------------------------------- EoS
def test_conformal_inversion_closed_form(orc):
    eos = orc.Eos(0)
    for s in [1e-3, 0.1, 5.0, 80.0, 3000.0]:
        ok, x, m, th = eos.invert_point(0, [s, 0, 0, 0], [580 / HBARC, 0, 0, 0], 0)
        assert ok and m == 0
        T = (s / (4 * C_CONF)) ** (1 / 3)
        assert abs(x[0] - T) / T < 1e-13 and np.all(x[1:] == 0)
        assert abs(th[12] - 4 * T / 3) / T < 1e-12        # dw/ds = 4T/3
        assert th[10] == 1 / 3
    for e in [0.01, 3.0, 400.0]:
        ok, x, m, th = eos.invert_point(1, [e, 0, 0, 0], [580 / HBARC, 0, 0, 0], 0)
        T = (e / (3 * C_CONF)) ** 0.25
        assert ok and abs(x[0] - T) / T < 1e-13
        assert abs(th[5] - 4 * C_CONF * T ** 3) / th[5] < 1e-12


def _thermo_at(eos, member, T, mB, mQ, mS):
    return eos.full_thermo(member, [T, mB, mQ, mS])


def test_conformal_roundtrip_with_charges(orc):
    eos = orc.Eos(0)
    rng = np.random.default_rng(7)
    for _ in range(20):
        T = rng.uniform(0.3, 3.0); mu = rng.uniform(-1.0, 1.0, 3)
        r = _thermo_at(eos, 0, T, *mu)
        tgt = [r[1], r[2], r[3], r[4]]                        # s, B, S, Q
        ok, x, m, th = eos.invert_point(0, tgt, [T * 1.3, 0, 0, 0], 0)
        assert ok and m == 0
        assert np.allclose(x, [T, *mu], rtol=1e-10, atol=1e-10)


def test_enthalpy_derivatives_by_finite_differences(orc):
    """dw/ds at fixed (B,S,Q) and dw/dB at fixed (s,S,Q): pins the Schur
    complements of src/eos_derivatives.cpp (index placement, transposes)."""
    eos = orc.Eos(0)
    rng = np.random.default_rng(11)
    for _ in range(6):
        T = rng.uniform(0.6, 2.0); mu = rng.uniform(-0.8, 0.8, 3)
        r = _thermo_at(eos, 0, T, *mu)
        dens = np.array([r[1], r[2], r[3], r[4]])             # s, B, S, Q
        ok, x, m, th = eos.invert_point(0, list(dens), [T, *mu], 0)
        assert ok

        def w_at(dv):
            ok2, x2, _, th2 = eos.invert_point(0, list(dens + dv), [T, *mu], 0)
            assert ok2
            return th2[11]
        for k, idx in [(0, 12), (1, 13), (2, 14), (3, 15)]:    # s, B, S, Q -> dwds, dwdB, dwdS, dwdQ
            h = 1e-5 * max(abs(dens[k]), 1.0)
            dv = np.zeros(4); dv[k] = h
            fd = (w_at(dv) - w_at(-dv)) / (2 * h)
            assert abs(fd - th[idx]) <= 2e-6 * max(1.0, abs(th[idx])), (k, fd, th[idx])


def test_alpha_derivatives_by_finite_differences(orc):
    """d alpha_c / d(s, B, S, Q), alpha_c = mu_c / T for c = B, S, Q, each at
    fixed values of the other three densities (the fields the diffusion MRFs
    read, src/eos_derivatives.cpp): central differences along the inversion
    on the conformal EoS with all three charges non-zero.  Pins the index
    placement of the 12 fields (charge major, density minor) and their
    Schur complements."""
    eos = orc.Eos(0)
    rng = np.random.default_rng(13)
    for _ in range(6):
        T = rng.uniform(0.6, 2.0); mu = rng.uniform(-0.8, 0.8, 3)
        r = _thermo_at(eos, 0, T, *mu)
        dens = np.array([r[1], r[2], r[3], r[4]])             # s, B, S, Q
        ok, x, m, th = eos.invert_point(0, list(dens), [T, *mu], 0)
        assert ok

        def alpha_at(dv):
            ok2, x2, _, th2 = eos.invert_point(0, list(dens + dv), [T, *mu], 0)
            assert ok2
            return np.array([th2[1], th2[2], th2[3]]) / th2[0]    # mu_B, mu_S, mu_Q over T
        for k in range(4):
            h = 1e-5 * max(abs(dens[k]), 1.0)
            dv = np.zeros(4); dv[k] = h
            fd = (alpha_at(dv) - alpha_at(-dv)) / (2 * h)
            for c in range(3):
                got = th[16 + 4 * c + k]
                assert abs(fd[c] - got) <= 2e-6 * max(1.0, abs(got)), (c, k, fd[c], got)


def test_tanh_conformal_cs2(orc):
    """cs2 of the tanh-conformal EoS: at mu = 0 it is s/(T chi_TT); at mu != 0
    the adiabatic derivative dp/de along fixed n/s, by finite differences."""
    # a table chain carries tanh_conformal as member 1; use a tiny table
    tab, dims, mins, maxs = tiny_table()
    eos = orc.Eos(1, 150 / HBARC, dims, mins, maxs, tab)
    for T in [0.5, 1.0, 1.5]:
        r = _thermo_at(eos, 1, T, 0, 0, 0)
        assert abs(r[6] - r[1] / (T * r[16])) < 1e-12
    T, mu = 1.2, np.array([0.3, -0.2, 0.1])
    r = _thermo_at(eos, 1, T, *mu)
    dens = np.array([r[1], r[2], r[3], r[4]])

    def pe(scale):
        ok, x, m, th = eos.invert_point(0, list(dens * scale), [T, *mu], 1)
        assert ok and m == 1
        return th[9], th[4]
    p1, e1 = pe(1 + 1e-6); p0, e0 = pe(1 - 1e-6)
    assert abs((p1 - p0) / (e1 - e0) - r[6]) < 1e-6


def tiny_table():
    """Small tanh-conformal table in the 17-field layout (for EoS tests)."""
    from ccake_amd import eos_tables
    return eos_tables.tanh_conformal_table(dims=(8, 5, 5, 5), T_range=(0.05, 0.8), mu_max=0.45)


def test_table_interpolation_exact_at_nodes_and_multilinear(orc):
    tab, dims, mins, maxs = tiny_table()
    eos = orc.Eos(1, 150 / HBARC, dims, mins, maxs, tab)
    t = tab.reshape(*dims, 17)
    axes = [np.linspace(mins[a], maxs[a], dims[a]) for a in range(4)]
    rng = np.random.default_rng(5)
    for _ in range(10):
        idx = [rng.integers(0, d) for d in dims]
        r = _thermo_at(eos, 0, *[axes[a][idx[a]] for a in range(4)])
        assert np.allclose(r, t[tuple(idx)], rtol=1e-12, atol=1e-14)
    # midpoint along T is the average of the two nodes (multilinear)
    i = [3, 2, 2, 2]
    pt = [0.5 * (axes[0][3] + axes[0][4]), axes[1][2], axes[2][2], axes[3][2]]
    r = _thermo_at(eos, 0, *pt)
    assert np.allclose(r, 0.5 * (t[3, 2, 2, 2] + t[4, 2, 2, 2]), rtol=1e-12)


def test_table_chain_inversion_roundtrip(orc):
    """Round trip in the thermodynamically stable part of the table (the
    tanh-modulated pressure has cs2 < 0 at low T and large mu, where the
    chain moves on to the fallback EoS and the inverse is not unique)."""
    tab, dims, mins, maxs = tiny_table()
    eos = orc.Eos(1, 150 / HBARC, dims, mins, maxs, tab)
    rng = np.random.default_rng(9)
    for _ in range(10):
        T = rng.uniform(2.0, 3.8); mu = rng.uniform(-0.8, 0.8, 3)
        r = _thermo_at(eos, 0, T, *mu)
        ok, x, m, th = eos.invert_point(0, [r[1], r[2], r[3], r[4]], [T * 0.9, mu[0] * 0.5, 0, 0], 0)
        assert ok and m == 0
        r2 = _thermo_at(eos, 0, *x)
        assert np.allclose([r2[1], r2[2], r2[3], r2[4]], [r[1], r[2], r[3], r[4]], rtol=1e-9, atol=1e-9)
        assert np.allclose(x, [T, *mu], rtol=1e-8, atol=1e-8)


# ---------------------------------------------------------------- transport
def test_transport_closed_forms(orc):
    cfg = config.lumpy_params()
    ev = _event(orc, cfg, 3)
    T = np.array([0.6, 0.17 / config.HBARC_GEVFM, 1.5]); w = np.array([2.0, 3.0, 40.0])
    ev.set("ORC_F_T", T); ev.set("ORC_F_W", w); ev.set("ORC_F_P", [0.5, 0.7, 10.0]); ev.set("ORC_F_CS2", 0.25)
    ev.viscosities()
    eta = 0.08 * w / T
    assert np.allclose(ev.field("ORC_F_ETA_PI"), eta, rtol=1e-15)
    assert np.allclose(ev.field("ORC_F_TAU_PI"), np.maximum(5 * eta / w, 0.005), rtol=1e-15)
    zeta = 0.1 * np.exp(-(T - cfg["zeta_gauss_T0"]) ** 2 / (2 * cfg["zeta_gauss_width"] ** 2)) * w / T
    assert np.allclose(ev.field("ORC_F_ZETA_PI"), zeta, rtol=1e-14)
    assert abs(ev.field("ORC_F_ZETA_PI")[1] - 0.1 * w[1] / T[1]) < 1e-14   # peak of the Gaussian
    tPi = np.maximum(5 * zeta / ((1 / 3 - 0.25) ** 2 * w), 0.1)
    assert np.allclose(ev.field("ORC_F_TAU_BULK"), tPi, rtol=1e-14)
    assert np.allclose(ev.field("ORC_F_DELTA_PIPI"), ev.field("ORC_F_TAU_PI"))       # israel-stewart
    assert np.allclose(ev.field("ORC_F_DELTA_BB"), ev.field("ORC_F_TAU_BULK"))


def test_hydro_numbers_closed_forms_at_rest(orc):
    """EoM_default::compute_hydro_numbers (src/eom_default.cpp:1831-1995) on
    fluid cells at rest (u = 0, gamma = 1) with prescribed velocity gradients,
    pi^{ij} and Pi, at tau = 1.3: the MRF shear rate reduces to the textbook
    sigma^{ij} = -(d_i v^j + d_j v^i)/2 + delta_ij (1/tau + div v)/3, of which
    the reference's read-back keeps sigma^{yy} (stored at (2,2), read into
    (1,1); its time components then follow from u = 0, and tau^2 sigma^{etaeta}
    from tracelessness), so Kn_shear = tau_pi sqrt(2) |sigma^{yy}|;
    Kn_bulk = tau_Pi |theta| with theta = 1/tau + div v;
    Re^-1_shear = sqrt(pi:pi)/p with pi:pi = pi_xx^2 + 2 pi_xy^2 + pi_yy^2 + tau^4 (pi^{etaeta})^2;
    Re^-1_bulk = |Pi|/p."""
    rng = np.random.default_rng(11)
    n = 5
    cfg = config.base_params(t0=1.3, constant_eta_over_s=0.08)
    ev = _event(orc, cfg, n)
    ev.t = 1.3
    t = 1.3
    ev.set("ORC_F_GAMMA", np.ones(n))
    for f in ("UX", "UY", "VX", "VY"):
        ev.set("ORC_F_" + f, np.zeros(n))
    for f in ("T", "W", "S", "P", "SIGMA", "SIGMA_LAB", "MASS_S", "DWDS", "TAU_PI", "TAU_BULK", "ETA_PI",
              "ZETA_PI", "DELTA_PIPI", "DELTA_BB"):
        ev.set("ORC_F_" + f, rng.uniform(0.5, 2.0, n))
    g = rng.normal(size=(4, n))                     # d_x v^x, d_x v^y, d_y v^x, d_y v^y
    fid = orc.FIELDS
    for k, c in enumerate((0, 1, 3, 4)):
        ev.field(fid["ORC_F_GRADV0"] + c)[:] = g[k]
    pxx, pxy, pyy, pee = rng.normal(scale=0.1, size=(4, n))
    shv = {5: pxx, 6: pxy, 9: pxy, 10: pyy, 15: pee}
    for c in range(16):
        ev.field(fid["ORC_F_SHV0"] + c)[:] = shv.get(c, np.zeros(n))
    Pi = rng.normal(scale=0.1, size=n)
    ev.set("ORC_F_BULK", Pi)
    ev.eom()
    divv = g[0] + g[3]
    theta = 1.0 / t + divv
    assert np.allclose(ev.field("ORC_F_THETA"), theta, rtol=1e-13)
    s_yy = -g[3] + theta / 3.0
    p = ev.field("ORC_F_P")
    assert np.allclose(ev.field("ORC_F_SHEAR_KNUDSEN"), ev.field("ORC_F_TAU_PI") * np.sqrt(2.0) * np.abs(s_yy),
                       rtol=1e-12)
    assert np.allclose(ev.field("ORC_F_BULK_KNUDSEN"), ev.field("ORC_F_TAU_BULK") * np.abs(theta), rtol=1e-13)
    pipi = pxx ** 2 + 2 * pxy ** 2 + pyy ** 2 + t ** 4 * pee ** 2
    assert np.allclose(ev.field("ORC_F_INV_REYNOLDS_SHEAR"), np.sqrt(pipi) / p, rtol=1e-13)
    assert np.allclose(ev.field("ORC_F_INV_REYNOLDS_BULK"), np.abs(Pi) / p, rtol=1e-15)


# ---------------------------------------------------------------- evolution
def _gubser_event(orc, spacing, **kw):
    cfg = config.gubser_ideal_params(**kw)
    d = ic.gubser_ideal(spacing=spacing)
    n = len(d["x"])
    ev = _event(orc, cfg, n)
    for k, f in [("x", "X"), ("y", "Y"), ("e", "IN_E"), ("ux", "UX"), ("uy", "UY"),
                 ("rhoB", "IN_B"), ("rhoS", "IN_SS"), ("rhoQ", "IN_Q")]:
        ev.set("ORC_F_" + f, d[k])
    assert ev.initialize() == 0
    return ev, cfg, d


def test_ideal_invariants(orc):
    """Ideal conformal flow: extensive entropy and charges are constant
    (F_s = M_s = 0 and dN/dt = R_N^{-1} 0); Pi stays 0 (tau_Pi = 1e10)."""
    ev, cfg, d = _gubser_event(orc, 0.2)
    s0 = ev.field("ORC_F_EXT_S").copy()
    for _ in range(3):
        assert ev.advance_timestep(cfg["dt"], 2) == 0
    assert np.array_equal(ev.field("ORC_F_EXT_S"), s0)
    assert np.all(ev.field("ORC_F_EXT_B") == 0) and np.all(ev.field("ORC_F_BULK") == 0)
    assert np.all(ev.field("ORC_F_DEXT_S") == 0)


def test_radial_symmetry(orc):
    ev, cfg, d = _gubser_event(orc, 0.2)
    for _ in range(2):
        ev.advance_timestep(cfg["dt"], 2)
    x0, y0 = np.round(d["x"], 6), np.round(d["y"], 6)
    key = {(a, b): i for i, (a, b) in enumerate(zip(x0, y0))}
    mirror = np.array([key[(b, a)] for a, b in zip(x0, y0)])   # x <-> y
    e = ev.field("ORC_F_E")
    assert np.max(np.abs(e - e[mirror])) / e.max() < 1e-12
    flip = np.array([key[(-a if a != 0 else 0.0, b)] for a, b in zip(x0, y0)])
    assert np.max(np.abs(e - e[flip])) / e.max() < 1e-12


def test_rk2_update_arithmetic(orc):
    """Evolver::step_rk with prescribed derivatives: y = y0 + dt f."""
    ev = _event(orc, config.base_params(), 4)
    rng = np.random.default_rng(1)
    f = {k: rng.normal(size=4) for k in ("DEXT_S", "DEXT_BULK", "VX", "VY", "DUDT0")}
    y0 = {k: rng.uniform(1, 2, 4) for k in ("EXT_S", "EXT_BULK", "X", "Y", "UX")}
    for k, v in y0.items():
        ev.set("ORC_F_" + k, v)
    for k, v in f.items():
        ev.set("ORC_F_" + k, v)
    ev.cache_state()
    ev.rk_update(0.25, 1.0)
    assert np.allclose(ev.field("ORC_F_EXT_S"), y0["EXT_S"] + 0.25 * f["DEXT_S"], rtol=0, atol=1e-15)
    assert np.allclose(ev.field("ORC_F_X"), y0["X"] + 0.25 * f["VX"], rtol=0, atol=1e-15)
    assert np.allclose(ev.field("ORC_F_UX"), y0["UX"] + 0.25 * f["DUDT0"], rtol=0, atol=1e-15)
    assert abs(ev.t - 1.25) < 1e-15


def test_gubser_short_evolution(orc):
    """Physics check on a coarse lattice: the evolved energy density follows
    the analytic ideal Gubser solution within the SPH smoothing error
    already present at tau0 (radii where that error is small)."""
    ev, cfg, d = _gubser_event(orc, 0.1)
    for _ in range(20):
        assert ev.advance_timestep(cfg["dt"], 2) == 0
    r = np.hypot(ev.field("ORC_F_X"), ev.field("ORC_F_Y"))
    ea = ic.gubser_energy(ev.t, r)
    m = r < 2.5
    rel = np.abs(ev.field("ORC_F_E")[m] - ea[m]) / ea[m]
    assert rel.max() < 0.07


def test_freeze_status_machine(orc):
    """check_freeze_out_status transitions (include/freeze_out.h:455-542)."""
    cfg = config.base_params(e_freeze=1.0, particlization=1)
    ev = _event(orc, cfg, 7)
    ev.set("ORC_F_E", [2.0, 0.5, 0.5, 2.0, 0.5, 2.0, 0.5])
    fz = ev.ifield("ORC_I_FREEZE"); bt = ev.ifield("ORC_I_BTRACK")
    fz[:] = [0, 0, 1, 1, 1, 1, 4]
    bt[:] = [5, 5, -1, 5, 5, 5, 5]
    frz1 = ev.field("ORC_F_FRZ1_E"); frz1[:] = [0, 0, 0, 1.0, 0.9, 3.0, 0]
    ev.freeze_out_particles()            # first step: status check + snapshot only
    # 0 stays 0 above e_FO; 0 -> 1 below; 1 -> 3 without neighbours; 1 stays 1
    # when e rose; 1 -> 3 below e_FO; 1 -> 0 above e_FO and falling; 4 stays 4
    assert list(ev.ifield("ORC_I_FREEZE")) == [0, 1, 3, 1, 3, 0, 4]
    assert np.array_equal(ev.field("ORC_F_FRZ2_E"), ev.field("ORC_F_E"))


def test_hydro_numbers_closed_forms_with_flow(orc):
    """compute_hydro_numbers (src/eom_default.cpp:1831-1995) on moving fluid
    cells, u = (u^x, 0), gamma = sqrt(1 + u_x^2), v = u / gamma, with
    prescribed d_i v^j, pi^{ij} and Pi at tau = 1.3.  The expansion rate is
    theta = d_mu u^mu + u^tau / tau = dgamma/dt + gamma div v + gamma / tau
    (the comoving dgamma/dt = u^x (du^x/dt) / gamma from the EoM's du/dt).
    With u^y = 0 the projector's y row is the unit row, so the textbook
    sigma^{yy} = d_y u^y - theta / 3 with d_y u^y = gamma d_y v^y.  The
    reference's read-back keeps sigma^{yy} (into the (1,1) slot) and
    sigma^{y eta} = 0 and rebuilds the rest from u: then
    sigma_{mu nu} sigma^{mu nu} = 2 (sigma^{yy})^2 / gamma^4, so
    Kn_shear = tau_pi sqrt(2) |sigma^{yy}| / gamma^2, and Kn_bulk =
    tau_Pi |theta|.  The flow-dependent parts pinned here are the MRF
    shear rate's u-terms and theta's dgamma/dt and gamma factors."""
    rng = np.random.default_rng(12)
    n = 6
    cfg = config.base_params(t0=1.3, constant_eta_over_s=0.08)
    ev = _event(orc, cfg, n)
    ev.t = 1.3
    t = 1.3
    ux = rng.uniform(-0.8, 0.8, n)
    gam = np.sqrt(1.0 + ux ** 2)
    ev.set("ORC_F_GAMMA", gam)
    ev.set("ORC_F_UX", ux); ev.set("ORC_F_UY", np.zeros(n))
    ev.set("ORC_F_VX", ux / gam); ev.set("ORC_F_VY", np.zeros(n))
    for f in ("T", "W", "S", "P", "SIGMA", "SIGMA_LAB", "MASS_S", "DWDS", "TAU_PI", "TAU_BULK", "ETA_PI",
              "ZETA_PI", "DELTA_PIPI", "DELTA_BB"):
        ev.set("ORC_F_" + f, rng.uniform(0.5, 2.0, n))
    g = rng.normal(scale=0.3, size=(4, n))          # d_x v^x, d_x v^y, d_y v^x, d_y v^y
    fid = orc.FIELDS
    for k, c in enumerate((0, 1, 3, 4)):
        ev.field(fid["ORC_F_GRADV0"] + c)[:] = g[k]
    # a shear tensor orthogonal to u (pi^{0i} = v_j pi^{ji}) and traceless
    pxx, pxy, pyy = rng.normal(scale=0.05, size=(3, n))
    shv = np.zeros((16, n))
    shv[5] = pxx; shv[6] = shv[9] = pxy; shv[10] = pyy
    shv[1] = shv[4] = (ux / gam) * pxx
    shv[2] = shv[8] = (ux / gam) * pxy
    shv[0] = (ux / gam) ** 2 * pxx
    shv[15] = (shv[0] - pxx - pyy) / t ** 2
    for c in range(16):
        ev.field(fid["ORC_F_SHV0"] + c)[:] = shv[c]
    ev.set("ORC_F_BULK", rng.normal(scale=0.1, size=n))
    ev.eom()
    dux = ev.field(fid["ORC_F_DUDT0"])
    assert np.all(np.abs(dux) > 1e-3)                # the EoM's du/dt is live
    theta = ux * dux / gam + gam * (g[0] + g[3]) + gam / t
    assert np.allclose(ev.field("ORC_F_THETA"), theta, rtol=1e-12)
    s_yy = gam * g[3] - theta / 3.0
    kn = ev.field("ORC_F_SHEAR_KNUDSEN")
    assert np.allclose(kn, ev.field("ORC_F_TAU_PI") * np.sqrt(2.0) * np.abs(s_yy) / gam ** 2, rtol=1e-10)
    assert np.allclose(ev.field("ORC_F_BULK_KNUDSEN"), ev.field("ORC_F_TAU_BULK") * np.abs(theta), rtol=1e-12)
    # plausible mistakes fall far outside: gamma factors dropped, the
    # Eulerian/comoving dgamma/dt term dropped, the at-rest formula
    for wrong in (ev.field("ORC_F_TAU_PI") * np.sqrt(2.0) * np.abs(g[3] - theta / 3.0) / gam ** 2,
                  ev.field("ORC_F_TAU_PI") * np.sqrt(2.0) * np.abs(s_yy),
                  ev.field("ORC_F_TAU_PI") * np.sqrt(2.0) * np.abs(gam * g[3] - (theta - ux * dux / gam) / 3.0) / gam ** 2):
        assert np.max(np.abs(wrong - kn) / kn) > 1e-3
