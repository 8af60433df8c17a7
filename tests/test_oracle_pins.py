"""Closed-form pins of oracle parts that no reference fixture covers.

1. The bulk-pressure equations of motion (src/eom_default.cpp:306-347,
   761-763, and the entropy term of 327): in the interior of a uniform
   fluid at rest on a large lattice every gradient vanishes and u = 0, so
   each interior particle is a 0+1D Bjorken fluid.  With the reference's
   F_extensive_bulk and F_extensive_entropy at u = 0, sigma = sigma_lab / tau,
   conserved charges (dN/dt = 0) and delta_PiPi = tau_Pi (the IS mode, so
   zeta_tilde = zeta):
       d Pi  / d tau = -(zeta / tau + Pi) / tau_Pi - Pi / tau
       d s   / d tau = -s / tau - Pi / (tau T)      (de/dtau = -(e+p+Pi)/tau)
       d rho / d tau = -rho / tau
   solved here with an adaptive Runge-Kutta method, the EoS taken from the
   oracle's own point inversion (pinned by test_oracle.py).
2. The gradient pair sums (src/sph_workstation.cpp:796-986) on linear
   profiles: on a uniform lattice the SPH operators reproduce the slope of
   a linear field to the lattice's discretisation error, for grad E,
   grad P, grad Pi, grad v, divshear and gradshear.
Each test also shows that plausible mistakes in the equations (a dropped
term, a flipped sign, a factor) move the answer far outside its tolerance.
"""
import numpy as np
import pytest
from scipy.integrate import solve_ivp

import ccake_amd
from ccake_amd import config, eos_tables

T_FO = 150.0 / config.HBARC_MEVFM


def lattice(half, a):
    n = int(round(2 * half / a)) + 1
    c = -half + a * np.arange(n)
    x, y = np.meshgrid(c, c, indexing="xy")
    return x.ravel(), y.ravel()


def load(o, x, y, e, rhoB=None):
    z = np.zeros_like(x)
    b = z if rhoB is None else rhoB
    for f, v in (("X", x), ("Y", y), ("IN_E", e), ("UX", z), ("UY", z), ("IN_B", b), ("IN_SS", z), ("IN_Q", z)):
        o.set("ORC_F_" + f, v)


# ---------------------------------------------------------------- bulk
@pytest.fixture(scope="module")
def bjorken(orc):
    """A uniform fluid at rest (e = 6 GeV/fm^3, rho_B = 0.1 fm^-3) on the 0.1 fm
    lattice over [-4, 4]^2 with the table EoS, shear off, the Gaussian zeta/s
    of the lumpy workload scaled up tenfold (so the bulk pressure grows to a
    few percent of p), 50 RK2 steps from tau0 = 0.6 fm/c."""
    tab, dims, mins, maxs = eos_tables.tanh_conformal_table()
    eos = orc.Eos(1, T_FO, dims, mins, maxs, tab)
    cfg = config.lumpy_params(eta_mode=config.ETA_CONSTANT, constant_eta_over_s=0.0, using_shear=0, particlization=0,
                              zeta_gauss_A=1.0)
    assert cfg["using_shear"] == 0 and cfg["delta_PiPi_mode"] == config.MODE_IS
    x, y = lattice(4.0, 0.1)
    e0 = 6.0 / config.HBARC_GEVFM
    o = orc.Event(cfg, eos, len(x))
    load(o, x, y, np.full_like(x, e0), np.full_like(x, 0.1))
    assert o.initialize() == 0
    s0 = o.field("ORC_F_S").copy()
    b0 = o.field("ORC_F_RHOB").copy()
    steps = 50
    for _ in range(steps):
        assert o.advance_timestep(cfg["dt"], 2) == 0
    inner = (np.abs(x) < 1.0) & (np.abs(y) < 1.0)
    return dict(o=o, eos=eos, cfg=cfg, inner=inner, s0=float(np.median(s0[inner])), b0=float(np.median(b0[inner])),
                t1=o.t)


def thermo_at(eos, s, rhoB):
    ok, _, _, th = eos.invert_point(0, (s, rhoB, 0.0, 0.0), (0.8, 0.01, 0.0, 0.0))
    assert ok
    return th


def bjorken_ode(eos, cfg, s0, b0, t0, t1, sign_zeta=1.0, delta=None, entropy_term=1.0, relax=1.0):
    """The interior's ODE; the keyword arguments plant mistakes."""
    gA, gT0, gW, p_exp = cfg["zeta_gauss_A"], cfg["zeta_gauss_T0"], cfg["zeta_gauss_width"], cfg["cs2_dependent_zeta_p"]

    def rhs(tau, yv):
        s, Pi = yv
        th = thermo_at(eos, s, b0 * t0 / tau)
        T, cs2, w = th[0], th[10], th[11]
        zeta = gA * np.exp(-(T - gT0) ** 2 / (2 * gW * gW)) * w / T     # Gaussian zeta/s times s (s = w/T)
        cc = min(cs2, 1.0)
        pw = (1.0 / 3.0 - cc) ** p_exp
        tau_Pi = 1e10 if pw < 1e-10 else max(5.0 * zeta / (pw * w), 0.1)
        d_PiPi = tau_Pi if delta is None else delta * tau_Pi
        zt = zeta + (d_PiPi - tau_Pi) * Pi
        dPi = -(sign_zeta * zt / tau + relax * Pi) / tau_Pi - Pi / tau
        ds = -s / tau - entropy_term * Pi / (tau * T)
        return [ds, dPi]

    sol = solve_ivp(rhs, (t0, t1), [s0, 0.0], method="DOP853", rtol=1e-11, atol=1e-14)
    return sol.y[0, -1], sol.y[1, -1]


def interior_at_t1(bj):
    """s, Pi and rho_B of the interior at the step's end time t1 from the
    evolved extensive quantities (the thermodynamic fields belong to the
    last stage's time t1 - dt/2): x = x_ext sigma_lab / (gamma t1)."""
    o, inner, t1 = bj["o"], bj["inner"], bj["t1"]
    sl = o.field("ORC_F_SIGMA_LAB")[inner] / o.field("ORC_F_GAMMA")[inner] / t1
    return (o.field("ORC_F_EXT_S")[inner] * sl, o.field("ORC_F_EXT_BULK")[inner] * sl,
            o.field("ORC_F_EXT_B")[inner] * sl)


def test_bulk_pressure_follows_the_bjorken_ode(bjorken):
    o, inner = bjorken["o"], bjorken["inner"]
    s_sim, Pi_sim, b_sim = interior_at_t1(bjorken)
    p_m = float(np.median(o.field("ORC_F_P")[inner]))
    # the interior is uniform and at rest (the edges' rarefaction reaches it
    # only through the kernel's overlap over the 100 stages, at 1e-8)
    assert np.max(np.abs(o.field("ORC_F_UX")[inner])) < 1e-6
    assert np.ptp(s_sim) <= 1e-8 * np.max(s_sim) and np.ptp(Pi_sim) <= 1e-6 * np.max(np.abs(Pi_sim))
    t0, t1 = bjorken["cfg"]["t0"], bjorken["t1"]
    s_ode, Pi_ode = bjorken_ode(bjorken["eos"], bjorken["cfg"], bjorken["s0"], bjorken["b0"], t0, t1)
    s_m, Pi_m, b_m = float(np.median(s_sim)), float(np.median(Pi_sim)), float(np.median(b_sim))
    print(f"tau {t1:.3f}: s {s_m:.8e} (ODE {s_ode:.8e}), Pi {Pi_m:.6e} (ODE {Pi_ode:.6e}), Pi/p {Pi_m / p_m:.4f}")
    assert abs(b_m * t1 / (bjorken["b0"] * t0) - 1) < 1e-9         # rho tau conserved
    assert Pi_m < 0 and abs(Pi_m / p_m) > 0.01                      # the bulk pressure is live
    assert abs(s_m / s_ode - 1) < 1e-4
    assert abs(Pi_m / Pi_ode - 1) < 1e-4


@pytest.mark.parametrize("mistake", [dict(sign_zeta=-1.0), dict(delta=2.0 / 3.0), dict(entropy_term=0.0),
                                     dict(relax=2.0), dict(delta=0.0)])
def test_bulk_ode_pin_rejects_planted_mistakes(bjorken, mistake):
    """Each planted mistake (the zeta term's sign, delta_PiPi = 2 tau_Pi / 3
    or 0 instead of tau_Pi, the bulk term of the entropy equation dropped,
    the relaxation term doubled) moves Pi or s by far more than the 1e-4
    the oracle meets."""
    s_sim, Pi_sim, _ = interior_at_t1(bjorken)
    s_m, Pi_m = float(np.median(s_sim)), float(np.median(Pi_sim))
    s_bad, Pi_bad = bjorken_ode(bjorken["eos"], bjorken["cfg"], bjorken["s0"], bjorken["b0"], bjorken["cfg"]["t0"],
                                bjorken["t1"], **mistake)
    dev = max(abs(s_m / s_bad - 1), abs(Pi_m / Pi_bad - 1))
    print(mistake, f"deviation {dev:.3e}")
    assert dev > 1e-3


# ---------------------------------------------------------------- gradients
@pytest.fixture(scope="module")
def gradients(orc):
    """Linear profiles on the 0.05 fm lattice over [-3, 3]^2 (h = 0.3 fm,
    shear on): each gradient pair sum of the oracle at the interior
    particles (|x|, |y| < 2 fm, more than 2h from the edges)."""
    x, y = lattice(3.0, 0.05)
    n = len(x)
    cfg = config.base_params(hT=0.3, t0=1.0, using_shear=1)
    o = orc.Event(cfg, orc.Eos(0), n)
    load(o, x, y, np.full_like(x, 1.0))
    L = orc.lib()
    assert L.orc_build_neighbors(o.h, 0) == 0
    L.orc_smooth_fields(o.h)              # sigma_lab (masses 1, gamma 1)
    prof = dict(
        E=(2.0, 0.3, -0.5), P=(1.0, -0.2, 0.45), BULK=(-0.1, 0.05, 0.02),
        VX=(0.01, 0.03, 0.07), VY=(-0.02, 0.05, -0.04),
        PI11=(0.3, 0.11, -0.13), PI12=(0.1, 0.17, 0.19), PI22=(0.25, -0.23, 0.29),
        PI01=(0.05, 0.31, -0.37), PI02=(-0.04, 0.41, 0.43))
    lin = {k: c0 + cx * x + cy * y for k, (c0, cx, cy) in prof.items()}
    o.set("ORC_F_E", lin["E"]); o.set("ORC_F_P", lin["P"]); o.set("ORC_F_BULK", lin["BULK"])
    o.set("ORC_F_VX", lin["VX"]); o.set("ORC_F_VY", lin["VY"])
    shv = np.zeros((16, n))
    shv[5] = lin["PI11"]; shv[6] = shv[9] = lin["PI12"]; shv[10] = lin["PI22"]
    shv[1] = shv[4] = lin["PI01"]; shv[2] = shv[8] = lin["PI02"]
    for c in range(16):
        o.set(ccake_amd.FIELDS["CCAKE_F_SHV0"] + c, shv[c])
    L.orc_smooth_gradients(o.h)
    inner = (np.abs(x) < 2.0) & (np.abs(y) < 2.0)
    g = lambda name, c: o.field(ccake_amd.FIELDS["CCAKE_F_" + name] + c)[inner]
    return dict(g=g, prof=prof, v=(lin["VX"][inner], lin["VY"][inner]))


# the lattice's discretisation error of a first moment (h / a = 6): 5e-5
TOL_GRAD = 5e-4


def check(vals, want, scale):
    err = float(np.max(np.abs(vals - want))) / scale
    return err


def test_linear_gradients_are_the_slopes(gradients):
    g, pr = gradients["g"], gradients["prof"]
    rows = []
    for name, field in (("E", "GRADE0"), ("P", "GRADP0"), ("BULK", "GRADBULK0")):
        for c, want in enumerate(pr[name][1:]):
            rows.append((f"{field}+{c}", check(g(field, c), want, abs(want))))
    # grad v: GRADV0 + 3 d + e = d_d v^e
    for d in range(2):
        for e, comp in enumerate(("VX", "VY")):
            want = pr[comp][1 + d]
            rows.append((f"GRADV0+{3 * d + e}", check(g("GRADV0", 3 * d + e), want, abs(want))))
    # divshear_j = d_i pi^{j i}
    div = (pr["PI11"][1] + pr["PI12"][2], pr["PI12"][1] + pr["PI22"][2])
    for j in range(2):
        rows.append((f"DIVSHEAR0+{j}", check(g("DIVSHEAR0", j), div[j], abs(div[j]))))
    # gradshear_d = v^j d_j pi^{0 d} (v of the owner particle)
    vx, vy = gradients["v"]
    for d, comp in enumerate(("PI01", "PI02")):
        want = vx * pr[comp][1] + vy * pr[comp][2]
        rows.append((f"GRADSHEAR0+{d}", check(g("GRADSHEAR0", d), want, float(np.max(np.abs(want))))))
    print(" ".join(f"{k}={v:.1e}" for k, v in rows))
    assert all(v < TOL_GRAD for _, v in rows), rows


def test_gradient_pins_reject_planted_mistakes(gradients):
    """The answers of plausible mistakes -- a flipped sign, the transposed
    velocity gradient, divshear over the other index pair, a factor
    sigma_lab^2 -- are each far outside TOL_GRAD of the oracle's values."""
    g, pr = gradients["g"], gradients["prof"]
    sl2 = (1.0 / 0.05 ** 2) ** 2
    far = 10 * TOL_GRAD
    assert check(g("GRADE0", 0), -pr["E"][1], abs(pr["E"][1])) > far
    assert check(g("GRADV0", 1), pr["VX"][2], abs(pr["VX"][2])) > far           # d_y v^x for d_x v^y
    div_t = pr["PI11"][1] + pr["PI12"][1]                                       # d_j pi^{j x} summed wrongly
    assert check(g("DIVSHEAR0", 0), div_t, abs(div_t)) > far
    assert check(g("GRADP0", 0), pr["P"][1] * sl2, abs(pr["P"][1] * sl2)) > far


# ---------------------------------------------------------------- RK4
def test_rk4_step_is_the_references_sequence(orc):
    """One step of Evolver::advance_timestep_rk4 (src/evolver.cpp:519-561)
    on an ideal conformal fluid at rest: in the interior nothing moves and
    the extensive entropy is constant, so only E_z evolves, with the rate
    g(t) = p(s_ext sigma_lab / t) m / sigma_lab (src/eom_default.cpp:855)
    that depends on t alone.  The reference's step then gives exactly
        E_z(t0 + dt) = E_z(t0) + dt (g(t0+dt/2)/6 + 2 g(t0+dt/2)/6
                                     + 2 g(t0+dt/2)/6 + 2 g(t0+dt)/6),
    weights summing to 7/6 (the first stage's derivative is not kept); the
    classical step, dt (g(t0) + 4 g(t0+dt/2) + g(t0+dt)) / 6, differs by
    far more than the rounding."""
    x, y = lattice(3.0, 0.1)
    cfg = config.gubser_ideal_params(dt=0.05)
    eos = orc.Eos(0)
    o = orc.Event(cfg, eos, len(x))
    load(o, x, y, np.full_like(x, 10.0 / config.HBARC_GEVFM))
    assert o.initialize() == 0
    inner = (np.abs(x) < 1.0) & (np.abs(y) < 1.0)
    t0, dt = o.t, cfg["dt"]
    ez0 = o.field("ORC_F_EZ")[inner].copy()
    sext = o.field("ORC_F_EXT_S")[inner].copy()
    sl = o.field("ORC_F_SIGMA_LAB")[inner].copy()
    m = o.field("ORC_F_MASS_S")[inner].copy()
    assert o.advance_timestep(dt, 4) == 0
    assert abs(o.t - (t0 + dt)) < 1e-15
    assert np.max(np.abs(o.field("ORC_F_UX")[inner])) < 1e-10       # at rest to rounding
    assert np.array_equal(o.field("ORC_F_EXT_S")[inner], sext)

    def g(t):
        out = np.empty(len(sext))
        for k in range(len(sext)):
            ok, _, _, th = eos.invert_point(0, (sext[k] * sl[k] / t, 0.0, 0.0, 0.0), (1.0, 0.0, 0.0, 0.0))
            assert ok
            out[k] = th[9] * m[k] / sl[k]
        return out

    gh, g1, g0 = g(t0 + 0.5 * dt), g(t0 + dt), g(t0)
    want = ez0 + dt * (gh / 6.0 + 2.0 * gh / 6.0 + 2.0 * gh / 6.0 + 2.0 * g1 / 6.0)
    classical = ez0 + dt * (g0 + 4.0 * gh + g1) / 6.0
    got = o.field("ORC_F_EZ")[inner]
    err = np.max(np.abs(got - want) / np.abs(want - ez0))
    far = np.max(np.abs(classical - want) / np.abs(want - ez0))
    print(f"RK4 E_z: relative error {err:.2e}; classical RK4 would be {far:.2e} away")
    # the Newton roots of g's evaluations and the step's own differ at 1e-12
    assert err < 1e-10 and far > 0.1


def test_rk4_update_writes_the_extensive_bulk_into_bulk(orc):
    """update_rk4 stores the updated extensive bulk pressure into the bulk
    field as well (src/evolver.cpp:318): after a fourth-order step of a
    bulk-viscous fluid the two fields are equal (and non-zero)."""
    tab, dims, mins, maxs = eos_tables.tanh_conformal_table()
    eos = orc.Eos(1, T_FO, dims, mins, maxs, tab)
    cfg = config.lumpy_params(eta_mode=config.ETA_CONSTANT, constant_eta_over_s=0.0, using_shear=0, particlization=0,
                              zeta_gauss_A=1.0)
    x, y = lattice(2.0, 0.1)
    o = orc.Event(cfg, eos, len(x))
    load(o, x, y, np.full_like(x, 6.0 / config.HBARC_GEVFM), np.full_like(x, 0.1))
    assert o.initialize() == 0
    assert o.advance_timestep(cfg["dt"], 4) == 0
    b, eb = o.field("ORC_F_BULK"), o.field("ORC_F_EXT_BULK")
    assert np.any(eb != 0.0) and np.array_equal(b, eb)
