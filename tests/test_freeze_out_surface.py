"""Freeze-out hypersurface (FreezeOut::bsqsvfreezeout / bsqsvinterpolate,
include/freeze_out.h:549-930; rows as Output::print_freeze_out writes them,
src/output.cpp:502-560).

Row layout: [particle id, freeze-out step, then the 40 columns of
freeze_out.dat]: divEener, divE[0..1], gsub, u[0..1], swsub, bulk, shear
00/11/22/33/12, tau, r[0..1], s*T^3, e, T, muB, muS, muQ, w, cs2, rhoB,
rhoS, rhoQ, diffusion[12], eos flag.

The oracle is pinned three ways: (1) a hand-built sequence of states where
the expected snapshot, time and row set of every particle follow from the
reference's timing rules; (2) brute force: the state of every particle is
recorded after every step and each row must carry, bit for bit, the values
of the step the snapshot logic had to pick; (3) physics: on ideal Gubser
flow the rows sit on the analytic isotherm, carry the analytic flow, and
their normal points along the analytic spacetime gradient of e, and the
conformal identities s = (e+p)/T, w = 4e/3, cs2 = 1/3 hold.  The GPU rows
then equal the oracle's to 1e-12.
"""
import math

import numpy as np
import pytest

import ccake_amd
from ccake_amd import config, ic

HB = config.HBARC_GEVFM
C_CONF = math.pi ** 2 * (2 * (3 * 3 - 1) + 3.5 * 3 * 2.5) / 90.0
E_FO_GEV = 3.36          # GeV/fm^3: the isotherm starts near r = 2.5 fm on the tau0 = 1 Gubser profile
COLS = ["divEener", "divE0", "divE1", "gsub", "u0", "u1", "swsub", "bulk", "shv00", "shv11", "shv22",
        "shv33", "shv12", "tau", "r0", "r1", "sT3", "e", "T", "muB", "muS", "muQ", "w", "cs2",
        "rhoB", "rhoS", "rhoQ"] + [f"diff{k}" for k in range(12)] + ["eos_table"]
C = {k: 2 + i for i, k in enumerate(COLS)}
# snapshot columns and the oracle field each one copies
SNAP_FIELDS = {"u0": "UX", "u1": "UY", "bulk": "BULK", "r0": "X", "r1": "Y", "e": "E", "muB": "MUB",
               "muS": "MUS", "muQ": "MUQ", "w": "W", "cs2": "CS2", "rhoB": "RHOB", "rhoS": "RHOS",
               "rhoQ": "RHOQ"}
SHEAR = {"shv00": 0, "shv11": 5, "shv22": 10, "shv33": 15, "shv12": 6}


def gubser_cfg(**kw):
    return config.gubser_ideal_params(particlization=1, e_freeze=E_FO_GEV / HB, dt=0.02, **kw)


def oracle_event(orc, cfg, d):
    eos = orc.Eos(0)
    ev = orc.Event(cfg, eos, len(d["x"]))
    ev._eos_keep = eos
    for k, f in [("x", "X"), ("y", "Y"), ("e", "IN_E"), ("ux", "UX"), ("uy", "UY"),
                 ("rhoB", "IN_B"), ("rhoS", "IN_SS"), ("rhoQ", "IN_Q")]:
        ev.set("ORC_F_" + f, d[k])
    assert ev.initialize() == 0
    return ev


def check_layout(rows, n):
    ids, steps = rows[:, 0].astype(np.int64), rows[:, 1].astype(np.int64)
    assert np.all(rows[:, 0] == ids) and np.all(rows[:, 1] == steps)
    assert len(np.unique(ids)) == len(ids), "a particle appears twice"
    assert np.all((ids >= 0) & (ids < n))
    key = steps * (n + 1) + ids
    assert np.all(np.diff(key) > 0), "rows not ordered by (step, id)"
    return ids, steps


# ------------------------------------------------------------------ (1) hand-built timing cases
def test_surface_timing_cases(orc):
    """Six particles driven through freeze_out_particles by hand (e_FO = 1,
    dt = 0.05, call k at tau = 1 + 0.05 (k-1)); every other field takes a
    fresh random value at every call, so a row names the call it came from:
      P0  e 2, 1.5, 0.95, 0.9       -> status 1 at call 3, 3 at call 4:
                                       row at step 4 from call 3 (closer to e_FO), tau = t3
      P1  as P0 but btrack -1 at 4  -> row at step 4 from call 3, tau = t3 - dt
      P2  as P0 but btrack 0 at 4   -> frozen without a row (freeze_no_neighbors)
      P3  e 2, 1.05, 0.5, 0.4       -> row at step 4 from call 2 (frz2), tau = t2
      P4  e 2, 1.5, 0.95, 0.97, 0.93 -> e rose at call 4 (stays 1); row at step 5 from call 4
      P5  e 0.9, 0.95 with btrack -1 at call 2 -> second-step interpolation: row at
          step 2 from call 2 (frz1, taken before interpolating), tau = t2 - dt."""
    cfg = config.base_params(e_freeze=1.0, particlization=1, dt=0.05)
    n = 6
    eos = orc.Eos(0)
    ev = orc.Event(cfg, eos, n)
    rng = np.random.default_rng(7)
    E = {1: [2.0, 2.0, 2.0, 2.0, 2.0, 0.9], 2: [1.5, 1.5, 1.5, 1.05, 1.5, 0.95],
         3: [0.95, 0.95, 0.95, 0.5, 0.95, 0.5], 4: [0.9, 1.2, 0.9, 0.4, 0.97, 0.5],
         5: [0.5, 0.5, 0.5, 0.5, 0.93, 0.5]}
    BT = {k: [5] * 6 for k in E}
    BT[2][5] = -1
    BT[4][1] = -1
    BT[4][2] = 0
    fz = ev.ifield("ORC_I_FREEZE")
    fz[:] = 0
    ev.set("ORC_F_MASS_S", rng.uniform(0.5, 2.0, n))
    hist = {}
    names = ["UX", "UY", "BULK", "X", "Y", "MUB", "MUS", "MUQ", "W", "CS2", "RHOB", "RHOS", "RHOQ", "T",
             "S", "SIGMA_LAB", "THETA", "SHV_NABLA_U", "GRADP0", "GRADE0"]
    for k in sorted(E):
        ev.t = 1.0 + 0.05 * (k - 1)
        st = {}
        for f in names:
            v = rng.uniform(0.5, 2.0, n) * (0.3 if f in ("UX", "UY") else 1.0)
            ev.set("ORC_F_" + f, v)
            st[f] = v
        g1 = rng.uniform(0.5, 2.0, n)
        ev.field("ORC_F_GRADP0")[:] = st["GRADP0"]
        ev.field(ev_fid(orc, "GRADP0", 1))[:] = g1
        ev.field(ev_fid(orc, "GRADE0", 1))[:] = g1[::-1].copy()
        shv = rng.uniform(-0.2, 0.2, (16, n))
        for c in range(16):
            ev.field(ev_fid(orc, "SHV0", c))[:] = shv[c]
        st["SHV"] = shv
        ev.set("ORC_F_E", E[k])
        ev.ifield("ORC_I_BTRACK")[:] = BT[k]
        ev.freeze_out_particles()
        hist[k] = st
        if k == 4:
            assert ev.ifield("ORC_I_BTRACK")[2] == -1 and ev.ifield("ORC_I_FREEZE")[2] == 4
    rows = ev.freeze_out_surface()
    ids, steps = check_layout(rows, n)
    # (particle, step, call whose state the row holds, tau)
    t = {k: 1.0 + 0.05 * (k - 1) for k in E}
    want = [(5, 2, 2, t[2] - 0.05), (0, 4, 3, t[3]), (1, 4, 3, t[3] - 0.05), (3, 4, 2, t[2]), (4, 5, 4, t[4])]
    assert [(int(a), int(b)) for a, b in zip(ids, steps)] == [(p, s) for p, s, _, _ in want]
    assert list(ev.ifield("ORC_I_FREEZE")) == [4, 4, 4, 4, 4, 4]
    for row, (p, s, call, tau) in zip(rows, want):
        st = hist[call]
        assert row[C["tau"]] == tau
        assert row[C["e"]] == E[call][p]
        for col, f in SNAP_FIELDS.items():
            if f != "E":
                assert row[C[col]] == st[f][p], (p, col)
        for col, c in SHEAR.items():
            assert row[C[col]] == st["SHV"][c][p], (p, col)
        assert row[C["T"]] == st["T"][p] * 0.1973269804
        assert np.all(row[C["diff0"]:C["diff11"] + 1] == 0.0)
        check_normals(row)


def ev_fid(orc, base, c):
    return orc.FIELDS["ORC_F_" + base] + c


def check_normals(row):
    """|n_tau^2 - |n|^2| = 1 for the normalised spacetime gradient of e
    (metric diag(1, -1, -1)) and gsub = sqrt(1 + |u|^2) (u^mu u_mu = 1)."""
    inv = row[C["divEener"]] ** 2 - row[C["divE0"]] ** 2 - row[C["divE1"]] ** 2
    assert abs(abs(inv) - 1.0) < 1e-12
    u2 = row[C["u0"]] ** 2 + row[C["u1"]] ** 2
    assert abs(row[C["gsub"]] ** 2 - (1.0 + u2)) < 1e-12 * (1.0 + u2)


# ------------------------------------------------------------------ (2) brute force over the step history
SNAPSHOT_STATE = ["X", "Y", "UX", "UY", "E", "T", "MUB", "MUS", "MUQ", "RHOB", "RHOS", "RHOQ", "W", "CS2",
                  "BULK", "S", "SIGMA_LAB", "THETA", "SHV_NABLA_U"]


def test_surface_rows_hold_the_right_step(orc):
    """Viscous conformal Gubser event (0.1 fm lattice, eta/s = 0.08): the
    state of every particle is recorded after every step.  A particle has a
    row iff it reached status 4 through status 3 with neighbours, and the
    row carries bit for bit the state after step k-1 or k-2 (k the row's
    step; after step 2 and 1 at the second step), whichever has e closer to
    e_FO, with tau the time of that step."""
    cfg = config.base_params(t0=1.0, particlization=1, e_freeze=E_FO_GEV / HB, dt=0.02,
                             constant_eta_over_s=0.08, zeta_mode=config.ZETA_CONSTANT,
                             constant_zeta_over_s=0.0)
    d = ic.gubser_ideal(spacing=0.1)
    n = len(d["x"])
    ev = oracle_event(orc, cfg, d)
    efc = cfg["e_freeze"]
    hist = {}
    status = {0: ev.ifield("ORC_I_FREEZE").copy()}
    via3 = np.zeros(n, bool)
    fo_step = np.full(n, -1)
    steps = 16
    for k in range(1, steps + 1):
        before = ev.ifield("ORC_I_FREEZE").copy()
        assert ev.advance_timestep(cfg["dt"], 2) == 0
        st = {f: ev.field("ORC_F_" + f).copy() for f in SNAPSHOT_STATE}
        st["SHV"] = {c: ev.field(ev_fid(orc, "SHV0", c)).copy() for c in SHEAR.values()}
        st["t"] = ev.t
        hist[k] = st
        after = ev.ifield("ORC_I_FREEZE").copy()
        assert np.all(ev.ifield("ORC_I_BTRACK")[after < 4] > 0)     # full lattice: nobody isolated
        newly = (before < 4) & (after == 4)
        via3 |= newly
        fo_step[newly] = k
        status[k] = after
    rows = ev.freeze_out_surface()
    ids, rsteps = check_layout(rows, n)
    print("rows", len(rows), "steps", np.unique(rsteps))
    assert len(rows) > 100
    assert np.array_equal(np.sort(ids), np.nonzero(via3)[0])
    assert np.array_equal(rsteps, fo_step[ids])
    for row, i, k in zip(rows, ids, rsteps):
        a, b = (k, k - 1) if k == 2 else (k - 1, k - 2)
        ea, eb = hist[a]["E"][i], hist[b]["E"][i]
        pick = a if abs(ea - efc) < abs(eb - efc) else b
        st = hist[pick]
        assert row[C["tau"]] == st["t"]
        for col, f in SNAP_FIELDS.items():
            assert row[C[col]] == st[f][i], (i, k, col)
        for col, c in SHEAR.items():
            assert row[C[col]] == st["SHV"][c][i], (i, k, col)
        check_normals(row)
    assert np.any(np.abs(rows[:, C["shv12"]]) > 0) and np.any(np.abs(rows[:, C["shv33"]]) > 0)


# ------------------------------------------------------------------ (3) physics on ideal Gubser flow
def gubser_physics(rows, ext_s):
    c = rows[:, 2:]
    k = lambda name: c[:, C[name] - 2]
    tau, r0, r1 = k("tau"), k("r0"), k("r1")
    r = np.hypot(r0, r1)
    e = k("e")
    efc = E_FO_GEV / HB
    # on the isotherm: the closer of two snapshots one step apart
    assert np.all(np.abs(e / efc - 1) < 0.03)
    # the analytic solution at the row's (tau, r), within the SPH smoothing error of the tau0 profile
    assert np.max(np.abs(e / ic.gubser_energy(tau, r) - 1)) < 0.07
    ur = np.hypot(k("u0"), k("u1"))
    assert np.max(np.abs(ur - ic.gubser_ur(tau, r))) < 0.02
    assert np.max(np.abs(k("u0") * r1 - k("u1") * r0)) < 1e-3            # radial flow
    # normal: (n_tau, n) along (-d_tau e, -grad e) of the analytic solution
    h = 1e-6
    dte = (ic.gubser_energy(tau + h, r) - ic.gubser_energy(tau - h, r)) / (2 * h)
    dre = (ic.gubser_energy(tau, r + h) - ic.gubser_energy(tau, r - h)) / (2 * h)
    nr = (k("divE0") * r0 + k("divE1") * r1) / r
    assert np.all(k("divEener") > 0) and np.all(nr > 0)
    assert np.max(np.abs(np.arctan2(k("divEener"), nr) - np.arctan2(-dte, -dre))) < 0.08
    assert np.max(np.abs(k("divE0") * r1 - k("divE1") * r0) / np.abs(nr)) < 0.01
    # conformal identities (T in GeV, e in fm^-4): e = 3 C T^4, s = (e + p)/T, w = 4e/3, cs2 = 1/3
    T = k("T") / HB
    assert np.max(np.abs(T / (e / (3 * C_CONF)) ** 0.25 - 1)) < 1e-12
    s = k("sT3") / k("T") ** 3
    assert np.max(np.abs(s / (4 * e / (3 * T)) - 1)) < 1e-12
    assert np.max(np.abs(k("w") / (4 * e / 3) - 1)) < 1e-12
    assert np.max(np.abs(k("cs2") - 1.0 / 3.0)) < 1e-14
    # swsub = (entropy mass)/(rest-frame SPH density sigma_lab/(gamma tau)): swsub * s
    # is the particle's specific entropy, constant in ideal flow, so the ratio to
    # the initial one is the same on every row (dropping the 1/tau of the
    # hyperbolic density would spread it by the rows' tau range, 4 to 38 percent);
    # its offset from 1 is the smoothing error of the tau0 profile (3.5 percent at r = 1 fm)
    ratio = k("swsub") * s / ext_s
    assert np.max(np.abs(ratio - 1)) < 0.07 and np.ptp(ratio) < 0.01
    assert np.all(k("bulk") == 0) and np.all(k("shv12") == 0) and np.all(k("eos_table") == 0)
    for row in rows:
        check_normals(row)


def test_surface_on_ideal_gubser_isotherm(orc):
    """Ideal Gubser flow (0.1 fm lattice), e_FO = 3.36 GeV/fm^3, 20 steps of
    0.02 fm/c: the rows lie on the analytic isotherm with the analytic flow
    and normal."""
    cfg = gubser_cfg()
    d = ic.gubser_ideal(spacing=0.1)
    ev = oracle_event(orc, cfg, d)
    for _ in range(20):
        assert ev.advance_timestep(cfg["dt"], 2) == 0
    rows = ev.freeze_out_surface()
    ids, steps = check_layout(rows, len(d["x"]))
    print("rows", len(rows))
    assert len(rows) > 300
    assert np.all(steps >= 3)        # status 1 then 3 takes two steps below e_FO
    taus = np.round(rows[:, C["tau"]], 12)
    assert set(taus) <= {round(1.0 + 0.02 * j, 12) for j in range(21)}
    gubser_physics(rows, ev.field("ORC_F_EXT_S")[ids])


def test_surface_requires_particlization(orc):
    cfg = config.gubser_ideal_params(particlization=0)
    d = ic.gubser_ideal(spacing=0.5)
    ev = oracle_event(orc, cfg, d)
    assert len(ev.freeze_out_surface()) == 0


# ------------------------------------------------------------------ GPU parity
def gpu_and_oracle(dev, orc, cfg, d, steps):
    eos = ccake_amd.Eos(dev, 0)
    g = ccake_amd.Event(dev, eos, cfg, len(d["x"]))
    ccake_amd.load_event(g, d)
    g.initialize()
    o = oracle_event(orc, cfg, d)
    for _ in range(steps):
        g.advance_timestep(cfg["dt"], 2)
        assert o.advance_timestep(cfg["dt"], 2) == 0
    gr, orr = g.freeze_out_surface(), o.freeze_out_surface()
    fz = g.download_int("CCAKE_I_FREEZE")
    g.close()
    eos.close()
    return gr, orr, fz, o


def assert_rows_equal(gr, orr):
    print("rows gpu", len(gr), "oracle", len(orr), flush=True)
    assert gr.shape == orr.shape
    assert np.array_equal(gr[:, :2], orr[:, :2]), "row sets differ"
    errs = {}
    for name, c in C.items():
        a, b = gr[:, c], orr[:, c]
        s = np.max(np.abs(b)) if len(b) else 0.0
        errs[name] = float(np.max(np.abs(a - b)) / s) if s > 0 else float(np.max(np.abs(a - b), initial=0.0))
    print("worst column:", max(errs.items(), key=lambda kv: kv[1]), flush=True)
    bad = {k: v for k, v in errs.items() if not v <= 1e-12}
    assert not bad, bad


@pytest.mark.gpu
def test_surface_parity_ideal_gubser(dev, orc):
    """The GPU hypersurface of the ideal Gubser run equals the oracle's
    (same rows, columns to 1e-12) and passes the same physics checks."""
    d = ic.gubser_ideal(spacing=0.1)
    gr, orr, fz, o = gpu_and_oracle(dev, orc, gubser_cfg(), d, 20)
    assert_rows_equal(gr, orr)
    assert np.array_equal(fz, o.ifield("ORC_I_FREEZE"))
    gubser_physics(gr, o.field("ORC_F_EXT_S")[gr[:, 0].astype(int)])


@pytest.mark.gpu
def test_surface_parity_viscous_gubser(dev, orc):
    """Viscous conformal run (eta/s = 0.08) with shear in the rows: GPU
    rows equal the oracle's."""
    cfg = config.base_params(t0=1.0, particlization=1, e_freeze=E_FO_GEV / HB, dt=0.02,
                             constant_eta_over_s=0.08, zeta_mode=config.ZETA_CONSTANT,
                             constant_zeta_over_s=0.0)
    d = ic.gubser_ideal(spacing=0.1)
    gr, orr, fz, o = gpu_and_oracle(dev, orc, cfg, d, 16)
    assert_rows_equal(gr, orr)
    assert np.any(gr[:, C["shv12"]] != 0)
