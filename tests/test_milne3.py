"""The 3+1D Milne path (D = 3): the oracle's pins on the CPU and the HIP
library's parity with it on the GPU.

Oracle pins (no GPU):
1. Pair sets: the cell list equals O(N^2) brute force on a jittered
   (x, y, eta) lattice, with the predicate fl(fl(fl(dx^2) + fl(dy^2)) +
   fl(dz^2)) <= rc^2 of the build contract.
2. The smoothing weights: sigma_lab of a particle is the sum of
   SPHkernel<3> values over its neighbours, or of W_2D(r_perp, hT)
   W_1D(|d eta|, hEta) with the factorized kernel (src/kernel.cpp:129-170),
   the one-dimensional and two-dimensional kernels being pinned by the
   reference's known answers (tests/test_oracle.py).
3. A uniform fluid at rest in Cartesian coordinates seen in Milne
   coordinates (u^eta = -sinh(eta_s)/tau, gamma = cosh(eta_s)): an exact
   static solution of ideal hydrodynamics in which the tau^2 metric factors,
   the geometric factor u^eta u^eta tau / gamma and the 2 gamma u^eta / tau
   term of the acceleration must cancel.  The SPH particles sit uniformly
   in z = tau0 sinh(eta_s); T stays constant and u^eta follows
   -sinh(eta_s)/tau to the SPH discretisation error.
4. A viscous uniform fluid at rest, viscous in 3+1D and in 2+1D: its
   interior is the same 0+1D Bjorken flow in both, so the 3+1D shear and
   bulk equations at u = 0 must reproduce the 2+1D ones (whose shear
   equations are pinned by the viscous Gubser profiles).
The shear equations' terms with u^eta != 0 are pinned only by the GPU's
agreement with the oracle (both restate src/eom_default.cpp:882-1412): on
the static fluid of (3) with viscosity the reference's equations do not
keep the shear at zero (the eta component of F_i0_sigma + M_i0_sigma du/dt
does not vanish on that flow), so no closed form holds for them.
"""
import numpy as np
import pytest

import ccake_amd
from ccake_amd import config, eos_tables, ic
from parity_metrics import DERIVS3, STATE3, field_err, pp_tol, whole_field

T_FO = 150.0 / config.HBARC_MEVFM
TOL = 1e-12


def oracle_load(oev, d):
    for k, f in [("x", "X"), ("y", "Y"), ("eta", "ETA"), ("e", "IN_E"), ("ux", "UX"), ("uy", "UY"),
                 ("ueta", "UETA"), ("rhoB", "IN_B"), ("rhoS", "IN_SS"), ("rhoQ", "IN_Q")]:
        if k in d:
            oev.set("ORC_F_" + f, d[k])


def pair_set(pi, pj):
    return set(zip(pi.tolist(), pj.tolist()))


def jittered_lattice3(seed=7, n_side=9, a=0.2, jitter=0.05):
    rng = np.random.default_rng(seed)
    c = a * np.arange(n_side)
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    pts = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    pts = pts + rng.uniform(-jitter, jitter, size=pts.shape)
    # exact-boundary pairs: a particle at exactly rc along eta and along a diagonal
    extra = np.array([[0.0, 0.0, 10.0], [0.0, 0.0, 10.6], [0.3, 0.4, 20.0], [0.3, 0.4, 20.0 + np.sqrt(0.36 - 0.25)]])
    return np.concatenate([pts, extra], axis=0)


def milne_cfg(**kw):
    return config.lumpy_params(**dict(dict(dim=3, hEta=0.2, particlization=0), **kw))


# ---------------------------------------------------------------- oracle pins
def test_oracle_pairs_3d_equal_brute_force(orc):
    pts = jittered_lattice3()
    n = len(pts)
    cfg = milne_cfg()
    eos = orc.Eos(0, T_FO)
    sets = []
    for brute in (False, True):
        o = orc.Event(cfg, eos, n)
        o.set("ORC_F_X", pts[:, 0]); o.set("ORC_F_Y", pts[:, 1]); o.set("ORC_F_ETA", pts[:, 2])
        assert o.build_neighbors(brute) == 0
        sets.append(pair_set(*o.pairs()))
    assert sets[0] == sets[1] and len(sets[0]) > 0
    # the exact-boundary pair at |d eta| = rc is in, by the <= of the predicate
    assert (n - 4, n - 3) in sets[0] and (n - 3, n - 4) in sets[0]
    # symmetry
    assert all((j, i) in sets[0] for i, j in sets[0])


@pytest.mark.parametrize("factorized", [0, 1])
def test_oracle_smoothing_weights_3d(orc, factorized):
    """sigma_lab of each of a few particles equals the kernel sum over its
    neighbours computed here from the KAT-pinned 1D / 2D / 3D kernels."""
    pts = np.array([[0.0, 0.0, 0.0], [0.2, 0.1, 0.05], [-0.3, 0.2, -0.1], [0.1, -0.4, 0.3], [0.5, 0.3, 0.2]])
    cfg = milne_cfg(factorized_kernel=factorized)
    hT, hE = cfg["hT"], cfg["hEta"]
    eos = orc.Eos(0, T_FO)
    o = orc.Event(cfg, eos, len(pts))
    o.set("ORC_F_X", pts[:, 0]); o.set("ORC_F_Y", pts[:, 1]); o.set("ORC_F_ETA", pts[:, 2])
    o.set("ORC_F_GAMMA", np.ones(len(pts)))
    assert o.build_neighbors(False) == 0
    o.initial_sigma()
    L = orc.lib()

    def W(d):
        if factorized:
            return L.orc_kernel(2, float(np.hypot(d[0], d[1])), hT) * L.orc_kernel(1, float(abs(d[2])), hE)
        return L.orc_kernel(3, float(np.linalg.norm(d)), hT)

    rc = 2 * hT
    for i in range(len(pts)):
        want = W(np.zeros(3))
        for j in range(len(pts)):
            d = pts[i] - pts[j]
            if j != i and float(d @ d) <= rc * rc:
                want += W(d)
        assert o.field("ORC_F_SIGMA_LAB")[i] == pytest.approx(want, rel=1e-13)
    if factorized:   # W(0) = W_2D(0) W_1D(0) = (10/(7 pi hT^2)) (2/(3 hEta))
        assert W(np.zeros(3)) == pytest.approx(10 / (7 * np.pi * hT**2) * 2 / (3 * hE), rel=1e-14)


def static_milne_event(t0=1.0, a=0.15, half=1.5, dz=0.06, eta_max=1.4, e_gev=5.0):
    """A uniform fluid at rest in (t, z) on Milne coordinates: particles
    uniform in z = t0 sinh(eta_s), u^eta = -sinh(eta_s)/t0."""
    c = np.round(np.arange(-half, half + a / 2, a), 12)
    zmax = np.sinh(eta_max) * t0
    zs = np.arange(-np.floor(zmax / dz) * dz, zmax + dz / 2, dz)
    etas = np.arcsinh(zs / t0)
    Z, Y, X = np.meshgrid(etas, c, c, indexing="ij")
    x, y, eta = X.ravel(), Y.ravel(), Z.ravel()
    zero = np.zeros_like(x)
    return dict(x=x, y=y, eta=eta, e=np.full_like(x, e_gev / config.HBARC_GEVFM), ux=zero, uy=zero.copy(),
                ueta=-np.sinh(eta) / t0, rhoB=zero.copy(), rhoS=zero.copy(), rhoQ=zero.copy())


def static_cfg():
    return config.base_params(dim=3, t0=1.0, dt=0.01, particlization=0, constant_eta_over_s=0.0,
                              zeta_mode=config.ZETA_CONSTANT, constant_zeta_over_s=0.0, hT=0.3, hEta=0.15)


def check_static(get, t, sel, T0):
    T = get("T")[sel]
    eta = get("ETA")[sel]
    ue = get("UETA")[sel]
    d = dict(T=float(np.max(np.abs(T / T0 - 1))),
             ueta=float(np.max(np.abs(ue + np.sinh(eta) / t))),
             ux=float(max(np.max(np.abs(get("UX")[sel])), np.max(np.abs(get("UY")[sel])))))
    print("static fluid in Milne coordinates:", d, flush=True)
    # measured on the oracle: T 3e-5, u^eta 5e-4 (of |u^eta| <= 0.3), u_perp 3e-7
    assert d["T"] < 2e-4 and d["ueta"] < 3e-3 and d["ux"] < 1e-5
    return d


def test_oracle_static_fluid_in_milne_coordinates(orc):
    d = static_milne_event()
    cfg = static_cfg()
    o = orc.Event(cfg, orc.Eos(0, T_FO), len(d["x"]))
    oracle_load(o, d)
    assert o.initialize() == 0
    sel = (np.abs(d["x"]) < 0.31) & (np.abs(d["y"]) < 0.31) & (np.abs(d["eta"]) < 0.4)
    T0 = o.field("ORC_F_T")[sel].copy()
    for _ in range(10):
        assert o.advance_timestep(cfg["dt"], 2) == 0
    check_static(lambda f: o.field("ORC_F_" + f), o.t, sel, T0)


def test_oracle_viscous_rest_3d_equals_2d(orc):
    """Centre particle of a viscous conformal fluid at rest: 3+1D against
    2+1D after 5 RK2 steps (T, pi^{xx} = pi^{yy}, pi^{eta eta})."""
    eos = orc.Eos(0, T_FO)
    cfg2 = config.base_params(dim=2, t0=0.6, dt=0.02, particlization=0, constant_eta_over_s=0.2,
                              zeta_mode=config.ZETA_CONSTANT, constant_zeta_over_s=0.0)
    cfg3 = dict(cfg2, dim=3, hEta=0.3)
    e0 = 10.0 / config.HBARC_GEVFM
    x, y = ic.lattice(0.1, 2.0)
    o2 = orc.Event(cfg2, eos, len(x))
    o2.set("ORC_F_X", x); o2.set("ORC_F_Y", y); o2.set("ORC_F_IN_E", np.full_like(x, e0))
    x3, y3, z3 = ic.lattice3(0.1, 1.6, 0.1, 1.6)
    o3 = orc.Event(cfg3, eos, len(x3))
    o3.set("ORC_F_X", x3); o3.set("ORC_F_Y", y3); o3.set("ORC_F_ETA", z3); o3.set("ORC_F_IN_E", np.full_like(x3, e0))
    assert o2.initialize() == 0 and o3.initialize() == 0
    for _ in range(5):
        assert o2.advance_timestep(cfg2["dt"], 2) == 0 and o3.advance_timestep(cfg3["dt"], 2) == 0
    c2 = int(np.argmin(x * x + y * y))
    c3 = int(np.argmin(x3 * x3 + y3 * y3 + z3 * z3))
    sh = orc.FIELDS["ORC_F_SHV0"]
    pairs = [("T", o2.field("ORC_F_T")[c2], o3.field("ORC_F_T")[c3]),
             ("pi11", o2.field(sh + 5)[c2], o3.field(sh + 5)[c3]),
             ("pi22", o2.field(sh + 10)[c2], o3.field(sh + 10)[c3]),
             ("pi33", o2.field(sh + 15)[c2], o3.field(sh + 15)[c3])]
    for name, a, b in pairs:
        print(name, a, b, flush=True)
        assert abs(b - a) <= 2e-5 * abs(a), name     # edge effects reach the centre only at ~1e-7 by step 5
    assert abs(pairs[1][1]) > 0.1                     # the shear has built up


# ---------------------------------------------------------------- GPU parity
class Pair3:
    def __init__(self, dev, orc, cfg, d, table=None):
        n = len(d["x"])
        if table is None:
            self.geos, self.oeos = ccake_amd.Eos(dev, 0, T_FO), orc.Eos(0, T_FO)
        else:
            tab, dims, mins, maxs = table
            self.geos = ccake_amd.Eos(dev, 1, T_FO, dims, mins, maxs, tab)
            self.oeos = orc.Eos(1, T_FO, dims, mins, maxs, tab)
        if cfg["particlization"]:
            cfg = dict(cfg, e_freeze=self.geos.efreeze(T_FO))
        self.cfg = cfg
        self.g = ccake_amd.Event(dev, self.geos, cfg, n)
        ccake_amd.load_event(self.g, d)
        self.o = orc.Event(cfg, self.oeos, n)
        oracle_load(self.o, d)

    def close(self):
        self.g.close()
        self.geos.close()


def compare3(g, o, labels, what, tol=TOL):
    errs, pps = {}, {}
    for lab in labels:
        b0, c = (lab.split("+") + ["0"])[:2]
        f = ccake_amd.FIELDS["CCAKE_F_" + b0] + int(c)
        a, b = g.download(f), o.field(f)
        errs[lab] = whole_field(a, b)
        pps[lab] = field_err(lab, a, b, tol)
    print(what, "worst whole-field:", max(errs.items(), key=lambda kv: kv[1]),
          "worst per-particle:", max(pps.items(), key=lambda kv: kv[1]), flush=True)
    bad = {k: v for k, v in errs.items() if not v <= tol}
    bad.update({k: v for k, v in pps.items() if not v <= pp_tol(k, tol)})
    assert not bad, f"{what}: {bad}"


@pytest.fixture(scope="module")
def table():
    return eos_tables.tanh_conformal_table()


@pytest.mark.gpu
@pytest.mark.parametrize("factorized", [0, 1])
def test_gpu_pairs_3d_bit_exact(dev, orc, factorized):
    pts = jittered_lattice3()
    n = len(pts)
    cfg = milne_cfg(factorized_kernel=factorized)
    geos = ccake_amd.Eos(dev, 0, T_FO)
    g = ccake_amd.Event(dev, geos, cfg, n)
    for f, k in (("CCAKE_F_X", 0), ("CCAKE_F_Y", 1), ("CCAKE_F_ETA", 2)):
        g.upload(f, pts[:, k])
    g.build_neighbors()
    gi, gj = g.pairs()
    o = orc.Event(cfg, orc.Eos(0, T_FO), n)
    o.set("ORC_F_X", pts[:, 0]); o.set("ORC_F_Y", pts[:, 1]); o.set("ORC_F_ETA", pts[:, 2])
    assert o.build_neighbors(True) == 0
    oi, oj = o.pairs()
    assert len(gi) == len(oi) > 0
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj)
    assert np.array_equal(g.download_int("CCAKE_I_BTRACK"), o.ifield("ORC_I_BTRACK"))
    g.close(); geos.close()


@pytest.mark.gpu
@pytest.mark.parametrize("factorized", [0, 1])
def test_gpu_milne_lumpy_parity(dev, orc, table, factorized):
    """3+1D lumpy viscous BSQ event on the table EoS (0.1 fm and 0.1 in eta,
    21 x 21 x 21 lattice, 9,261 particles): initialize, one derivative
    evaluation and 5 RK2 steps against the oracle, every state and
    derivative field to 1e-12 (whole field and per particle), pair counts,
    freeze status and EoS members equal."""
    d = ic.milne_lumpy(spacing=0.1, half_width=1.0, eta_spacing=0.1, eta_max=1.0, center_sd=0.5)
    p = Pair3(dev, orc, milne_cfg(factorized_kernel=factorized), d, table)
    p.g.initialize()
    assert p.o.initialize() == 0
    compare3(p.g, p.o, STATE3, "after initialize")
    p.g.get_time_derivatives()
    assert p.o.get_time_derivatives() == 0
    compare3(p.g, p.o, STATE3 + DERIVS3, "after one derivative evaluation")
    for _ in range(5):
        p.g.advance_timestep(p.cfg["dt"], 2)
        assert p.o.advance_timestep(p.cfg["dt"], 2) == 0
    compare3(p.g, p.o, STATE3 + DERIVS3, "after 5 steps")
    for name in ("BTRACK", "FREEZE", "EOS_ID"):
        assert np.array_equal(p.g.download_int("CCAKE_I_" + name), p.o.ifield("ORC_I_" + name)), name
    assert abs(p.g.t - p.o.t) < 1e-14
    p.close()


@pytest.mark.gpu
def test_gpu_static_fluid_in_milne_coordinates(dev):
    d = static_milne_event()
    cfg = static_cfg()
    geos = ccake_amd.Eos(dev, 0, T_FO)
    g = ccake_amd.Event(dev, geos, cfg, len(d["x"]))
    ccake_amd.load_event(g, d)
    g.initialize()
    sel = (np.abs(d["x"]) < 0.31) & (np.abs(d["y"]) < 0.31) & (np.abs(d["eta"]) < 0.4)
    T0 = g.download("CCAKE_F_T")[sel].copy()
    g.run_steps(cfg["dt"], 10)
    check_static(lambda f: g.download("CCAKE_F_" + f), g.t, sel, T0)
    g.close(); geos.close()


@pytest.mark.gpu
def test_gpu_milne_run_to_freeze_out(dev, orc, table):
    """A small 3+1D lumpy event with the reference's energy cutoff runs until
    continue_evolution ends it (every particle frozen out) on the GPU and on
    the oracle, in the same number of steps within 3, with the frozen counts
    along the way within 1 % of the particles."""
    d = ic.milne_lumpy(spacing=0.2, half_width=2.0, eta_spacing=0.2, eta_max=2.0, center_sd=0.8,
                       e_cutoff_gev=0.15)
    p = Pair3(dev, orc, milne_cfg(particlization=1), d, table)
    n = len(d["x"])
    p.g.initialize()
    assert p.o.initialize() == 0
    steps_g = steps_o = 0
    frozen = []
    while True:
        kg = p.g.continue_evolution(max_tau=20.0)
        ko = bool(p.o.continue_evolution(20.0))
        if not kg and not ko:
            break
        if kg:
            p.g.advance_timestep(p.cfg["dt"], 2)
            steps_g += 1
        if ko:
            assert p.o.advance_timestep(p.cfg["dt"], 2) == 0
            steps_o += 1
        if kg and ko and steps_g % 20 == 0:
            frozen.append((p.g.reduce("CCAKE_R_FROZEN"), float(np.sum(p.o.ifield("ORC_I_FREEZE") >= 4))))
    print(f"{n} particles: steps GPU {steps_g}, oracle {steps_o}; frozen along the way {frozen}", flush=True)
    assert abs(steps_g - steps_o) <= 3
    assert all(abs(a - b) <= 0.01 * n for a, b in frozen)
    assert p.g.reduce("CCAKE_R_FROZEN") == n
    p.close()


@pytest.mark.gpu
def test_gpu_milne_rejects_unbuilt_combinations(dev):
    geos = ccake_amd.Eos(dev, 0, T_FO)
    for kw in (dict(using_diffusion=1), dict(pair_precision=1), dict(factorized_kernel=1, hEta=0.0)):
        with pytest.raises(ccake_amd.CcakeError):
            ccake_amd.Event(dev, geos, milne_cfg(**kw), 100)
    with pytest.raises(ccake_amd.CcakeError):
        ccake_amd.Batch(dev, geos, milne_cfg(), [100, 100])
    geos.close()
