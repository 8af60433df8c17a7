"""BSQ diffusion on the GPU against the oracle (1e-12 relative per field,
as for the rest of the path): initial conditions, one evaluation of the
time derivatives, several RK2 steps and the freeze-out rows, for the
default kappa (a non-symmetric kappa/T^2 matrix, regulator on), the
baryon coth form with tau_q = C_B/T, shear viscosity on top, and the table
EoS.  Every call goes through the C-ABI (libccake_amd.so)."""
import numpy as np
import pytest

import ccake_amd
from ccake_amd import config, eos_tables, ic
from test_gpu_parity import DERIVS, STATE, TRANSPORT, T_FO, Pair, compare, fid, report_all

pytestmark = pytest.mark.gpu

KAPPA = [0.05, -0.02, 0.01, 0.03, 0.10, 0.00, 0.00, 0.02, 0.08]
# fields set by alpha = mu/T and its EoS derivatives
ALPHA = ([f"DALPHA0+{k}" for k in range(12)] + [f"GRAD_ALPHA0+{k}" for k in range(6)]
         + [f"KN_DIFF0+{k}" for k in range(3)])
DIFF = ([f"Q0+{k}" for k in range(12)] + [f"EXT_Q0+{k}" for k in range(9)]
        + [f"KAPPA_Q0+{k}" for k in range(9)] + [f"TAU_Q0+{k}" for k in range(9)])
DIFF_DERIVS = ([f"DEXT_Q0+{k}" for k in range(9)] + [f"GRAD_QA0+{k}" for k in range(3)]
               + [f"DIV_QA0+{k}" for k in range(3)])


def compare_inverse_reynolds(p, what):
    """|q_c| / |n_c| divides by a root-solver output that passes through
    zero where a charge changes sign: there its relative error is that of
    n_c, not of the path.  Compared as |q_c| (RE |n_c|) over all particles
    and as RE where |n_c| >= 0.1 max |n_c|, both to 1e-12."""
    for c, rho in enumerate(("RHOB", "RHOS", "RHOQ")):
        a, b = p.g.download(fid(f"RE_DIFF0+{c}")), p.o.field(fid(f"RE_DIFF0+{c}"))
        na, nb = np.abs(p.g.download(fid(rho))), np.abs(p.o.field(fid(rho)))
        qa, qb = a * na, b * nb
        s = np.max(np.abs(qb))
        if s == 0.0:
            assert np.max(np.abs(qa)) == 0.0
            continue
        assert np.max(np.abs(qa - qb)) <= 1e-12 * s, (what, c, "q")
        big = nb >= 0.1 * np.max(nb)
        assert np.max(np.abs(a[big] - b[big])) <= 1e-12 * np.max(np.abs(b[big])), (what, c, "RE")


def initial_q(d, seed=3, scale=2.0):
    """q^x, q^y of each charge: |q_c| = scale |n_c| U(0, 1), random direction
    (scale 2: the regulator caps a share of them)."""
    rng = np.random.default_rng(seed)
    n = len(d["x"])
    q = np.zeros((12, n))
    for c, rho in enumerate((d["rhoB"], d["rhoS"], d["rhoQ"])):
        ang = rng.uniform(0.0, 2.0 * np.pi, n)
        mag = scale * np.abs(rho) * rng.uniform(0.0, 1.0, n)
        q[4 * c + 1] = mag * np.cos(ang)
        q[4 * c + 2] = mag * np.sin(ang)
    return q


def diffusion_pair(dev, orc, cfg, d, q, eos_kind=0, table=None):
    print("diffusion pair:", len(d["x"]), "particles, mode", cfg["diffusion_mode"],
          "shear", cfg["using_shear"], "eos", eos_kind, flush=True)
    p = Pair(dev, orc, cfg, d, eos_kind=eos_kind, table=table)
    for k in range(12):
        p.g.upload(fid(f"Q0+{k}"), q[k])
        p.o.field(fid(f"Q0+{k}"))[:] = q[k]
    return p


def check_run(p, what, steps, alpha_tol=1e-12):
    """alpha_tol applies to the fields set by alpha = mu/T alone.  On the
    tanh-conformal table they inherit the conditioning of the EoS root: the
    derivatives d alpha_c / d(s, n) are 1 / (Schur complement) terms whose
    complements nearly cancel (|dalpha_Q/ds| reaches 3e7 fm; moving the root
    by 1e-12 relative moves it by 2e-3 of its maximum), and grad alpha sums
    differences of nearly equal alpha, so those fields are held to 1e-10
    there (measured: 1.3e-11 at most); every other field, and all fields with
    the conformal EoS, to 1e-12."""
    p.initialize()
    report_all(p.g, p.o, what + " init")
    compare(p.g, p.o, STATE + ["Q0+%d" % k for k in range(12)] + ["EXT_Q0+%d" % k for k in range(9)],
            what=what + " init")
    p.g.get_time_derivatives()
    assert p.o.get_time_derivatives() == 0
    report_all(p.g, p.o, what + " derivatives")
    compare(p.g, p.o, STATE + TRANSPORT + DERIVS + DIFF + DIFF_DERIVS, what=what + " derivatives")
    compare(p.g, p.o, ALPHA, tol=alpha_tol, what=what + " derivatives, alpha")
    compare_inverse_reynolds(p, what + " derivatives")
    assert np.max(np.abs(p.o.field(fid("DEXT_Q0+0")))) > 0.0
    p.step(steps)
    report_all(p.g, p.o, what + f" {steps} steps")
    compare(p.g, p.o, STATE + DIFF + DIFF_DERIVS + ["DUDT0+0", "DUDT0+1", "DEXT_B", "DEXT_SS", "DEXT_Q"],
            what=what + f" {steps} steps")
    compare(p.g, p.o, ALPHA, tol=alpha_tol, what=what + f" {steps} steps, alpha")
    compare_inverse_reynolds(p, what + f" {steps} steps")


def test_diffusion_default_kappa_with_regulator(dev, orc):
    d = ic.lumpy(spacing=0.1)
    cfg = config.lumpy_params(particlization=0, using_shear=0, constant_eta_over_s=0.0,
                              using_diffusion=1, kappa_over_T2=KAPPA, diffusion_regulator=1)
    p = diffusion_pair(dev, orc, cfg, d, initial_q(d))
    check_run(p, "diffusion default", 5)
    p.close()


def test_diffusion_coth_kappa_with_shear(dev, orc):
    d = ic.lumpy(spacing=0.1)
    d["rhoS"][:] = 0.0
    d["rhoQ"][:] = 0.0
    cfg = config.lumpy_params(particlization=0, using_diffusion=1,
                              diffusion_mode=config.DIFF_KAPPA_B_COTH,
                              diffusion_relax_mode=config.TAUQ_CB_OVER_T, C_B=0.4)
    assert cfg["using_shear"] == 1
    p = diffusion_pair(dev, orc, cfg, d, initial_q(d, seed=4, scale=0.5))
    check_run(p, "diffusion coth+shear", 5)
    assert np.max(p.o.field(fid("KAPPA_Q0+0"))) > 0.0
    p.close()


def test_diffusion_on_the_table_eos(dev, orc):
    table = eos_tables.tanh_conformal_table()
    d = ic.lumpy(spacing=0.1)
    cfg = config.lumpy_params(particlization=0, using_diffusion=1, kappa_over_T2=KAPPA)
    p = diffusion_pair(dev, orc, cfg, d, initial_q(d, seed=5, scale=0.5), eos_kind=1, table=table)
    check_run(p, "diffusion table", 3, alpha_tol=1e-10)
    p.close()


def test_freeze_out_rows_carry_the_diffusion_currents(dev, orc):
    """Lumpy run through freeze-out with diffusion on: the hypersurface rows
    (q^mu of B, S, Q in columns 27-38) equal the oracle's."""
    d = ic.lumpy(spacing=0.1)
    cfg = config.lumpy_params(using_diffusion=1, kappa_over_T2=KAPPA, diffusion_regulator=1)
    p = diffusion_pair(dev, orc, cfg, d, initial_q(d))
    p.initialize()
    p.step(25)
    gr, orr = p.g.freeze_out_surface(), p.o.freeze_out_surface()
    print("rows gpu", len(gr), "oracle", len(orr), flush=True)
    assert len(orr) > 20 and gr.shape == orr.shape
    assert np.array_equal(gr[:, :2], orr[:, :2])
    for c in range(2, ccake_amd.FO_ROW):
        a, b = gr[:, c], orr[:, c]
        s = np.max(np.abs(b))
        err = np.max(np.abs(a - b)) / s if s > 0 else np.max(np.abs(a - b))
        assert err <= 1e-12, (c, err)
    assert np.max(np.abs(orr[:, 2 + 27:2 + 39])) > 0.0          # the currents are in the rows
    p.close()


def test_phase_calls_equal_the_fused_stage_with_diffusion(dev, orc):
    """The phase-by-phase C-ABI (build_neighbors ... eom, each phase its own
    call) gives the derivatives of the fused stage, diffusion and shear on:
    q^mu rebuilt and regulated in reset_pi_tensor, kappa/tau_q in
    viscosities, the diffusion sums in smooth_gradients."""
    d = ic.lumpy(spacing=0.1)
    cfg = config.lumpy_params(particlization=0, using_diffusion=1, kappa_over_T2=KAPPA, diffusion_regulator=1)
    q = initial_q(d)
    n = len(d["x"])
    eos = ccake_amd.Eos(dev, 0, T_FO)
    evs = []
    for _ in range(2):
        ev = ccake_amd.Event(dev, eos, cfg, n)
        ccake_amd.load_event(ev, d)
        for k in range(12):
            ev.upload(fid(f"Q0+{k}"), q[k])
        ev.initialize()
        evs.append(ev)
    a, b = evs
    a.get_time_derivatives()
    for phase in ("build_neighbors", "gamma_velocity", "smooth_fields"):
        getattr(b, phase)()
    b.eos_invert(0)
    for phase in ("reset_pi_tensor", "viscosities", "smooth_gradients", "eom"):
        getattr(b, phase)()
    worst = 0.0
    for lab in STATE + TRANSPORT + DERIVS + DIFF + DIFF_DERIVS + ALPHA:
        x, y = a.download(fid(lab)), b.download(fid(lab))
        s = np.max(np.abs(x))
        e = float(np.max(np.abs(x - y)) / s) if s > 0 else float(np.max(np.abs(x - y)))
        worst = max(worst, e)
        assert e <= 1e-12, (lab, e)
    print("phase calls vs fused stage, worst", worst, flush=True)
    a.close(); b.close(); eos.close()
