"""Settings of the time-step path and the benchmark workloads.

`base_params()` restates the defaults of the reference's Settings for the
fields the path reads (include/defaults.h:14-66, include/settings.h), with
every transport mode set explicitly: the reference's defaults.h declares
several of them as `bool` initialised from strings, so a missing key falls
back to DEFAULT with a warning (src/transport_coefficients.cpp:9-283).
"""
from __future__ import annotations

HBARC_GEVFM = 0.1973269804          # include/constants.h:15
HBARC_MEVFM = 197.3269804           # include/constants.h:16

# transport modes (same numbering as include/ccake_amd.h)
ETA_DEFAULT, ETA_CONSTANT = 0, 1
TAUPI_DEFAULT, TAUPI_MINVAL, TAUPI_GUBSER = 0, 1, 2
ZETA_DEFAULT, ZETA_CONSTANT, ZETA_CS2_DEPENDENT, ZETA_GAUSSIAN = 0, 1, 2, 3
MODE_DEFAULT, MODE_DISABLED, MODE_IS = 0, 1, 2

EOS_CONFORMAL, EOS_TABLE = 0, 1
DIFF_KAPPA_DEFAULT, DIFF_KAPPA_B_COTH = 0, 2
TAUQ_DEFAULT, TAUQ_CB_OVER_T = 0, 1


def base_params(**kw):
    p = dict(
        dim=2, rk_order=2, hT=0.3, hEta=0.3, t0=1.0, dt=0.05, e_freeze=0.0,
        particlization=1, using_shear=1, using_vorticity=0,
        input_initial_shear=0, bulk_from_trace=0,
        prohibit_unstable_cs2=1, prohibit_acausal_cs2=1,
        eta_mode=ETA_CONSTANT, constant_eta_over_s=0.08,
        shear_relax_mode=TAUPI_DEFAULT,
        zeta_mode=ZETA_CS2_DEPENDENT, constant_zeta_over_s=0.005,
        cs2_dependent_zeta_A=1.67552, cs2_dependent_zeta_p=2.0,
        modulate_zeta_with_tanh=1,
        zeta_gauss_A=0.0, zeta_gauss_T0=0.0, zeta_gauss_width=1.0,
        delta_pipi_mode=MODE_IS, tau_pipi_mode=MODE_DISABLED, lambda_piPi_mode=MODE_DISABLED,
        phi6_mode=MODE_DISABLED, phi7_mode=MODE_DISABLED,
        delta_PiPi_mode=MODE_IS, lambda_Pipi_mode=MODE_DISABLED, phi1_mode=MODE_DISABLED,
        phi3_mode=MODE_DISABLED,
        cells_per_rc=3,
        using_diffusion=0, diffusion_mode=DIFF_KAPPA_DEFAULT, diffusion_relax_mode=TAUQ_DEFAULT,
        kappa_over_T2=[0.0] * 9, C_B=0.0,
        diffusion_regulator=0, regulator_chi0=10.0, regulator_e0=0.1, regulator_xi0=0.01, regulator_rq_max=1.0,
        pair_precision=0, factorized_kernel=0, x_subcells=8,
    )
    p.update(kw)
    # input.cpp:390-407: eta/s <= 1e-3 disables shear evolution
    if p["eta_mode"] == ETA_CONSTANT and p["constant_eta_over_s"] <= 1e-3:
        p["using_shear"] = 0
    return p


def gubser_ideal_params(**kw):
    """Ideal 2+1D Gubser flow: conformal EoS, eta/s = zeta/s = 0, hT = 0.3 fm,
    RK2, dt = 0.005 fm/c from tau0 = 1 fm/c, no particlization."""
    p = base_params(t0=1.0, dt=0.005, hT=0.3, particlization=0,
                    eta_mode=ETA_CONSTANT, constant_eta_over_s=0.0,
                    zeta_mode=ZETA_CONSTANT, constant_zeta_over_s=0.0)
    p.update(kw)
    if p["eta_mode"] == ETA_CONSTANT and p["constant_eta_over_s"] <= 1e-3:
        p["using_shear"] = 0
    return p


def lumpy_params(**kw):
    """Viscous BSQ lumpy event: table EoS, eta/s = 0.08, zeta/s a Gaussian
    of height 0.1 at T = 0.17 GeV (width 0.02 GeV), RK2, freeze-out at
    T = 150 MeV.  tau0 = 0.6 fm/c and dt = 0.02 fm/c are chosen here (the
    workload description leaves them open)."""
    p = base_params(t0=0.6, dt=0.02, hT=0.3, particlization=1,
                    eta_mode=ETA_CONSTANT, constant_eta_over_s=0.08,
                    zeta_mode=ZETA_GAUSSIAN, zeta_gauss_A=0.1,
                    zeta_gauss_T0=0.170/HBARC_GEVFM, zeta_gauss_width=0.020/HBARC_GEVFM)
    p.update(kw)
    return p
