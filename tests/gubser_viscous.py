"""Semi-analytic Israel-Stewart Gubser flow (conformal, eta/s = 0.2,
tau_pi = 5 eta/(e+P)), restated from its defining equations (Marrochio et
al., Phys. Rev. C 91, 014903, eqs. 2-3), as tabulated by the reference's
misc/Gubser_checks/ac/Gubser_semi-analytical_tabulator.wls:
  dT/drho / T + (2/3) tanh rho = (1/3) pi_bar tanh rho,
  (c/T)(eta/s)(dpi_bar/drho + (4/3) pi_bar^2 tanh rho) + pi_bar = (4/3)(eta/s)/T tanh rho,
with T(0) = 1.2, pi_bar(0) = 0, c = 5, and the map to Milne coordinates
rho = asinh(-(1 - tau^2 + r^2)/(2 tau)).  Solved here with scipy's DOP853;
tests/test_gubser_viscous.py pins the result against the reference's
tabulated profiles (tests/golden/gubser_viscous)."""
from __future__ import annotations

import math
import os

import numpy as np
from scipy.integrate import solve_ivp

HBARC = 0.19732698          # GeV fm, as in the tabulator
ETA_S = 0.2
C_TAU = 5.0
T_HAT0 = 1.2
NC, NF = 3.0, 2.5
ALPHA = 3 * math.pi ** 2 / 90 * (2 * (NC ** 2 - 1) + 3.5 * NC * NF)   # e / T^4
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "gubser_viscous")


def _rhs(rho, y):
    T, p = y
    th = math.tanh(rho)
    dT = T * th * (p / 3.0 - 2.0 / 3.0)
    dp = ((4.0 / 3.0) * ETA_S / T * th - p) * T / (C_TAU * ETA_S) - (4.0 / 3.0) * p * p * th
    return [dT, dp]


class Solution:
    """T_hat(rho) and pi_bar(rho) on [-12, 12] (dense output)."""

    def __init__(self):
        kw = dict(method="DOP853", rtol=1e-12, atol=1e-14, dense_output=True)
        self.fwd = solve_ivp(_rhs, (0.0, 12.0), [T_HAT0, 0.0], **kw)
        self.bwd = solve_ivp(_rhs, (0.0, -12.0), [T_HAT0, 0.0], **kw)

    def at(self, rho):
        rho = np.asarray(rho, dtype=np.float64)
        out = np.empty((2,) + rho.shape)
        m = rho >= 0
        if np.any(m):
            out[:, m] = self.fwd.sol(rho[m])
        if np.any(~m):
            out[:, ~m] = self.bwd.sol(rho[~m])
        return out[0], out[1]

    def fields(self, tau, x, y):
        """T (GeV), u^x, u^y, pi^xx, pi^yy, pi^xy (GeV/fm^3), tau^2 pi^etaeta (GeV/fm^3)."""
        x = np.asarray(x, dtype=np.float64)
        y = np.asarray(y, dtype=np.float64)
        r2 = x * x + y * y
        r = np.sqrt(r2)
        rho = np.arcsinh(-(1.0 - tau * tau + r2) / (2.0 * tau))
        That, pbar = self.at(rho)
        T = HBARC * That / tau
        ur = np.sinh(np.arctanh(2.0 * tau * r / (1.0 + tau * tau + r2)))
        with np.errstate(invalid="ignore", divide="ignore"):
            cx = np.where(r > 0, x / np.where(r > 0, r, 1.0), 0.0)
            cy = np.where(r > 0, y / np.where(r > 0, r, 1.0), 0.0)
        pihat = 4.0 / 3.0 * ALPHA * HBARC * That ** 4 * pbar
        pixx = -0.5 / tau ** 4 * (1.0 + cx * cx * ur * ur) * pihat
        piyy = -0.5 / tau ** 4 * (1.0 + cy * cy * ur * ur) * pihat
        pixy = -0.5 / tau ** 4 * cx * cy * ur * ur * pihat
        return T, ur * cx, ur * cy, pixx, piyy, pixy, pihat / tau ** 4


def golden(name):
    """A tabulated profile of the reference: columns x, y, T, u^x, u^y,
    pi^xx, pi^yy, pi^xy, tau^2 pi^etaeta."""
    return np.loadtxt(os.path.join(GOLDEN, name))


def initial_condition(sol, spacing=0.05, half_width=5.0, tau0=1.0):
    """Lattice initial condition in the units the library takes: e and
    shear in fm^-4 (GeV/fm^3 divided by hbar c), u dimensionless."""
    from ccake_amd import ic
    x, y = ic.lattice(spacing, half_width)
    T, ux, uy, pxx, pyy, pxy, _ = sol.fields(tau0, x, y)
    Tf = T / HBARC
    e = ALPHA * Tf ** 4                                  # conformal: e = 3p = alpha T^4
    z = np.zeros_like(x)
    return dict(x=x, y=y, e=e, ux=ux, uy=uy, rhoB=z, rhoS=z.copy(), rhoQ=z.copy(),
                pixx=pxx / HBARC, pixy=pxy / HBARC, piyy=pyy / HBARC)


def params(**kw):
    """Settings of the check: conformal EoS, eta/s = 0.2, tau_pi = 5 eta/w
    (the tabulator's c), delta_pipi = 4 tau_pi / 3 and no other second-order
    terms (eqs. 2-3), zeta = 0, initial shear from the profile."""
    from ccake_amd import config
    p = config.base_params(t0=1.0, dt=0.005, hT=0.3, particlization=0,
                           eta_mode=config.ETA_CONSTANT, constant_eta_over_s=ETA_S,
                           shear_relax_mode=config.TAUPI_GUBSER,
                           zeta_mode=config.ZETA_CONSTANT, constant_zeta_over_s=0.0,
                           delta_pipi_mode=config.MODE_DEFAULT, tau_pipi_mode=config.MODE_DISABLED,
                           lambda_piPi_mode=config.MODE_DISABLED, phi7_mode=config.MODE_DISABLED,
                           input_initial_shear=1, using_shear=1)
    p.update(kw)
    return p
