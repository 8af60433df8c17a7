"""Synthetic initial conditions for the benchmark workloads.

Energy densities are returned in fm^-4 (the reference's internal unit:
src/BSQHydro.cpp:272-275 divides the file's GeV/fm^3 by hbar c), charges
in fm^-3.  Formulas are restated from their physics definitions; the ideal
Gubser profile is the closed form of utilities/ideal_gubser_ic_generator.py
(eps_a, kappa, velocity_x/y) and misc/Gubser_checks/plot_gubser_profiles.py.
"""
from __future__ import annotations

import numpy as np

from .config import HBARC_GEVFM


def lattice(spacing, half_width):
    """Square lattice over [-L, L]^2 with the given spacing, x fastest.
    Coordinates are the nearest doubles to their decimal values (as read
    from a text initial-condition file), so that pair distances of exact
    multiples of the spacing round the same way on every machine."""
    n = int(round(2 * half_width / spacing)) + 1
    g = np.round(-half_width + spacing * np.arange(n), 12) + 0.0
    x, y = np.meshgrid(g, g, indexing="xy")
    return x.ravel().copy(), y.ravel().copy()


def gubser_energy(tau, r, q=1.0, eps0_gev=80.0):
    """Ideal Gubser energy density (fm^-4):
    eps = eps0 (2q)^{8/3} / (tau^{4/3} [1 + 2q^2(tau^2+r^2) + q^4 (tau^2-r^2)^2]^{4/3})."""
    den = 1.0 + 2.0 * q**2 * (tau**2 + r**2) + q**4 * (tau**2 - r**2) ** 2
    return (eps0_gev / HBARC_GEVFM) * (2.0 * q) ** (8.0 / 3.0) / (tau ** (4.0 / 3.0) * den ** (4.0 / 3.0))


def gubser_ur(tau, r, q=1.0):
    """Radial four-velocity u^r = sinh(artanh(2 q^2 tau r / (1 + q^2 tau^2 + q^2 r^2)))."""
    return np.sinh(np.arctanh(2.0 * q**2 * tau * r / (1.0 + (q * tau) ** 2 + (q * r) ** 2)))


def gubser_ideal(spacing=0.05, half_width=5.0, tau0=1.0, q=1.0, eps0_gev=80.0):
    x, y = lattice(spacing, half_width)
    r = np.sqrt(x * x + y * y)
    e = gubser_energy(tau0, r, q, eps0_gev)
    ur = gubser_ur(tau0, r, q)
    with np.errstate(invalid="ignore", divide="ignore"):
        ux = np.where(r > 0, ur * x / np.where(r > 0, r, 1.0), 0.0)
        uy = np.where(r > 0, ur * y / np.where(r > 0, r, 1.0), 0.0)
    z = np.zeros_like(x)
    return dict(x=x, y=y, e=e, ux=ux, uy=uy, rhoB=z, rhoS=z.copy(), rhoQ=z.copy())


def lumpy(seed=0xCCA4E001, spacing=0.02, half_width=5.0, n_blobs=50, sigma=0.4,
          center_sd=3.0, amp_median_gev=20.0, amp_logsd=0.5, charge_frac=0.05, e_cutoff_gev=0.0):
    """Lumpy event: e = sum of n_blobs 2D Gaussians (width sigma) with centres
    ~ N(0, center_sd) per axis and amplitudes ~ LogNormal(ln amp_median, amp_logsd)
    GeV/fm^3; rho_X = sum_k f_{X,k} blob_k with f ~ U(-charge_frac, charge_frac)
    (in fm^-3 per GeV/fm^3).  Draw order per blob: c_x, c_y, A, f_B, f_S, f_Q.
    Lattice points with e <= e_cutoff_gev (GeV/fm^3) are dropped, as the
    reference's initial-condition readers do (src/BSQHydro.cpp:129, 402-405;
    default 0.15 GeV/fm^3 in include/defaults.h:31)."""
    rng = np.random.Generator(np.random.MT19937(seed))
    x, y = lattice(spacing, half_width)
    e = np.zeros_like(x)
    rb, rs, rq = np.zeros_like(x), np.zeros_like(x), np.zeros_like(x)
    for _ in range(n_blobs):
        cx = rng.normal(0.0, center_sd)
        cy = rng.normal(0.0, center_sd)
        A = float(np.exp(rng.normal(np.log(amp_median_gev), amp_logsd)))
        fB, fS, fQ = rng.uniform(-charge_frac, charge_frac, size=3)
        blob = A * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * sigma * sigma))
        e += blob
        rb += fB * blob
        rs += fS * blob
        rq += fQ * blob
    keep = e > e_cutoff_gev
    x, y, e, rb, rs, rq = x[keep], y[keep], e[keep], rb[keep], rs[keep], rq[keep]
    z = np.zeros_like(x)
    return dict(x=x, y=y, e=e / HBARC_GEVFM, ux=z, uy=z.copy(), rhoB=rb, rhoS=rs, rhoQ=rq)
