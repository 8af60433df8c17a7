"""Synthetic initial conditions for the benchmark workloads.

Energy densities are returned in fm^-4 (the reference's internal unit:
src/BSQHydro.cpp:272-275 divides the file's GeV/fm^3 by hbar c), charges
in fm^-3.  Formulas are restated from their physics definitions; the ideal
Gubser profile is the closed form of utilities/ideal_gubser_ic_generator.py
(eps_a, kappa, velocity_x/y) and misc/Gubser_checks/plot_gubser_profiles.py.
"""
from __future__ import annotations

import re

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


def lattice3(spacing, half_width, eta_spacing, eta_max):
    """(x, y, eta_s) lattice: the square lattice of `lattice` at every eta
    node of [-eta_max, eta_max] (x fastest, eta slowest)."""
    x2, y2 = lattice(spacing, half_width)
    m = int(round(2 * eta_max / eta_spacing)) + 1
    ge = np.round(-eta_max + eta_spacing * np.arange(m), 12) + 0.0
    x = np.tile(x2, m)
    y = np.tile(y2, m)
    eta = np.repeat(ge, len(x2))
    return x, y, eta


def plateau(eta, eta_flat=2.0, eta_sigma=0.5):
    """Longitudinal Gaussian-plateau profile: 1 for |eta| <= eta_flat/2,
    exp(-(|eta| - eta_flat/2)^2 / (2 eta_sigma^2)) beyond."""
    d = np.maximum(np.abs(eta) - 0.5 * eta_flat, 0.0)
    return np.exp(-d * d / (2.0 * eta_sigma * eta_sigma))


def milne_lumpy(seed=0xCCA4E001, spacing=0.1, half_width=3.0, eta_spacing=0.1, eta_max=4.0, eta_flat=2.0,
                eta_sigma=0.5, n_blobs=20, sigma=0.4, center_sd=1.0, amp_median_gev=20.0, amp_logsd=0.5,
                charge_frac=0.05, e_cutoff_gev=0.0):
    """3+1D lumpy event in Milne coordinates (tau, x, y, eta_s): the
    transverse lumpy profile of `lumpy` (same draws, same units) times the
    longitudinal Gaussian plateau of `plateau`, on the (x, y, eta_s)
    lattice of `lattice3`; charges the same blobs scaled; u = 0.  Points at
    or below e_cutoff_gev (GeV/fm^3) are dropped."""
    rng = np.random.Generator(np.random.MT19937(seed))
    x, y, eta = lattice3(spacing, half_width, eta_spacing, eta_max)
    et = np.zeros_like(x)
    rb, rs, rq = np.zeros_like(x), np.zeros_like(x), np.zeros_like(x)
    for _ in range(n_blobs):
        cx = rng.normal(0.0, center_sd)
        cy = rng.normal(0.0, center_sd)
        A = float(np.exp(rng.normal(np.log(amp_median_gev), amp_logsd)))
        fB, fS, fQ = rng.uniform(-charge_frac, charge_frac, size=3)
        blob = A * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * sigma * sigma))
        et += blob
        rb += fB * blob
        rs += fS * blob
        rq += fQ * blob
    h = plateau(eta, eta_flat, eta_sigma)
    e, rb, rs, rq = et * h, rb * h, rs * h, rq * h
    keep = e > e_cutoff_gev
    x, y, eta, e, rb, rs, rq = x[keep], y[keep], eta[keep], e[keep], rb[keep], rs[keep], rq[keep]
    z = np.zeros_like(x)
    return dict(x=x, y=y, eta=eta, e=e / HBARC_GEVFM, ux=z, uy=z.copy(), ueta=z.copy(), rhoB=rb, rhoS=rs, rhoQ=rq)


# ---------------------------------------------------------------- file formats
# Readers of the reference's initial-condition files, the Python twins of
# read_ccake / read_ICCING in ccake_amd/host/ccake_io.hpp: the same fields,
# units, cutoff and charge switches, and the same treatment of short or
# malformed lines, because numbers are taken the way std::istringstream >>
# takes them (libstdc++ num_get: optional sign, digits, point, digits,
# exponent; an incomplete or empty number reads as 0 and fails the line; a
# read at the end of the line fails and leaves the variable as it was).

_WS = " \t\n\v\f\r"
_DIGITS = "0123456789"
_FULL = re.compile(r"[+-]?(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?\Z")


class _LineStream:
    """std::istringstream over one line, read with operator>>(double&)."""

    def __init__(self, text, pos=0, eof=False):
        self.s, self.pos, self.failed = text, pos, eof

    def read(self, current):
        if self.failed:
            return current
        s, p, n = self.s, self.pos, len(self.s)
        while p < n and s[p] in _WS:
            p += 1
        if p >= n:                          # sentry fails: value unchanged
            self.failed, self.pos = True, p
            return current
        q = p
        if q < n and s[q] in "+-":
            q += 1
        while q < n and s[q] in _DIGITS:
            q += 1
        if q < n and s[q] == ".":
            q += 1
            while q < n and s[q] in _DIGITS:
                q += 1
        if q < n and s[q] in "eE":
            q += 1
            if q < n and s[q] in "+-":
                q += 1
            while q < n and s[q] in _DIGITS:
                q += 1
        tok = s[p:q]
        self.pos = q
        if _FULL.match(tok):
            return float(tok)
        self.failed = True                  # nothing or an incomplete number: 0
        return 0.0


def _ignore_to_hash(line):
    """istream::ignore(256, '#') on the header line: the stream after it."""
    k = line.find("#", 0, 256)
    if k >= 0:
        return _LineStream(line, k + 1)
    if len(line) > 256:
        return _LineStream(line, 256)
    return _LineStream(line, len(line), eof=True)


def read_ccake(path, e_cutoff_gev=0.15, baryon=True, strange=True, electric=True, initial_diffusion=False,
               input_as_entropy=False):
    """BSQHydro::read_ccake (src/BSQHydro.cpp:160-425) for 2+1D: header
    "#<ignored> dx dy deta <ignored> xmin ymin etamin", then per line x y eta
    e[GeV/fm^3] rhoB rhoS rhoQ ux uy ueta Pi pi^xx pi^xy pi^xeta pi^yy
    pi^yeta pi^etaeta (GeV/fm^3 for e, Pi and pi), and with
    initial_diffusion 12 more columns qB0..3 qS0..3 qQ0..3.  Blank and '#'
    lines after the header are skipped; points with e <= e_cutoff_gev are
    dropped; disabled charges read as 0.  Returns the load_event dict (e, Pi
    and pi in fm^-4; shv as 16 rows, pi^{mu nu} row-major) with 'diffusion'
    (n x 3 x 4) and 'grid' (dx, dy, deta, xmin, ymin, etamin)."""
    if input_as_entropy:
        raise NotImplementedError("input_as_entropy initial conditions are not supported on this path")
    hc = HBARC_GEVFM
    names = ["x", "y", "eta", "col4", "rB", "rS", "rQ", "ux", "uy", "ueta", "bulk",
             "pxx", "pxy", "pxe", "pyy", "pye", "pee"]
    v = dict.fromkeys(names, 0.0)
    q = [0.0] * 12
    rows, grid, headers = [], (0.0,) * 6, []
    with open(path) as f:
        lines = f.read().split("\n")
    if lines and lines[-1] == "":
        lines.pop()
    for k, line in enumerate(lines):
        if k == 0:
            headers.append(line)
            st = _ignore_to_hash(line)
            g = [0.0] * 8
            for j in range(8):
                g[j] = st.read(g[j])
            grid = (g[1], g[2], g[3], g[5], g[6], g[7])
            continue
        stripped = line.lstrip(" \t\r\n")
        if not stripped or stripped[0] == "#":
            continue
        st = _LineStream(line)
        for nme in names:
            v[nme] = st.read(v[nme])
        if initial_diffusion:
            for j in range(12):
                q[j] = st.read(q[j])
        e = v["col4"]
        v["col4"] = 0.0
        e /= hc
        for nme in ("pxx", "pxy", "pxe", "pyy", "pye", "pee", "bulk"):
            v[nme] /= hc
        if e > e_cutoff_gev / hc:
            diff = [qq if on else 0.0 for qq, on in zip(q, [baryon] * 4 + [strange] * 4 + [electric] * 4)] \
                if initial_diffusion else [0.0] * 12
            rows.append((v["x"], v["y"], e, v["rB"] if baryon else 0.0, v["rS"] if strange else 0.0,
                         v["rQ"] if electric else 0.0, v["ux"], v["uy"], v["bulk"],
                         v["pxx"], v["pxy"], v["pxe"], v["pyy"], v["pye"], v["pee"], diff))
    n = len(rows)
    out = {k: np.array([r[j] for r in rows], dtype=np.float64)
           for j, k in enumerate(["x", "y", "e", "rhoB", "rhoS", "rhoQ", "ux", "uy", "bulk"])}
    shv = np.zeros((16, n))
    for j, (a, b) in enumerate([(1, 1), (1, 2), (1, 3), (2, 2), (2, 3), (3, 3)]):
        col = np.array([r[9 + j] for r in rows], dtype=np.float64)
        shv[4 * a + b] = col
        shv[4 * b + a] = col
    out["shv"] = shv
    out["diffusion"] = np.array([r[15] for r in rows], dtype=np.float64).reshape(n, 3, 4)
    out["grid"] = grid
    out["headers"] = headers
    return out


def read_iccing(path, e_cutoff_gev=0.15, baryon=True, strange=True, electric=True):
    """BSQHydro<2, EoM_default>::read_ICCING (src/BSQHydro.cpp:83-139): header
    "<ignored> dx dy <ignored> xmin ymin", then x y e[GeV/fm^3] rhoB rhoS
    rhoQ per line, fluid at rest; every line after the header is a point."""
    hc = HBARC_GEVFM
    x = y = e = rB = rS = rQ = 0.0
    rows, grid, headers = [], (0.0,) * 6, []
    with open(path) as f:
        lines = f.read().split("\n")
    if lines and lines[-1] == "":
        lines.pop()
    for k, line in enumerate(lines):
        st = _LineStream(line)
        if k == 0:
            headers.append(line)
            g = [0.0] * 6
            for j in range(6):
                g[j] = st.read(g[j])
            grid = (g[1], g[2], 0.0, g[4], g[5], 0.0)
            continue
        x, y, e, rB, rS, rQ = (st.read(c) for c in (x, y, e, rB, rS, rQ))
        e /= hc
        if e > e_cutoff_gev / hc:
            rows.append((x, y, e, rB if baryon else 0.0, rS if strange else 0.0, rQ if electric else 0.0))
    n = len(rows)
    out = {k: np.array([r[j] for r in rows], dtype=np.float64)
           for j, k in enumerate(["x", "y", "e", "rhoB", "rhoS", "rhoQ"])}
    out["ux"], out["uy"], out["bulk"] = np.zeros(n), np.zeros(n), np.zeros(n)
    out["shv"] = np.zeros((16, n))
    out["diffusion"] = np.zeros((n, 3, 4))
    out["grid"] = grid
    out["headers"] = headers
    return out


def write_ccake(path, d, grid=(0.0, 0.0, 0.0, 0.0, 0.0, 0.0)):
    """Writes an initial condition in the ccake format read above (e, Pi and
    pi converted back to GeV/fm^3; 17 significant digits)."""
    hc = HBARC_GEVFM
    n = len(d["x"])
    shv = d.get("shv", np.zeros((16, n)))
    bulk = d.get("bulk", np.zeros(n))
    dx, dy, deta, xmin, ymin, etamin = grid
    with open(path, "w") as f:
        f.write(f"#0 {dx!r} {dy!r} {deta!r} 0 {xmin!r} {ymin!r} {etamin!r}\n")
        for i in range(n):
            vals = [d["x"][i], d["y"][i], 0.0, d["e"][i] * hc, d["rhoB"][i], d["rhoS"][i], d["rhoQ"][i],
                    d["ux"][i], d["uy"][i], 0.0, bulk[i] * hc, shv[5][i] * hc, shv[6][i] * hc, shv[7][i] * hc,
                    shv[10][i] * hc, shv[11][i] * hc, shv[15][i] * hc]
            f.write(" ".join("%.17g" % float(v) for v in vals) + "\n")
