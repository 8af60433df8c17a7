"""Synthetic 4D equation-of-state tables.

`tanh_conformal_table` fills a uniform (T, muB, muQ, muS) grid, T slowest,
with the 17 fields of the reference's table layout (p, s, B, S, Q, e, cs2,
chiBB, chiQQ, chiSS, chiBQ, chiBS, chiQS, chiTB, chiTQ, chiTS, chiTT;
src/eos_table.cpp:30-34) from the tanh-modulated conformal pressure
p = t(T) c T0^4 x^2, x = (T/T0)^2 + sum (mu/mu0)^2, t = (1 + tanh((T-Tc)/Ts))/2
(include/eos_tanh_conformal.h), in FP64 and fm units.  The second
derivatives are analytic; cs2 is the adiabatic dp/de along fixed n_X/s.
"""
from __future__ import annotations

import math

import numpy as np

from .config import HBARC_MEVFM

NC, NF = 3.0, 2.5
C_CONF = math.pi ** 2 * (2.0 * (NC * NC - 1.0) + (7.0 / 2.0) * NC * NF) / 90.0


def tanh_conformal_fields(T, muB, muQ, muS, c=C_CONF, T0=1.0, muB0=1.0, muQ0=1.0, muS0=1.0,
                          Tc=220.0 / HBARC_MEVFM, Ts=120.0 / HBARC_MEVFM):
    """The 17 forward fields at arrays of points (fm units)."""
    T, muB, muQ, muS = np.broadcast_arrays(*(np.asarray(a, dtype=np.float64) for a in (T, muB, muQ, muS)))
    cp = c * T0 ** 4
    x = (T / T0) ** 2 + (muB / muB0) ** 2 + (muS / muS0) ** 2 + (muQ / muQ0) ** 2
    pc = cp * x * x
    sc = 4 * cp * T * x / T0 ** 2
    Bc = 4 * cp * muB * x / muB0 ** 2
    Sc = 4 * cp * muS * x / muS0 ** 2
    Qc = 4 * cp * muQ * x / muQ0 ** 2
    arg = (T - Tc) / Ts
    th = np.tanh(arg)
    tf = 0.5 * (1 + th)
    tp = 0.5 / np.cosh(arg) ** 2 / Ts
    tpp = -2.0 * th * tp / Ts
    r = np.empty(T.shape + (17,))
    r[..., 0] = tf * pc
    r[..., 1] = tf * sc + tp * pc
    r[..., 2] = tf * Bc
    r[..., 3] = tf * Sc
    r[..., 4] = tf * Qc
    r[..., 5] = r[..., 1] * T + muB * r[..., 2] + muQ * r[..., 4] + muS * r[..., 3] - r[..., 0]
    r[..., 7] = tf * cp * (8 * muB ** 2 / muB0 ** 4 + 4 * x / muB0 ** 2)
    r[..., 8] = tf * cp * (8 * muQ ** 2 / muQ0 ** 4 + 4 * x / muQ0 ** 2)
    r[..., 9] = tf * cp * (8 * muS ** 2 / muS0 ** 4 + 4 * x / muS0 ** 2)
    r[..., 10] = tf * 8 * cp * muB * muQ / (muB0 ** 2 * muQ0 ** 2)
    r[..., 11] = tf * 8 * cp * muB * muS / (muB0 ** 2 * muS0 ** 2)
    r[..., 12] = tf * 8 * cp * muS * muQ / (muS0 ** 2 * muQ0 ** 2)
    r[..., 13] = tf * 8 * cp * T * muB / (T0 ** 2 * muB0 ** 2) + tp * Bc
    r[..., 14] = tf * 8 * cp * T * muQ / (T0 ** 2 * muQ0 ** 2) + tp * Qc
    r[..., 15] = tf * 8 * cp * T * muS / (T0 ** 2 * muS0 ** 2) + tp * Sc
    r[..., 16] = tf * cp * (8 * T ** 2 / T0 ** 4 + 4 * x / T0 ** 2) + 2 * tp * sc + tpp * pc
    r[..., 6] = adiabatic_cs2(T, muB, muQ, muS, r)
    return r


def adiabatic_cs2(T, muB, muQ, muS, r):
    """dp/de along d(n_X/s) = 0 from the Hessian of p over (T, muB, muQ, muS)."""
    H = np.stack([np.stack([r[..., 16], r[..., 13], r[..., 14], r[..., 15]], -1),
                  np.stack([r[..., 13], r[..., 7], r[..., 10], r[..., 11]], -1),
                  np.stack([r[..., 14], r[..., 10], r[..., 8], r[..., 12]], -1),
                  np.stack([r[..., 15], r[..., 11], r[..., 12], r[..., 9]], -1)], -2)
    s = r[..., 1]
    rhs = np.stack([np.ones_like(s), r[..., 2] / s, r[..., 4] / s, r[..., 3] / s], -1)
    d = np.linalg.solve(H, rhs[..., None])[..., 0]
    dp = r[..., 1] * d[..., 0] + r[..., 2] * d[..., 1] + r[..., 4] * d[..., 2] + r[..., 3] * d[..., 3]
    de = T + muB * rhs[..., 1] + muQ * rhs[..., 2] + muS * rhs[..., 3]
    return dp / de


def tanh_conformal_table(dims=(64, 32, 32, 32), T_range=(0.05, 0.8), mu_max=0.45):
    """Table over T in T_range and |muB|, |muQ|, |muS| <= mu_max (GeV).
    Returns (table[(nT*nB*nQ*nS)*17], dims, mins, maxs) in fm units."""
    g2f = 1000.0 / HBARC_MEVFM
    mins = [T_range[0] * g2f, -mu_max * g2f, -mu_max * g2f, -mu_max * g2f]
    maxs = [T_range[1] * g2f, mu_max * g2f, mu_max * g2f, mu_max * g2f]
    axes = [np.linspace(mins[a], maxs[a], dims[a]) for a in range(4)]
    tab = np.empty(tuple(dims) + (17,))
    B, Q, S = np.meshgrid(axes[1], axes[2], axes[3], indexing="ij")
    for i, T in enumerate(axes[0]):
        tab[i] = tanh_conformal_fields(np.full(B.shape, T), B, Q, S)
    return tab.reshape(-1), list(dims), mins, maxs


def preinverted_flat(fields_fm, s_axis, nB_axis):
    """A flat (s, nB) pre-inverted table (the input of CCAKE_EOS_PREINVERTED,
    the content of the reference's flat HDF5 file, src/eos_interpolator.cpp:
    291-527) from fields given in fm units on the (s, nB) grid.
    fields_fm: dict of (Ns, NB) arrays T, muB, muS, muQ [fm^-1], e, p [fm^-4],
    cs2, chiTT, chiTB, chiBB [fm^-2]; s_axis, nB_axis in fm^-3.
    Returns (dims, table): table is 12 x Ns x NB in MeV units (MeV^3, MeV,
    MeV^4, MeV^2) in the order s, nB, T, muB, muS, muQ, e, p, cs2, chiTT,
    chiTB, chiBB."""
    hc = HBARC_MEVFM
    Ns, NB = len(s_axis), len(nB_axis)
    S, B = np.meshgrid(np.asarray(s_axis, float), np.asarray(nB_axis, float), indexing="ij")
    f = {k: np.broadcast_to(np.asarray(v, float), (Ns, NB)) for k, v in fields_fm.items()}
    planes = [S * hc ** 3, B * hc ** 3, f["T"] * hc, f["muB"] * hc, f["muS"] * hc, f["muQ"] * hc,
              f["e"] * hc ** 4, f["p"] * hc ** 4, f["cs2"], f["chiTT"] * hc ** 2, f["chiTB"] * hc ** 2,
              f["chiBB"] * hc ** 2]
    return (Ns, NB, 1, 1), np.ascontiguousarray(np.stack(planes))


def preinverted_conformal(Ns=600, NB=4, s_range=(1e-3, 1e4), nB_range=(1e-4, 1e-1), c=C_CONF):
    """The conformal EoS at zero chemical potential (p = c T^4, s = 4 c T^3)
    as a flat pre-inverted table: T = (s/4c)^(1/3) on every nB column,
    chiTT = 12 c T^2, chiTB = 0, chiBB = 2 c T^2 (the mu^2 curvature of a
    free gas), cs2 = 1/3."""
    s = np.logspace(np.log10(s_range[0]), np.log10(s_range[1]), Ns)
    nB = np.logspace(np.log10(nB_range[0]), np.log10(nB_range[1]), NB)
    T = (s / (4.0 * c)) ** (1.0 / 3.0)
    T2 = np.repeat(T[:, None], NB, axis=1)
    z = np.zeros_like(T2)
    fields = dict(T=T2, muB=z, muS=z, muQ=z, e=3.0 * c * T2 ** 4, p=c * T2 ** 4, cs2=np.full_like(T2, 1.0 / 3.0),
                  chiTT=12.0 * c * T2 ** 2, chiTB=z, chiBB=2.0 * c * T2 ** 2)
    return preinverted_flat(fields, s, nB)


def load_flat_table_h5(path):
    """The reference's flat (s, nB) pre-inverted table file (HDF5: /dimensions
    = [Ns, NB], /data/{s, nB, T, muB, muS, muQ, e, p, cs2, chiTT, chiTB,
    chiBB}, MeV units; EoS_Interpolator::load_flat_table,
    src/eos_interpolator.cpp:291-527) read through libccake_hdf5.so:
    (table, (Ns, NB)) with the 12 arrays concatenated in that order, as
    ccake_amd.Eos takes them for a pre-inverted EoS."""
    import ctypes as C
    import os
    lib_path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libccake_hdf5.so")
    if not os.path.exists(lib_path):
        raise RuntimeError("libccake_hdf5.so is not built (no HDF5 library at build time)")
    L = C.CDLL(lib_path)
    f = L.ccake_h5_load_flat_table
    f.restype = C.c_int
    f.argtypes = [C.c_char_p, C.POINTER(C.c_double), C.c_size_t, C.POINTER(C.c_int), C.POINTER(C.c_int),
                  C.c_char_p, C.c_size_t]
    ns, nb = C.c_int(), C.c_int()
    err = C.create_string_buffer(512)
    st = f(path.encode(), None, 0, C.byref(ns), C.byref(nb), err, 512)
    if st == 1:
        raise ValueError(err.value.decode())
    out = np.zeros(12 * ns.value * nb.value)
    st = f(path.encode(), out.ctypes.data_as(C.POINTER(C.c_double)), out.size, C.byref(ns), C.byref(nb), err, 512)
    if st != 0:
        raise ValueError(err.value.decode() or f"status {st}")
    return out, (ns.value, nb.value)
