"""Per-particle comparison of a field against the oracle.

err(a, b) = max_i |a_i - b_i| / max(|b_i|, floor), with floor = frac * max|b|.
Positive-definite fields (energy, entropy, temperature, densities ...) use
frac = 0: every particle is held to the tolerance relative to its own value,
however small, so an error at a near-vacuum particle is not hidden by the
peak.  Signed fields (velocities, gradients, shear components, chemical
potentials) pass through zero, where a relative error is undefined; their
floor is frac = 1e-3 of the field's largest magnitude, and their
per-particle tolerance is TOL_SIGNED_PP (below).  The parity tests check
both this per-particle error and the whole-field error max|a-b| / max|b|.
"""
from __future__ import annotations

import numpy as np

SIGNED_FLOOR = 1e-3

# fields whose values are positive (or zero) on every particle
POSITIVE = {"GAMMA", "EXT_S", "SIGMA_LAB", "SIGMA", "SM_S", "T", "E", "S", "P", "CS2", "W", "DWDS",
            "IN_E", "IN_S", "MASS_S", "MASS_B", "MASS_SS", "MASS_Q", "ETA_PI", "TAU_PI", "ZETA_PI",
            "TAU_BULK", "DELTA_BB", "DELTA_PIPI", "LAMBDA_PIB", "TAU_PIPI", "PHI7"}


def floor_frac(label: str) -> float:
    return 0.0 if label.split("+")[0] in POSITIVE else SIGNED_FLOOR


def per_particle(a: np.ndarray, b: np.ndarray, frac: float) -> float:
    """max_i |a_i - b_i| / max(|b_i|, frac * max|b|); 0 for equal arrays."""
    if b.size == 0:
        return 0.0
    d = np.abs(a - b)
    if not np.any(d):
        return 0.0
    scale = np.maximum(np.abs(b), frac * float(np.max(np.abs(b))))
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(d == 0, 0.0, d / scale)
    return float(np.max(r))


# Per-particle tolerance of signed fields.  A signed field is a pair sum of
# terms of both signs (gradients, charge densities of blobs of both signs)
# or follows from one (chemical potentials, velocities): its rounding is
# relative to the sum of the terms' magnitudes, which can exceed the value
# by orders of magnitude.  Measured between two summation orders (the host
# build of the kernels against the oracle): up to 2e-11.
TOL_SIGNED_PP = 1e-10


def pp_tol(label: str, tol: float) -> float:
    """Per-particle tolerance of a field for a whole-field tolerance tol."""
    return tol if floor_frac(label) == 0.0 else max(tol, TOL_SIGNED_PP)


def whole_field(a: np.ndarray, b: np.ndarray) -> float:
    """max|a - b| / max|b| over a field."""
    s = float(np.max(np.abs(b))) if b.size else 0.0
    d = float(np.max(np.abs(a - b))) if b.size else 0.0
    return d if s == 0.0 else d / s


def field_err(label: str, a: np.ndarray, b: np.ndarray, tol: float) -> float:
    """Per-particle error of a field; non-finite oracle values must be
    matched exactly in position (NaN) and by the finite values to tol."""
    fin = np.isfinite(b)
    if not np.all(fin):
        if not np.array_equal(np.isfinite(a), fin):
            return float("inf")
        a, b = a[fin], b[fin]
    return per_particle(a, b, floor_frac(label))


# Field groups compared by the parity tests (labels: name or name+component).
STATE = ["X", "Y", "UX", "UY", "GAMMA", "VX", "VY", "EXT_S", "EXT_B", "EXT_SS", "EXT_Q", "EXT_BULK",
         "SIGMA_LAB", "SIGMA", "SM_S", "SM_B", "SM_SS", "SM_Q",
         "T", "MUB", "MUS", "MUQ", "E", "S", "RHOB", "RHOS", "RHOQ", "P", "CS2", "W",
         "DWDS", "DWDB", "DWDS_S", "DWDQ", "BULK"] + [f"SHV0+{c}" for c in range(16)]
TRANSPORT = ["ETA_PI", "TAU_PI", "ZETA_PI", "TAU_BULK", "DELTA_BB", "DELTA_PIPI", "LAMBDA_PIB",
             "LAMBDA_BPI", "PHI1", "PHI3", "PHI6", "PHI7", "TAU_PIPI"]
DERIVS = ([f"GRADP0+{c}" for c in range(2)] + [f"GRADE0+{c}" for c in range(2)]
          + [f"GRADBULK0+{c}" for c in range(2)] + [f"GRADSHEAR0+{c}" for c in range(2)]
          + [f"DIVSHEAR0+{c}" for c in range(2)] + ["GRADV0+0", "GRADV0+1", "GRADV0+3", "GRADV0+4"]
          + [f"DUDT0+{c}" for c in range(2)]
          + ["DEXT_S", "DEXT_B", "DEXT_SS", "DEXT_Q", "DEXT_BULK", "THETA", "SHV_NABLA_U", "DEZ"]
          + ["SHEAR_KNUDSEN", "BULK_KNUDSEN", "INV_REYNOLDS_SHEAR", "INV_REYNOLDS_BULK"]
          + [f"DEXT_SHV0+{c}" for c in range(5)])


# 3+1D (D = 3): the eta components, three-component gradients and the 3x3
# velocity gradient; the Knudsen and inverse-Reynolds diagnostics are 2+1D
# only.
STATE3 = STATE + ["ETA", "UETA", "VETA"]
DERIVS3 = ([f"GRADP0+{c}" for c in range(3)] + [f"GRADE0+{c}" for c in range(3)]
           + [f"GRADBULK0+{c}" for c in range(3)] + [f"GRADSHEAR0+{c}" for c in range(3)]
           + [f"DIVSHEAR0+{c}" for c in range(3)] + [f"GRADV0+{c}" for c in range(9)]
           + [f"DUDT0+{c}" for c in range(3)]
           + ["DEXT_S", "DEXT_B", "DEXT_SS", "DEXT_Q", "DEXT_BULK", "THETA", "SHV_NABLA_U", "DEZ"]
           + [f"DEXT_SHV0+{c}" for c in range(5)])
