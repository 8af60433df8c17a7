"""Error of the mixed-precision pair sums (pair_precision 1) against the FP64
path on the same event, per field.

The measure is per particle: |x - x_ref| / |x_ref| for fields that never
change sign, and |x - x_ref| / max(|x_ref|, 1e-3 max|x_ref|) for signed
fields (velocities, charge densities, shear and bulk pressure), whose
relative error at a zero crossing means nothing.  Each field reports the
largest, the 99th-percentile and the median per-particle error.
"""
from __future__ import annotations

import numpy as np

SIGNED_FLOOR = 1e-3

# freeze_out.dat columns of a hypersurface row (ccake_freeze_out_surface:
# particle id, step, then the 40 columns of Output<2>::print_freeze_out)
SURFACE_FIELDS = {
    "T": (20, False), "e": (19, False), "s": (18, False), "gamma": (5, False),
    "u_x": (6, True), "u_y": (7, True),
    "rho_B": (26, True), "rho_S": (27, True), "rho_Q": (28, True),
    "pi_00": (10, True), "pi_11": (11, True), "pi_22": (12, True), "pi_33": (13, True), "pi_12": (14, True),
    "Pi": (9, True), "tau": (15, False), "x": (16, True), "y": (17, True),
}


def rel_errors(x: np.ndarray, ref: np.ndarray, signed: bool) -> np.ndarray:
    d = np.abs(np.asarray(x, dtype=np.float64) - np.asarray(ref, dtype=np.float64))
    if d.size == 0:
        return d
    scale = np.abs(ref)
    if signed:
        scale = np.maximum(scale, SIGNED_FLOOR * float(np.max(np.abs(ref))))
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(d == 0, 0.0, d / scale)


def summarize(r: np.ndarray) -> dict:
    if r.size == 0:
        return {"max": 0.0, "p99": 0.0, "median": 0.0}
    return {"max": float(np.max(r)), "p99": float(np.percentile(r, 99)), "median": float(np.median(r))}


def surface_errors(rows_ref: np.ndarray, rows: np.ndarray) -> dict:
    """Per-field errors of the hypersurface rows of a mixed-precision run
    against the FP64 run's rows of the same event.  Rows are matched by
    particle id; the fields are compared where both runs froze the particle
    out in the same step (a particle that crosses e_FO a step earlier in one
    run carries the state of another time, which is not a precision error
    of the fields; their number is reported)."""
    ia = {int(r[0]): k for k, r in enumerate(rows_ref)}
    ib = {int(r[0]): k for k, r in enumerate(rows)}
    common = sorted(set(ia) & set(ib))
    same = [i for i in common if rows_ref[ia[i], 1] == rows[ib[i], 1]]
    a = rows_ref[[ia[i] for i in same]] if same else np.zeros((0, rows_ref.shape[1]))
    b = rows[[ib[i] for i in same]] if same else np.zeros((0, rows_ref.shape[1]))
    out = {"rows_ref": int(len(rows_ref)), "rows": int(len(rows)), "matched": len(common),
           "same_step": len(same), "fields": {}}
    for name, (c, signed) in SURFACE_FIELDS.items():
        out["fields"][name] = summarize(rel_errors(b[:, c], a[:, c], signed))
    return out


def state_errors(get_ref, get, fields: dict) -> dict:
    """Per-field errors of two events' states; fields maps a name to
    (field id, signed); get_ref / get download a field id."""
    return {name: summarize(rel_errors(get(fid), get_ref(fid), signed)) for name, (fid, signed) in fields.items()}
