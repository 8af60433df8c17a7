"""Error of the mixed-precision pair sums (pair_precision 1) against the FP64
path on the same event, per field.

The measure is per particle: |x - x_ref| / |x_ref| for fields that never
change sign, and |x - x_ref| / max(|x_ref|, 1e-3 max|x_ref|) for signed
fields (velocities, charge densities, shear and bulk pressure), whose
relative error at a zero crossing means nothing.  Each field reports the
largest, the 99th-percentile and the median per-particle error, and the
whole-field error max|x - x_ref| / max|x_ref|.  The four-velocity also
reports |u^mu - u_ref^mu| / u_ref^0 per particle ("u_mu").
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


def whole_field(x: np.ndarray, ref: np.ndarray) -> float:
    """max |x - ref| / max |ref| over the field."""
    if ref.size == 0:
        return 0.0
    return float(np.max(np.abs(np.asarray(x) - np.asarray(ref))) / (np.max(np.abs(ref)) or 1.0))


def four_velocity_errors(g, ux, uy, g_ref, ux_ref, uy_ref) -> np.ndarray:
    """|u^mu - u_ref^mu| / u_ref^0 per particle (Euclidean over the three
    components, against the time component, which is at least 1)."""
    d = np.sqrt((g - g_ref) ** 2 + (ux - ux_ref) ** 2 + (uy - uy_ref) ** 2)
    return d / g_ref


def summarize(r: np.ndarray) -> dict:
    if r.size == 0:
        return {"max": 0.0, "p99": 0.0, "median": 0.0}
    return {"max": float(np.max(r)), "p99": float(np.percentile(r, 99)), "median": float(np.median(r))}


def surface_errors(rows_ref: np.ndarray, rows: np.ndarray) -> dict:
    """Per-field errors of the hypersurface rows of a mixed-precision run
    against the FP64 run's rows of the same event.  Rows are matched by
    particle id; the fields are compared where both runs froze the particle
    out in the same step and interpolated from the same snapshot (same row
    time).  A particle that crosses e_FO a step earlier in one run, or whose
    two candidate snapshots are nearly equally close to e_FO, carries the
    state of another time in one run, which is not a precision error of the
    fields; the counts of both kinds are reported."""
    ia = {int(r[0]): k for k, r in enumerate(rows_ref)}
    ib = {int(r[0]): k for k, r in enumerate(rows)}
    common = sorted(set(ia) & set(ib))
    step_same = [i for i in common if rows_ref[ia[i], 1] == rows[ib[i], 1]]
    # the row's time tells which snapshot the interpolation took (the one
    # whose energy is closer to e_FO): where the two runs took different
    # snapshots the row describes another time, a discrete choice rather
    # than a precision error of the fields
    same = [i for i in step_same if rows_ref[ia[i], 15] == rows[ib[i], 15]]
    a = rows_ref[[ia[i] for i in same]] if same else np.zeros((0, rows_ref.shape[1]))
    b = rows[[ib[i] for i in same]] if same else np.zeros((0, rows_ref.shape[1]))
    out = {"rows_ref": int(len(rows_ref)), "rows": int(len(rows)), "matched": len(common),
           "same_step": len(step_same), "same_snapshot": len(same), "fields": {}}
    for name, (c, signed) in SURFACE_FIELDS.items():
        out["fields"][name] = summarize(rel_errors(b[:, c], a[:, c], signed))
        out["fields"][name]["whole"] = whole_field(b[:, c], a[:, c])
    out["fields"]["u_mu"] = summarize(four_velocity_errors(b[:, 5], b[:, 6], b[:, 7], a[:, 5], a[:, 6], a[:, 7]))
    return out


def state_errors(get_ref, get, fields: dict) -> dict:
    """Per-field errors of two events' states; fields maps a name to
    (field id, signed); get_ref / get download a field id."""
    return {name: summarize(rel_errors(get(fid), get_ref(fid), signed)) for name, (fid, signed) in fields.items()}
