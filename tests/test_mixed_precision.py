"""Mixed-precision pair sums (ccake_params.pair_precision = 1): FP32 pair
terms with compensated FP32 accumulation, FP64 everywhere else.  The FP64
GPU path on the same event is the reference here (it matches the oracle to
1e-12, tests/test_gpu_parity.py).  Tolerances are written in each test.

What must not change: the pair set (the predicate runs in FP64 on FP64
positions), so btrack and the pair count equal the FP64 path's exactly.
"""
import numpy as np
import pytest

import ccake_amd
from ccake_amd import config, eos_tables, ic, precision

T_FO = 150.0 / config.HBARC_MEVFM
F = ccake_amd.FIELDS


def fid(label):
    if "+" in label:
        base, off = label.split("+")
        return F["CCAKE_F_" + base] + int(off)
    return F["CCAKE_F_" + label]


@pytest.fixture(scope="module")
def dev():
    d = ccake_amd.Device(0)
    yield d
    d.close()


@pytest.fixture(scope="module")
def table_eos(dev):
    tab, dims, mins, maxs = eos_tables.tanh_conformal_table()
    e = ccake_amd.Eos(dev, config.EOS_TABLE, T_FO, dims, mins, maxs, tab)
    yield e
    e.close()


def pair_of_events(dev, eos, cfg, d):
    evs = []
    for pp in (0, 1):
        ev = ccake_amd.Event(dev, eos, dict(cfg, pair_precision=pp), len(d["x"]))
        ccake_amd.load_event(ev, d)
        ev.initialize()
        evs.append(ev)
    return evs


def whole(a, b):
    return float(np.max(np.abs(a - b)) / (np.max(np.abs(b)) or 1.0))


# whole-field bounds after one derivative evaluation (FP32 rounding of the
# terms is 6e-8; a term's factors and its spline add a few roundings, the
# gradient sums cancel by up to the ratio of their terms to their result)
SMOOTH_FIELDS = ["SIGMA_LAB", "SIGMA", "S", "RHOB", "RHOS", "RHOQ", "E", "T", "P"]
GRAD_FIELDS = ["GRADP0", "GRADP0+1", "GRADE0", "GRADE0+1", "GRADBULK0", "GRADV0", "GRADV0+1", "GRADV0+3",
               "GRADV0+4", "DIVSHEAR0", "DIVSHEAR0+1", "GRADSHEAR0", "GRADSHEAR0+1"]
TOL_SMOOTH = 1e-6
TOL_GRAD = 2e-5


@pytest.mark.gpu
def test_mixed_pair_sums_against_fp64(dev, table_eos):
    """Lumpy viscous BSQ event on the table EoS (0.1 fm lattice with the
    reference's energy cutoff, 8,077 particles): after initialization and
    one derivative evaluation the smoothed fields are within 1e-6 and the
    gradient sums within 2e-5 (whole-field relative) of the FP64 path;
    btrack and the stage's pair count are equal."""
    d = ic.lumpy(spacing=0.1, e_cutoff_gev=0.15)
    a, b = pair_of_events(dev, table_eos, config.lumpy_params(), d)
    for ev in (a, b):
        ev.pair_total(reset=True)
        ev.get_time_derivatives()
    assert a.pair_total() == b.pair_total() > 0
    assert np.array_equal(a.download_int("CCAKE_I_BTRACK"), b.download_int("CCAKE_I_BTRACK"))
    errs = {lab: whole(b.download(fid(lab)), a.download(fid(lab))) for lab in SMOOTH_FIELDS + GRAD_FIELDS}
    print("whole-field errors after one derivative evaluation:",
          " ".join(f"{k}={v:.1e}" for k, v in errs.items()), flush=True)
    for lab in SMOOTH_FIELDS:
        assert errs[lab] <= TOL_SMOOTH, (lab, errs[lab])
    for lab in GRAD_FIELDS:
        assert errs[lab] <= TOL_GRAD, (lab, errs[lab])
    assert max(errs.values()) > 0.0        # the FP32 kernels ran
    a.close(); b.close()


@pytest.mark.gpu
def test_mixed_precision_freeze_out_surface(dev, table_eos):
    """The same event run to complete freeze-out in both precisions
    (SPHWorkstation::continue_evolution ends each run): the runs take the
    same number of steps (within 1) and nearly every particle freezes out
    in the same step from the same snapshot (the row times agree).  On
    those rows: T within 1e-6 per particle; e and s
    within 1e-5; u^mu within 1e-5 of u^0 per particle; the charge densities
    within 1e-5 of their largest magnitude (they pass through zero); every
    column within 2e-3 per particle with signed columns measured against
    1e-3 of their largest magnitude, that is within 2e-6 of the largest
    magnitude where a signed field crosses zero.  (Measured on this event
    with the FP32 blocks folded every 32 candidates, twice the default
    interval: T 1.3e-7, e and s 1e-6
    at most, medians 1e-8 to 3e-7; the largest errors sit where a signed
    field is near zero (rho_Q 9.6e-4 of the floor, 9.6e-7 of its largest
    magnitude), and the flow's sensitivity amplifies the FP32 terms' 1e-7
    differences over the 337 steps.)"""
    d = ic.lumpy(spacing=0.1, e_cutoff_gev=0.15)
    cfg = dict(config.lumpy_params(), particlization=1, e_freeze=table_eos.efreeze(T_FO))
    evs = pair_of_events(dev, table_eos, cfg, d)
    steps = []
    for ev in evs:
        k = 0
        while ev.continue_evolution(max_tau=30.0):
            ev.advance_timestep(cfg["dt"], 2)
            k += 1
        steps.append(k)
    assert abs(steps[0] - steps[1]) <= 1, steps
    rows = [ev.freeze_out_surface() for ev in evs]
    r = precision.surface_errors(rows[0], rows[1])
    print(f"steps {steps}, rows {r['rows_ref']} / {r['rows']}, same step {r['same_step']}, "
          f"same snapshot {r['same_snapshot']}", flush=True)
    for k, v in r["fields"].items():
        print(f"  {k}: max {v['max']:.2e} p99 {v['p99']:.2e} median {v['median']:.2e} whole {v.get('whole', 0):.2e}",
              flush=True)
    assert r["same_step"] >= 0.99 * r["rows_ref"] and r["same_snapshot"] >= 0.99 * r["rows_ref"]
    f = r["fields"]
    assert f["T"]["max"] <= 1e-6, f["T"]
    assert f["e"]["max"] <= 1e-5 and f["s"]["max"] <= 1e-5, (f["e"], f["s"])
    assert f["u_mu"]["max"] <= 1e-5, f["u_mu"]
    for k in ("rho_B", "rho_S", "rho_Q"):
        assert f[k]["whole"] <= 1e-5, (k, f[k])
    for k, v in f.items():
        assert v["max"] <= 2e-3, (k, v)
    for ev in evs:
        ev.close()
