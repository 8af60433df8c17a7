"""A batch of events against the same events run alone (test
infrastructure): three lumpy viscous BSQ events of different sizes and
seeds on the conformal EoS with charges, initialize and STEPS steps; every
event's fields within TOL of its own run alone (whole-field relative) and
its btrack equal; prints "OK".  With --fo the events run to complete
freeze-out (particlization on): each event's freeze-out rows are compared,
and so is its final state, since an event whose continue flag has turned
false ends in the batch (later batch steps leave it unchanged) as the run
alone ends.

Usage (through tests/hostemu/emu.sh or on a GPU): run_batch.py SPACING STEPS TOL [--fo] [--table]
"""
from __future__ import annotations

import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import ccake_amd  # noqa: E402
from ccake_amd import config, eos_tables, ic  # noqa: E402

FIELDS = ["X", "Y", "UX", "UY", "SIGMA_LAB", "SIGMA", "S", "E", "T", "RHOB", "RHOS", "RHOQ", "BULK", "EXT_S",
          "GRADP0", "GRADE0", "GRADV0", "DUDT0"]


def main():
    spacing, steps, tol = float(sys.argv[1]), int(sys.argv[2]), float(sys.argv[3])
    fo, table = "--fo" in sys.argv, "--table" in sys.argv
    T_FO = 150.0 / config.HBARC_MEVFM
    dev = ccake_amd.Device(0)
    if table:
        tab, dims, mins, maxs = eos_tables.tanh_conformal_table()
        eos = ccake_amd.Eos(dev, config.EOS_TABLE, T_FO, dims, mins, maxs, tab)
    else:
        eos = ccake_amd.Eos(dev, config.EOS_CONFORMAL, T_FO)
    cfg = config.lumpy_params()
    if fo:
        cfg = dict(cfg, particlization=1, e_freeze=eos.efreeze(T_FO))
    ics = [ic.lumpy(seed=0xCCA4E001 + k, spacing=spacing, half_width=hw, e_cutoff_gev=0.15)
           for k, hw in enumerate((5.0, 3.0, 4.0))]
    counts = [len(d["x"]) for d in ics]
    print("events", counts, flush=True)
    b = ccake_amd.Batch(dev, eos, cfg, counts)
    ccake_amd.load_event(b, ccake_amd.concat_ics(ics))
    b.initialize()
    alone = []
    for d in ics:
        ev = ccake_amd.Event(dev, eos, cfg, len(d["x"]))
        ccake_amd.load_event(ev, d)
        ev.initialize()
        alone.append(ev)
    if fo:
        k = 0
        while b.continue_evolution(max_tau=30.0):
            keep = b.batch_continue_evolution(max_tau=30.0)
            for e, ev in enumerate(alone):
                if keep[e]:
                    ev.advance_timestep(cfg["dt"], 2)
            b.advance_timestep(cfg["dt"], 2)
            k += 1
        print("batch steps", k, flush=True)
    else:
        b.run_steps(cfg["dt"], steps)
        for ev in alone:
            ev.run_steps(cfg["dt"], steps)
    worst = 0.0
    ok = True
    for e, ev in enumerate(alone):
        sl = b.event_slice(e)
        if not fo:
            for lab in FIELDS:
                f = ccake_amd.FIELDS["CCAKE_F_" + lab]
                x, y = b.download(f)[sl], ev.download(f)
                err = float(np.max(np.abs(x - y)) / (np.max(np.abs(y)) or 1.0))
                worst = max(worst, err)
            same_bt = np.array_equal(b.download_int("CCAKE_I_BTRACK")[sl], ev.download_int("CCAKE_I_BTRACK"))
            ok = ok and same_bt
            print(f"event {e}: btrack equal {same_bt}", flush=True)
        else:
            for lab in FIELDS:
                f = ccake_amd.FIELDS["CCAKE_F_" + lab]
                x, y = b.download(f)[sl], ev.download(f)
                err = float(np.max(np.abs(x - y)) / (np.max(np.abs(y)) or 1.0))
                if err > tol:
                    print(f"event {e}: final {lab} differs by {err:.3g}", flush=True)
                worst = max(worst, err)
            rb = b.surface_by_event()[e]
            ra = ev.freeze_out_surface()
            same = rb.shape == ra.shape and np.array_equal(rb[:, :2], ra[:, :2])
            err = float(np.max(np.abs(rb[:, 2:] - ra[:, 2:]) / np.maximum(np.abs(ra[:, 2:]), 1e-300))) if same else 1.0
            err = min(err, float(np.max(np.abs(rb[:, 2:] - ra[:, 2:])) / np.max(np.abs(ra[:, 2:])))) if same else 1.0
            worst = max(worst, err)
            ok = ok and same
            print(f"event {e}: rows {len(rb)} / {len(ra)}, ids and steps equal {same}", flush=True)
    print("worst", worst, flush=True)
    frozen = b.batch_reduce("CCAKE_R_FROZEN")
    print("frozen per event", frozen.tolist(), flush=True)
    if ok and worst <= tol:
        print("OK", flush=True)


if __name__ == "__main__":
    main()
