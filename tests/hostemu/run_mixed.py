"""Mixed-precision pair sums (pair_precision 1) against the FP64 path on the
host emulation build (test infrastructure): a lumpy viscous BSQ event with
charges on the conformal EoS, initialize, one derivative evaluation and
STEPS steps; prints the whole-field relative error of each compared field
and "OK" when every one is within TOL and btrack is equal.

Usage (through tests/hostemu/emu.sh): run_mixed.py SPACING STEPS TOL
"""
from __future__ import annotations

import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import ccake_amd  # noqa: E402
from ccake_amd import config, ic  # noqa: E402

FIELDS = ["SIGMA_LAB", "SIGMA", "S", "E", "T", "RHOB", "UX", "UY", "BULK", "GRADP0", "GRADE0", "GRADV0",
          "DUDT0", "DEXT_S"]


def main():
    spacing, steps, tol = float(sys.argv[1]), int(sys.argv[2]), float(sys.argv[3])
    assert ccake_amd.LIB_PATH.endswith("libccake_emu.so"), ccake_amd.LIB_PATH
    dev = ccake_amd.Device(0)
    eos = ccake_amd.Eos(dev, 0, 150.0 / config.HBARC_MEVFM)
    d = ic.lumpy(spacing=spacing, e_cutoff_gev=0.15)
    cfg = config.lumpy_params()
    evs = []
    for pp in (0, 1):
        ev = ccake_amd.Event(dev, eos, dict(cfg, pair_precision=pp), len(d["x"]))
        ccake_amd.load_event(ev, d)
        ev.initialize()
        ev.get_time_derivatives()
        for _ in range(steps):
            ev.advance_timestep(cfg["dt"], 2)
        evs.append(ev)
    worst = 0.0
    for lab in FIELDS:
        f = ccake_amd.FIELDS["CCAKE_F_" + lab]
        a, b = evs[0].download(f), evs[1].download(f)
        e = float(np.max(np.abs(b - a)) / (np.max(np.abs(a)) or 1.0))
        print(f"{lab}: {e:.2e}", flush=True)
        worst = max(worst, e)
    same_bt = np.array_equal(evs[0].download_int("CCAKE_I_BTRACK"), evs[1].download_int("CCAKE_I_BTRACK"))
    print("particles", len(d["x"]), "worst", worst, "btrack equal", same_bt, flush=True)
    if worst <= tol and worst > 0.0 and same_bt:
        print("OK", flush=True)


if __name__ == "__main__":
    main()
