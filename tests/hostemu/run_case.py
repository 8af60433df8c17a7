"""Run one event through the host emulation build of the library and
compare it with the oracle (test infrastructure).

Cases:
  gubser  ideal Gubser flow, conformal EoS (51 x 51 particles by default)
  table   lumpy viscous BSQ event on the 64x32^3 tanh-conformal table
  table_fo  the same event with the reference's default energy cutoff,
          particlization on, run step by step
  milne   3+1D lumpy viscous BSQ event in Milne coordinates on the table EoS
          (isotropic 3D kernel); milne_fact the same with the factorized kernel
          W_2D(r_perp, hT) W_1D(|d eta|, hEta), hEta = 0.2 fm

Usage: run_case.py CASE SPACING STEPS [--no-oracle] [--tol X]
STEPS = -1 runs until continue_evolution ends the run (every particle
frozen out), as SPHWorkstation::continue_evolution does, and compares the
step count and the final state with the oracle's run.
Must run with the emulation library and AddressSanitizer loaded, e.g.
through tests/hostemu/emu.sh.  Every C-ABI call prints a line first, so a
report names the call that made it.
"""
from __future__ import annotations

import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

import ccake_amd  # noqa: E402
from ccake_amd import config, eos_tables, ic  # noqa: E402
from parity_metrics import DERIVS, DERIVS3, STATE, STATE3, field_err, pp_tol, whole_field  # noqa: E402

T_FO = 150.0 / config.HBARC_MEVFM


def say(*a):
    print(f"[{time.strftime('%H:%M:%S')}]", *a, flush=True)


def main():
    case, spacing, steps = sys.argv[1], float(sys.argv[2]), int(sys.argv[3])
    use_oracle = "--no-oracle" not in sys.argv
    tol = float(sys.argv[sys.argv.index("--tol") + 1]) if "--tol" in sys.argv else 1e-12
    assert ccake_amd.LIB_PATH.endswith("libccake_emu.so"), ccake_amd.LIB_PATH
    dev = ccake_amd.Device(0)
    table = None
    if case == "gubser":
        cfg = config.gubser_ideal_params()
        d = ic.gubser_ideal(spacing=spacing)
        geos = ccake_amd.Eos(dev, 0, T_FO)
    else:
        table = eos_tables.tanh_conformal_table()
        tab, dims, mins, maxs = table
        geos = ccake_amd.Eos(dev, 1, T_FO, dims, mins, maxs, tab)
        cfg = config.lumpy_params()
        if case.startswith("milne"):
            cfg = dict(cfg, dim=3, hEta=0.2, factorized_kernel=1 if case == "milne_fact" else 0, particlization=0)
            d = ic.milne_lumpy(spacing=spacing, half_width=1.0, eta_spacing=spacing, eta_max=1.0, center_sd=0.5)
        elif case == "table_fo":
            cfg = dict(cfg, particlization=1, e_freeze=geos.efreeze(T_FO))
            d = ic.lumpy(spacing=spacing, e_cutoff_gev=0.15)
        else:
            d = ic.lumpy(spacing=spacing)
    n = len(d["x"])
    say(f"case {case}: {n} particles, {steps} steps")
    ev = ccake_amd.Event(dev, geos, cfg, n)
    ccake_amd.load_event(ev, d)
    say("initialize")
    ev.initialize()
    say("initialize done")
    nsteps = 0
    if steps >= 0:
        for k in range(steps):
            ev.advance_timestep(cfg["dt"], 2)
            say(f"step {k + 1} done, t = {ev.t:.4f}")
        nsteps = steps
    else:
        while ev.continue_evolution(max_tau=30.0):
            ev.advance_timestep(cfg["dt"], 2)
            nsteps += 1
            if nsteps % 25 == 0:
                say(f"step {nsteps}, t = {ev.t:.4f}, frozen {ev.reduce('CCAKE_R_FROZEN'):.0f} of {n}")
        say(f"run ended after {nsteps} steps at t = {ev.t:.4f}")
    if use_oracle:
        import oracle
        oeos = oracle.Eos(0, T_FO) if table is None else oracle.Eos(1, T_FO, *table[1:], table[0])
        oev = oracle.Event(cfg, oeos, n)
        for k, f in [("x", "X"), ("y", "Y"), ("eta", "ETA"), ("e", "IN_E"), ("ux", "UX"), ("uy", "UY"),
                     ("ueta", "UETA"), ("rhoB", "IN_B"), ("rhoS", "IN_SS"), ("rhoQ", "IN_Q")]:
            if k in d:
                oev.set("ORC_F_" + f, d[k])
        assert oev.initialize() == 0
        if steps >= 0:
            for _ in range(steps):
                assert oev.advance_timestep(cfg["dt"], 2) == 0
        else:
            osteps = 0
            while not np.all(oev.ifield("ORC_I_FREEZE") >= 4) and oev.t <= 30.0:
                assert oev.advance_timestep(cfg["dt"], 2) == 0
                osteps += 1
            say(f"oracle run ended after {osteps} steps")
            if osteps != nsteps:
                raise SystemExit(f"step counts differ: {nsteps} against {osteps}")
        errs, whole = {}, {}
        st, dv = (STATE3, DERIVS3) if case.startswith("milne") else (STATE, DERIVS)
        for lab in st + (dv if steps != 0 else []):
            fid = ccake_amd.FIELDS["CCAKE_F_" + lab.split("+")[0]] + (int(lab.split("+")[1]) if "+" in lab else 0)
            a, b = ev.download(fid), oev.field(fid)
            errs[lab] = field_err(lab, a, b, tol)
            whole[lab] = whole_field(a, b)
        say("worst whole-field difference: %s %.2e" % max(whole.items(), key=lambda kv: kv[1]))
        worst = sorted(errs.items(), key=lambda kv: -kv[1])[:6]
        say("worst per-particle field differences against the oracle: "
            + ", ".join(f"{k} {v:.2e}" for k, v in worst))
        for name in ("BTRACK", "FREEZE", "EOS_ID"):
            if not np.array_equal(ev.download_int("CCAKE_I_" + name), oev.ifield("ORC_I_" + name)):
                raise SystemExit(f"{name} differs from the oracle")
        say("btrack, freeze and EoS member equal to the oracle's")
        bad = {k: v for k, v in errs.items() if not v <= pp_tol(k, tol)}
        bad.update({k: v for k, v in whole.items() if not v <= tol})
        if bad:
            raise SystemExit(f"fields beyond their tolerance ({tol} whole-field): {bad}")
    ev.close(); geos.close(); dev.close()
    say("OK")


if __name__ == "__main__":
    main()
