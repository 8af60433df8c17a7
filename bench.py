"""Throughput of the per-timestep SPH path on one MI355X.

Default workload: ideal 2+1D Gubser flow, FP64, 40,401 particles on the
0.05 fm lattice over [-5, 5]^2, conformal EoS, h = 0.3 fm, RK2 with
dt = 0.005 fm/c from tau0 = 1 fm/c.  `--workload lumpy40k` runs the viscous
BSQ lumpy event (50 Gaussian blobs, charges as scaled blobs, the
64x32x32x32 tanh-modulated conformal table EoS with its fallback chain,
eta/s = 0.08, Gaussian zeta/s, particlization at T = 150 MeV) on the
0.05 fm lattice (40,401 particles).  The 251,001-particle lumpy event is
not offered: its initialisation faulted the GPU and is held off it until
the cause is found.

A "step" is one RK2 time step of the whole event (two derivative
evaluations: neighbour build, smoothing, EoS inversion, transport,
gradients, equations of motion, RK update, freeze-out check).  Inputs are
resident in HBM before the timed region.

Prints one JSON line with the throughput in particle-steps/s, the roofline
of the dominant kernel (measured with HIP events on the library's stream
over the timed region) and the CPU oracle timed on a bounded sample on the
host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# MI355X published peaks
PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6

# Algorithmic work per unit, FMA = 2 flop, sqrt/div = 1 (frozen constants).
FLOP_PER_PAIR_SMOOTH = 28
FLOP_PER_PAIR_GRAD_SHEAR = 95
FLOP_PER_PAIR_GRAD_IDEAL = 62


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", choices=["gubser", "lumpy40k"], default="gubser")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle sample")
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--cells-per-rc", type=int, default=3, help="neighbour-grid cells per cutoff radius")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one plain 256 MiB device copy (reference kernel for HBM byte counters)")
    return ap.parse_args()


def workload(name):
    from ccake_amd import config, eos_tables, ic
    if name == "gubser":
        d = ic.gubser_ideal(spacing=0.05)
        return d, config.gubser_ideal_params(), None, dict(
            workload="ideal Gubser 2+1D, conformal EoS, 0.05 fm lattice", particles=len(d["x"]),
            hT=0.3, dt=0.005, tau0=1.0, rk=2)
    d = ic.lumpy(spacing=0.05)
    cfg = config.lumpy_params()
    table = eos_tables.tanh_conformal_table()
    return d, cfg, table, dict(
        workload="viscous BSQ 2+1D lumpy event, 0.05 fm lattice, 64x32x32x32 table EoS",
        particles=len(d["x"]), hT=cfg["hT"], dt=cfg["dt"], tau0=cfg["t0"], rk=2,
        eta_over_s=0.08, zeta_over_s="gaussian 0.1 at T=0.17 GeV, width 0.02 GeV", T_freeze_out_GeV=0.15)


def cpu_baseline(d, cfg, table, steps):
    """The oracle (OpenMP C restatement) on the same event: initialise
    untimed, then time `steps` RK2 steps of the full event."""
    import oracle
    T_FO = 150.0 / 197.3269804
    oracle.build()
    if table is None:
        eos = oracle.Eos(0, T_FO)
    else:
        tab, dims, mins, maxs = table
        eos = oracle.Eos(1, T_FO, dims, mins, maxs, tab)
    if cfg["particlization"]:
        cfg = dict(cfg, e_freeze=eos.efreeze(T_FO))
    n = len(d["x"])
    ev = oracle.Event(cfg, eos, n)
    for k, f in [("x", "X"), ("y", "Y"), ("e", "IN_E"), ("ux", "UX"), ("uy", "UY"),
                 ("rhoB", "IN_B"), ("rhoS", "IN_SS"), ("rhoQ", "IN_Q")]:
        ev.set("ORC_F_" + f, d[k])
    if ev.initialize() != 0:
        raise RuntimeError("oracle initialize failed")
    t0 = time.perf_counter()
    for _ in range(steps):
        if ev.advance_timestep(cfg["dt"], 2) != 0:
            raise RuntimeError("oracle step failed")
    sec = time.perf_counter() - t0
    return dict(value=n * steps / sec, unit="particle-steps/s", cores=int(oracle.get_threads()), kind="port",
                sample=f"{steps} RK2 step(s) of the full {n}-particle event after an untimed initialize "
                       f"({sec:.1f} s on {oracle.get_threads()} OpenMP threads)")


def copy_calibration(nbytes=256 << 20):
    """One plain device-to-device copy of known size through the HIP runtime
    (its copy kernel reads and writes `nbytes`); HBM byte counters of the
    path's kernels are scaled by what the counters read for it."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    if hip.hipMemset(a, 0, C.c_size_t(nbytes)) != 0 or hip.hipDeviceSynchronize() != 0:
        raise RuntimeError("hipMemset failed")
    if hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) != 0 or hip.hipDeviceSynchronize() != 0:   # 3: device to device
        raise RuntimeError("hipMemcpy failed")
    hip.hipFree(a)
    hip.hipFree(b)
    return nbytes


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    import ccake_amd
    from ccake_amd import config

    d, cfg, table, wl = workload(args.workload)
    cfg = dict(cfg, cells_per_rc=args.cells_per_rc)
    wl["cells_per_rc"] = args.cells_per_rc
    n = len(d["x"])
    T_FO = 150.0 / config.HBARC_MEVFM
    print(f"bench: {wl['workload']}: {n} particles, warmup {args.warmup}, steps {args.steps}",
          file=sys.stderr, flush=True)

    dev = ccake_amd.Device(0)
    if table is None:
        eos = ccake_amd.Eos(dev, config.EOS_CONFORMAL, T_FO)
    else:
        tab, dims, mins, maxs = table
        eos = ccake_amd.Eos(dev, config.EOS_TABLE, T_FO, dims, mins, maxs, tab)
    if cfg["particlization"]:
        cfg = dict(cfg, e_freeze=eos.efreeze(T_FO))
    ev = ccake_amd.Event(dev, eos, cfg, n)
    ccake_amd.load_event(ev, d)
    ev.initialize()
    if args.warmup:
        ev.run_steps(cfg["dt"], args.warmup)
    dev.sync()

    dev.profile(True)
    dev.profile_reset()
    dev.sync()
    t0 = time.perf_counter()
    done = ev.run_steps(cfg["dt"], args.steps)
    dev.sync()
    sec = time.perf_counter() - t0
    dev.profile(False)
    if done != args.steps:
        print(f"bench: evolution ended after {done} of {args.steps} steps (all particles frozen)",
              file=sys.stderr, flush=True)
    steps = max(done, 1)

    prof = {}
    for k in ccake_amd.KERNELS:
        if k == "CCAKE_NUM_KERNELS":
            continue
        ms, cnt = dev.profile_read(k)
        prof[k[len("CCAKE_K_"):]] = (ms, cnt)
    pairs = ev.last_pair_count()
    stages = 2 * steps
    print("bench: per-stage ms: " + " ".join(
        f"{k}={v[0] / max(v[1], 1):.3f}x{v[1]}" for k, v in prof.items() if v[1]), file=sys.stderr, flush=True)
    print(f"bench: {pairs} directed pairs per stage, {sec / steps * 1e3:.3f} ms per step", file=sys.stderr,
          flush=True)

    # dominant kernel group by measured time, with its algorithmic work
    dom = max(prof, key=lambda k: prof[k][0])
    gms, gcnt = prof["GRAD"]
    flop_grad = FLOP_PER_PAIR_GRAD_SHEAR if cfg["using_shear"] else FLOP_PER_PAIR_GRAD_IDEAL
    if dom == "SMOOTH":
        sms, scnt = prof["SMOOTH"]
        achieved = FLOP_PER_PAIR_SMOOTH * pairs / (sms / scnt * 1e-3) / 1e12
        kern, per = "smoothing (k_gather_smooth + k_smooth)", f"{FLOP_PER_PAIR_SMOOTH} flop/pair"
    else:
        # the gradient pair kernel (k_grad): the path's FP64 VALU-bound kernel
        achieved = flop_grad * pairs / (gms / gcnt * 1e-3) / 1e12
        kern, per = "gradients (k_grad)", f"{flop_grad} flop/pair"
    roofline = dict(bound="mfma", achieved=round(achieved, 3), peak=PEAK_FP64_TFLOPS, unit="TFLOP/s",
                    frac=round(achieved / PEAK_FP64_TFLOPS, 4), traffic=None,
                    kernel=kern, work=f"{per} x {pairs} pairs per launch", engine="FP64 VALU",
                    dominant_group=dom, avg_launch_ms=round((gms / gcnt) if dom != "SMOOTH" else
                                                            prof["SMOOTH"][0] / prof["SMOOTH"][1], 4))

    out = {
        "metric": "particle-steps/s",
        "value": n * steps / sec,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": sec / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": wl,
        "roofline": roofline,
        "stage_ms": {k: round(v[0] / max(v[1], 1), 4) for k, v in prof.items() if v[1]},
    }
    ev.close(); eos.close(); dev.close()
    if args.calibrate_copy:
        out["copy_calibration_bytes"] = copy_calibration()
    if not args.no_cpu:
        print("bench: timing the CPU oracle sample", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(d, cfg, table, args.cpu_steps)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
