"""Throughput of the per-timestep SPH path on one MI355X.

Default workload (`--workload lumpy`): the viscous BSQ 2+1D lumpy event on
the 0.02 fm lattice over [-5, 5]^2 (251,001 lattice points; the 200,081
above the reference's default energy cutoff of 0.15 GeV/fm^3 are kept, as
its initial-condition readers do, which is what lets the event reach
freeze-out), 50 Gaussian blobs with charges as scaled blobs, the
64x32x32x32 tanh-modulated conformal table EoS with its fallback chain,
eta/s = 0.08, Gaussian zeta/s (0.1 at T = 0.17 GeV, width 0.02 GeV),
particlization at T = 150 MeV, h = 0.3 fm, RK2 with dt = 0.02 fm/c from
tau0 = 0.6 fm/c.  `--workload gubser` runs ideal 2+1D Gubser flow on the
0.05 fm lattice (40,401 particles, conformal EoS, dt = 0.005 fm/c);
`--workload lumpy40k` the lumpy event on the 0.05 fm lattice, `--workload
lumpy1m` on the 0.01 fm lattice (1,002,001 lattice points).

A "step" is one RK2 time step of the whole event (two derivative
evaluations: neighbour build, smoothing, EoS inversion, transport,
gradients, equations of motion, RK update, freeze-out check).  Inputs are
resident in HBM before the timed region.

Prints one JSON line with the throughput in particle-steps/s over the timed
steps; the whole event run from initialize to complete freeze-out
(SPHWorkstation::continue_evolution) on the plain path, as events/s; the
roofline of the dominant kernel (measured with HIP events on the event's
stream over the timed region); the CPU oracle timed on a bounded sample of
the same event on the host cores, as a parallel and a reference-faithful
(serial EoS loop) variant with per-phase times; the same workload with
the mixed-precision pair sums, with the per-field error of its freeze-out
surface against the FP64 run ("mixed_precision"); a batch of lognormal
lumpy events run to freeze-out one at a time and as one batch, in FP64 and
mixed precision ("batching"); and a served multi-event measurement
("serving").
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

# Algorithmic work per unit, FMA = 2 flop, sqrt/div = 1 (frozen constants),
# per launched variant of the pair kernels.
# Smoothing, per directed pair: distance 5, sqrt 1, q 1, W 7, sigma_lab
# m W (2), then m ext W (3) for each of s, rho_B, rho_S, rho_Q: 28; the
# charge-free variant (k_smooth_slice<3>, an event without charges) 19.
FLOP_PER_PAIR_SMOOTH = {True: 28, False: 19}               # charges on / off
# Gradients, per directed pair (2D): 95 with shear on, 62 with it off; the
# bulk-free variant (no grad Pi sums: an ideal fluid with zeta = 0) skips 4.
FLOP_PER_PAIR_GRAD = {(True, True): 95, (True, False): 91, (False, True): 62, (False, False): 58}
# 3+1D (D = 3, isotropic kernel), per directed pair: smoothing 37 (offsets 3,
# distance 5, q and 1/q 9, spline 10, five accumulations 10); gradients 99
# with shear on (offsets 3, distance 5, q and 1/q 9, spline gradient 9,
# weight 4, 21 accumulations 39, shear 30), 69 with it off.
FLOP_PER_PAIR_SMOOTH3 = 37
FLOP_PER_PAIR_GRAD3 = {True: 99, False: 69}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", choices=["lumpy", "gubser", "lumpy40k", "lumpy1m"], default="lumpy")
    ap.add_argument("--no-event", action="store_true", help="skip the run of the whole event to freeze-out")
    ap.add_argument("--diffusion", action="store_true",
                    help="BSQ diffusion on (diagonal kappa/T^2 = 0.1, regulator on, q = 0 at tau0)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle sample")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--cells-per-rc", type=int, default=4,
                    help="neighbour-grid cells per cutoff radius (4 measured fastest on the 200k lumpy event: "
                         "6.29 ms per step against 6.56 with 3)")
    ap.add_argument("--x-subcells", type=int, default=8,
                    help="x cells per cell of a key row (8: measured fastest on the 200k lumpy event)")
    ap.add_argument("--no-serve", action="store_true", help="skip the served multi-event measurement")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision measurement")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched multi-event measurement")
    ap.add_argument("--no-milne", action="store_true", help="skip the 3+1D Milne event and served 3+1D stream")
    ap.add_argument("--no-milne-stream", action="store_true", help="skip the served 3+1D stream")
    ap.add_argument("--no-1m", action="store_true", help="skip the 0.01 fm lattice event (FP64 and mixed)")
    ap.add_argument("--batch-events", type=int, default=16)
    ap.add_argument("--large-batch", type=int, default=0,
                    help="also run a batch of this many lognormal lumpy events in mixed precision (512: "
                         "about 40M particles, several minutes)")
    ap.add_argument("--serve-events", type=int, default=24)
    ap.add_argument("--serve-slots", type=int, default=4)
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
    # points at or below the reference's default energy cutoff, 0.15 GeV/fm^3,
    # are dropped as its initial-condition readers do
    spacing = {"lumpy": 0.02, "lumpy1m": 0.01}.get(name, 0.05)
    d = ic.lumpy(spacing=spacing, e_cutoff_gev=0.15)
    cfg = config.lumpy_params()
    table = eos_tables.tanh_conformal_table()
    return d, cfg, table, dict(
        workload=f"viscous BSQ 2+1D lumpy event, {spacing} fm lattice over [-5,5]^2, e > 0.15 GeV/fm^3, "
                 "64x32x32x32 table EoS, run to freeze-out at T = 150 MeV",
        lattice_points=len(ic.lattice(spacing, 5.0)[0]), particles=len(d["x"]), hT=cfg["hT"], dt=cfg["dt"],
        tau0=cfg["t0"], rk=2, eta_over_s=0.08, zeta_over_s="gaussian 0.1 at T=0.17 GeV, width 0.02 GeV",
        T_freeze_out_GeV=0.15)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(d, cfg, table, steps):
    """The oracle (OpenMP C restatement) on the same event: initialise
    untimed, then time `steps` RK2 steps of the full event twice: with every
    per-particle loop parallel, and reference-faithful with the EoS inversion
    as a serial loop (src/sph_workstation.cpp:1737-1762).  Per-phase wall
    times use the reference's phase names."""
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
    variants = {}
    for name, serial in (("parallel", False), ("reference_serial_eos", True)):
        oracle.set_serial_eos(serial)
        oracle.phase_reset()
        t0 = time.perf_counter()
        for _ in range(steps):
            if ev.advance_timestep(cfg["dt"], 2) != 0:
                raise RuntimeError("oracle step failed")
        sec = time.perf_counter() - t0
        variants[name] = dict(particle_steps_per_s=round(n * steps / sec, 1), seconds=round(sec, 3),
                              phases_s={k: round(v[0], 4) for k, v in oracle.phase_times().items()})
        print(f"bench: CPU {name}: {sec:.2f} s for {steps} step(s)", file=sys.stderr, flush=True)
    oracle.set_serial_eos(False)
    thr = int(oracle.get_threads())
    return dict(value=variants["parallel"]["particle_steps_per_s"], unit="particle-steps/s", cores=thr,
                kind="port", cpu=cpu_model(),
                sample=f"{steps} RK2 step(s) of the full {n}-particle event after an untimed initialize, "
                       f"each variant; {thr} OpenMP threads (the EoS loop on one thread in the "
                       f"reference-faithful variant)",
                variants=variants)


# Compulsory HBM bytes per particle of the fused per-particle kernel k_eos<0>
# (EoS inversion + reset_pi_tensor + transport + gradient-record gather):
# reads 24 doubles (seed T, mu; s, rho; gamma, sigma, u; extensive Pi and 5
# shear; sigma_lab, x, y, m_s, v) + 2 int32 (perm, EoS member); writes 77
# doubles (31 thermo, Pi, 16 pi^{mu nu}, 13 transport, 16-double record)
# + 1 int32.
BYTES_PER_PARTICLE_EOS = 24 * 8 + 2 * 4 + 77 * 8 + 4

# The fused EoS kernel of a stage (k_eos<3, .>) also runs the smoothing sum
# pass: S slices x 6 partial sums read, gamma and btrack read, btrack and ten
# smoothed and local-rest-frame densities written.
def bytes_per_particle_eos(nslices):
    return BYTES_PER_PARTICLE_EOS + 48 * nslices + 8 + 8 + 4 + 80


KERNEL_OF_GROUP = {"SMOOTH": "k_smooth_slice", "GRAD": "k_grad_slice", "EOS": "k_eos"}
# With 4 or more slices per tile the FP64 2+1D pair kernels run a tile's
# slices as the waves of one workgroup and add them in LDS (k_smooth_tile,
# k_grad_tile; CCAKE_LDS_SLICES=0 turns that off), and the sum passes read
# one slice.
KERNEL_OF_GROUP_LDS = {"SMOOTH": "k_smooth_tile", "GRAD": "k_grad_tile", "EOS": "k_eos"}


def lds_tiles(nslices, cfg):
    return (nslices >= 4 and cfg.get("pair_precision", 0) == 0 and cfg.get("dim", 2) == 2
            and os.environ.get("CCAKE_LDS_SLICES", "1") != "0")


def traffic_of(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by the same
    counter's reading for a plain copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    d = json.load(open(path))
    k = d.get("kernels", {}).get(kernel)
    if not k:
        return None
    c = d["copy"]
    f_scale = c["bytes"] / (c["fetch_kib"] * 1024.0)
    w_scale = c["bytes"] / (c["write_kib"] * 1024.0)
    return round(k["fetch_kib"] * 1024.0 * f_scale + k["write_kib"] * 1024.0 * w_scale)


def make_roofline(prof, pairs, n, cfg, workload, charges, bulk, nslices):
    """Roofline of the dominant kernel group (by measured time) among the
    three with a work model: the smoothing and gradient pair kernels are
    FP64-VALU bound (work = flop per directed pair of the launched variant x
    pairs per launch; no matrix work on this path, so the bound is the FP64
    vector rate), the fused EoS kernel is HBM bound (work = compulsory bytes
    per particle x N)."""
    cand = {g: prof[g] for g in KERNEL_OF_GROUP if prof[g][1]}
    dom = max(cand, key=lambda g: cand[g][0])
    ms = cand[dom][0] / cand[dom][1]
    sec = max(ms * 1e-3, 1e-12)
    lds = lds_tiles(nslices, cfg)
    names = KERNEL_OF_GROUP_LDS if lds else KERNEL_OF_GROUP
    if dom == "EOS":
        bpp = bytes_per_particle_eos(1 if lds else nslices)
        achieved = bpp * n / sec / 1e9
        r = dict(bound="hbm", achieved=round(achieved, 2), peak=PEAK_HBM_GBS, unit="GB/s",
                 frac=round(achieved / PEAK_HBM_GBS, 4),
                 work=f"{bpp} B/particle x {n} particles per launch")
    else:
        shear = bool(cfg["using_shear"])
        fpp = FLOP_PER_PAIR_SMOOTH[charges] if dom == "SMOOTH" else FLOP_PER_PAIR_GRAD[(shear, bulk)]
        achieved = fpp * pairs / sec / 1e12
        r = dict(bound="valu", achieved=round(achieved, 3), peak=PEAK_FP64_TFLOPS, unit="TFLOP/s",
                 frac=round(achieved / PEAK_FP64_TFLOPS, 4), engine="FP64 VALU (no matrix work on this path)",
                 work=f"{fpp} flop/pair x {pairs:.0f} pairs per launch")
    r["traffic"] = traffic_of(names[dom], workload)
    r["kernel"] = names[dom]
    r["avg_launch_ms"] = round(ms, 4)
    return r


def whole_event(dev, eos, d, cfg, max_tau=30.0, rows=None):
    """One event from initialize until every particle has frozen out
    (SPHWorkstation::continue_evolution checked after every step, as
    BSQHydro::run does), wall-clocked on the host: the event latency.
    With `rows` (a list) the freeze-out hypersurface rows are appended to
    it after the clock stops."""
    import ccake_amd
    n = len(d["x"])
    ev = ccake_amd.Event(dev, eos, cfg, n)
    ccake_amd.load_event(ev, d)
    dev.sync()
    t0 = time.perf_counter()
    ev.initialize()
    steps = 0
    while ev.continue_evolution(max_tau=max_tau):
        ev.advance_timestep(cfg["dt"], 2)
        steps += 1
    dev.sync()
    sec = time.perf_counter() - t0
    tau = ev.t
    if rows is not None:
        rows.append(ev.freeze_out_surface())
    ev.close()
    return dict(seconds=round(sec, 4), events_per_s=round(1.0 / sec, 4), steps=steps, tau_end=round(tau, 4),
                particle_steps_per_s=round(n * steps / sec, 1))


def timed_steps(dev, eos, d, cfg, warmup, steps):
    """ms per RK2 step of the event over `steps` steps after `warmup`."""
    import ccake_amd
    ev = ccake_amd.Event(dev, eos, cfg, len(d["x"]))
    ccake_amd.load_event(ev, d)
    ev.initialize()
    if warmup:
        ev.run_steps(cfg["dt"], warmup)
    dev.sync()
    t0 = time.perf_counter()
    done = ev.run_steps(cfg["dt"], steps)
    dev.sync()
    sec = time.perf_counter() - t0
    ev.close()
    return sec / max(done, 1) * 1e3


def mixed_precision(dev, eos, d, cfg, warmup, steps, fp64_ms, fp64_event, fp64_rows):
    """The same workload with the mixed-precision pair sums (pair_precision
    1: FP32 pair terms, compensated FP32 accumulation, FP64 elsewhere):
    step time, the whole event to freeze-out, and the per-field error of its
    freeze-out hypersurface against the FP64 run's (rows matched by particle
    id; ccake_amd.precision gives the measures)."""
    mcfg = dict(cfg, pair_precision=1)
    ms = timed_steps(dev, eos, d, mcfg, warmup, steps)
    out = dict(ms_per_step=round(ms, 4), particle_steps_per_s=round(len(d["x"]) / ms * 1e3, 1),
               step_speedup_vs_fp64=round(fp64_ms / ms, 3))
    if fp64_event is not None:
        rows = []
        ev = whole_event(dev, eos, d, mcfg, rows=rows)
        out["event_to_freeze_out"] = ev
        out["event_speedup_vs_fp64"] = round(fp64_event["seconds"] / ev["seconds"], 3)
        out["freeze_out_error_vs_fp64"] = freeze_out_errors(fp64_rows[0], rows[0])
    return out


def freeze_out_errors(ref_rows, rows):
    """Per-field error of a freeze-out hypersurface against the FP64 run's
    (rows matched by particle id; ccake_amd.precision gives the measures)."""
    from ccake_amd import precision
    err = precision.surface_errors(ref_rows, rows)
    keep = ("T", "u_mu", "rho_B", "rho_S", "rho_Q", "s", "e", "pi_11", "pi_22", "pi_12", "Pi")
    return dict(
        rows=err["rows_ref"], same_step=err["same_step"], same_snapshot=err["same_snapshot"],
        measure="per particle: relative (signed fields against 1e-3 of their largest magnitude); "
                "u_mu: |du^mu| / u^0; whole: max|d| / max|x|",
        fields={k: {m: float(f"{v:.3g}") for m, v in err["fields"][k].items()} for k in keep})


def lattice_1m(dev, eos, cfg):
    """The lumpy event on the 0.01 fm lattice (1,002,001 points over
    [-5, 5]^2, the same blobs, EoS and transport as the 0.02 fm event; h
    stays 0.3 fm, so a particle has about 11,000 neighbours): FP64 and
    mixed-precision pair sums, each from initialize to complete freeze-out,
    and the per-field error of the mixed run's hypersurface against the
    FP64 run's."""
    d, _, _, wl = workload("lumpy1m")
    n = len(d["x"])
    out = dict(workload=wl["workload"], lattice_points=wl["lattice_points"], particles=n)
    fp64_rows, rows = [], []
    print(f"bench: 0.01 fm lattice, {n} particles, FP64 to freeze-out", file=sys.stderr, flush=True)
    out["fp64"] = whole_event(dev, eos, d, cfg, rows=fp64_rows)
    out["fp64"]["ms_per_step"] = round(out["fp64"]["seconds"] / max(out["fp64"]["steps"], 1) * 1e3, 3)
    print(f"bench: 1m fp64 {json.dumps(out['fp64'])}; mixed precision to freeze-out", file=sys.stderr, flush=True)
    out["mixed"] = whole_event(dev, eos, d, dict(cfg, pair_precision=1), rows=rows)
    out["mixed"]["ms_per_step"] = round(out["mixed"]["seconds"] / max(out["mixed"]["steps"], 1) * 1e3, 3)
    out["mixed"]["event_speedup_vs_fp64"] = round(out["fp64"]["seconds"] / out["mixed"]["seconds"], 3)
    out["mixed"]["freeze_out_error_vs_fp64"] = freeze_out_errors(fp64_rows[0], rows[0])
    return out


def large_batch(dev, eos, cfg, k, sample=32, seed=0xC3):
    """k lumpy events (lognormal sizes as in batched_run) in one batch with
    the mixed-precision pair sums, run until every event has frozen out,
    against the plain FP64 path (each event alone, one after the other)
    timed on the first `sample` events of the same draw: the plain path's
    events/s does not depend on how many events queue behind each other.
    The cell size and slice count are tuned on the full batch."""
    import numpy as np
    from ccake_amd import batching, precision
    print(f"bench: building {k} lumpy events", file=sys.stderr, flush=True)
    ics = batching.lumpy_batch_ics(k, seed=seed)
    n = [len(d["x"]) for d in ics]
    ntot = int(sum(n))
    print(f"bench: {ntot} particles; autotune", file=sys.stderr, flush=True)
    tune = batching.autotune(dev, eos, dict(cfg, pair_precision=1), ics, cells_per_rc=(3, 4), slices=(1, 2, 4))
    print(f"bench: autotune {json.dumps(tune)}; plain path on {sample} events", file=sys.stderr, flush=True)
    alone = batching.run_alone(dev, eos, cfg, ics[:sample])
    plain = batching.latency_summary(alone, int(sum(n[:sample])))
    print(f"bench: plain {json.dumps(plain)}; batch", file=sys.stderr, flush=True)

    def progress(steps, done):
        print(f"bench: batch step {steps}, {done} of {k} events frozen out", file=sys.stderr, flush=True)
    run = batching.run_batch(dev, eos, dict(cfg, cells_per_rc=tune["cells_per_rc"], pair_precision=1), ics,
                             slices=tune["slices"], progress=progress)
    r = batching.latency_summary(run, ntot)
    r["batch_steps"] = run["steps"]
    r["particle_steps_per_s"] = round(ntot * run["steps"] / run["wall_s"], 1)
    r["speedup_vs_plain"] = round(r["events_per_s"] / plain["events_per_s"], 3)
    errs = [precision.surface_errors(a, b)["fields"] for a, b in zip(alone["rows"], run["rows"][:sample])]
    r["freeze_out_error_vs_fp64"] = {f: float(f"{max(e[f]['max'] for e in errs):.3g}") for f in ("T", "u_mu", "s", "e")}
    r["freeze_out_error_vs_fp64"].update({f + "_whole": float(f"{max(e[f]['whole'] for e in errs):.3g}")
                                          for f in ("rho_B", "rho_S", "rho_Q")})
    r["freeze_out_error_events"] = f"the first {sample} events (their FP64 runs alone)"
    return dict(workload=f"{k} viscous BSQ lumpy events on the 0.02 fm lattice, lognormal sizes (median 80,000 "
                         "lattice points, [20,000, 300,000]), e > 0.15 GeV/fm^3, table EoS, mixed-precision pair "
                         "sums, each run to freeze-out, all arriving at once",
                particles_min=min(n), particles_max=max(n), particles_median=int(np.median(n)), autotune=tune,
                plain_fp64_sample=plain, batched_mixed=r)


def serving_requests(k, steps, seed=0x5E4F):
    """k ideal Gubser events on the 0.05 fm lattice with lognormal particle
    counts (median 25,000, clipped to [10,000, 60,000]) set by the lattice
    half-width; `steps` RK2 steps each; all arriving at once."""
    import numpy as np
    from ccake_amd import config, ic, serve
    rng = np.random.default_rng(seed)
    reqs = []
    for rid in range(k):
        n_want = float(np.clip(np.exp(rng.normal(np.log(25000.0), 0.5)), 10000, 60000))
        half = 0.05 * round((np.sqrt(n_want) - 1.0) / 2.0)
        reqs.append(serve.Request(rid=rid, ic=ic.gubser_ideal(spacing=0.05, half_width=half),
                                  params=config.gubser_ideal_params(), max_steps=steps,
                                  fields=("CCAKE_F_E", "CCAKE_F_UX", "CCAKE_F_UY")))
    return reqs


def serving_run(dev, events, slots, steps=40):
    """Events/s and latency of the same request batch run one event at a
    time (the plain path) and served `slots` events at a time, each on its
    own stream.  The served results must equal the sequential ones."""
    import numpy as np
    import ccake_amd
    from ccake_amd import serve
    reqs = serving_requests(events, steps)
    eos = ccake_amd.Eos(dev, 0)
    serve.Server(dev, eos, slots=slots).run(reqs[:min(len(reqs), slots)])   # warm-up (code objects, pools)
    out = {}
    runs = {}
    for name, sl in (("sequential", 1), ("served", slots)):
        dev.sync()
        t0 = time.perf_counter()
        res = serve.Server(dev, eos, slots=sl).run(reqs)
        wall = time.perf_counter() - t0
        runs[name] = res
        out[name] = dict(slots=sl, **{k: round(v, 4) if isinstance(v, float) else v
                                      for k, v in serve.summarize(res, wall).items()})
    eos.close()
    same = all(np.array_equal(a.fields[f], b.fields[f]) for a, b in zip(runs["sequential"], runs["served"])
               for f in a.fields)
    n = [len(r.ic["x"]) for r in reqs]
    out.update(events=events, steps_per_event=steps, particles_min=min(n), particles_max=max(n),
               particles_median=int(np.median(n)), results_equal=bool(same),
               speedup=round(out["served"]["events_per_s"] / out["sequential"]["events_per_s"], 3),
               workload="ideal Gubser 2+1D events, 0.05 fm lattice, lognormal particle counts, "
                        "all arriving at once")
    return out


def batched_run(dev, eos, cfg, k, seed=0xC3):
    """k lumpy events with lognormal sizes (median 80,000 lattice points,
    truncated to [20,000, 300,000]; 0.02 fm lattice, energy cutoff), all
    arriving at once: the plain path (each event alone, one after the other)
    against one batch of all of them with the autotuned cell size and slice
    count, in FP64 and with the mixed-precision pair sums.  Every event runs
    to complete freeze-out; latency is each event's completion time."""
    import numpy as np
    from ccake_amd import batching, precision
    ics = batching.lumpy_batch_ics(k, seed=seed)
    n = [len(d["x"]) for d in ics]
    ntot = int(sum(n))
    tune = batching.autotune(dev, eos, cfg, ics)
    tcfg = dict(cfg, cells_per_rc=tune["cells_per_rc"])
    alone = batching.run_alone(dev, eos, cfg, ics)
    out = dict(workload=f"{k} viscous BSQ lumpy events on the 0.02 fm lattice, lognormal sizes (median 80,000 "
                        "lattice points, [20,000, 300,000]), e > 0.15 GeV/fm^3, table EoS, each run to freeze-out, "
                        "all arriving at once",
               particles_min=min(n), particles_max=max(n), particles_median=int(np.median(n)),
               autotune=tune, plain=batching.latency_summary(alone, ntot))
    for name, pp in (("batched", 0), ("batched_mixed", 1)):
        run = batching.run_batch(dev, eos, dict(tcfg, pair_precision=pp), ics, slices=tune["slices"])
        r = batching.latency_summary(run, ntot)
        r["batch_steps"] = run["steps"]
        r["speedup_vs_plain"] = round(r["events_per_s"] / out["plain"]["events_per_s"], 3)
        same = all(a.shape == b.shape and np.array_equal(a[:, :2], b[:, :2]) for a, b in zip(alone["rows"], run["rows"]))
        r["rows_same_ids_and_steps"] = bool(same)
        if pp == 0 and same:
            r["max_column_difference_vs_plain"] = float(max(
                float(np.max(np.abs(b[:, 2:] - a[:, 2:])) / max(float(np.max(np.abs(a[:, 2:]))), 1e-300))
                for a, b in zip(alone["rows"], run["rows"]) if len(a)))
        if pp == 1:
            errs = [precision.surface_errors(a, b)["fields"] for a, b in zip(alone["rows"], run["rows"])]
            r["freeze_out_error_vs_fp64"] = {f: float(f"{max(e[f]['max'] for e in errs):.3g}")
                                             for f in ("T", "u_mu", "s", "e")}
            r["freeze_out_error_vs_fp64"].update({f + "_whole": float(f"{max(e[f]['whole'] for e in errs):.3g}")
                                                  for f in ("rho_B", "rho_S", "rho_Q")})
        out[name] = r
    out["served_stream"] = served_batch_stream(dev, eos, cfg, ics, tune, out["plain"])
    return out


def served_batch_stream(dev, eos, cfg, ics, tune, plain, rate=50.0, slots=3, window=0.1):
    """The same events as a Poisson stream (rate events/s) served by the
    continuous-batching loop (serve.BatchServer: each free worker forms a
    batch of the requests that have arrived, every event returned as soon as
    it has frozen out), in FP64 and mixed precision, against the plain path's
    events/s (each event alone, one after the other).  Three workers and a
    100 ms window measured best on the 16-event stream (windows of 50, 100,
    200 and 400 ms with 2 and 3 workers)."""
    import numpy as np
    from ccake_amd import serve
    arr = serve.poisson_arrivals(len(ics), rate)
    out = dict(workload=f"the batch's {len(ics)} events arriving as a Poisson stream at {rate:g} events/s, "
                        f"{slots} workers forming batches (window {window * 1e3:g} ms), each event returned "
                        "when it has frozen out", arrivals_span_s=round(float(arr[-1]), 3))
    for name, pp in (("fp64", 0), ("mixed", 1)):
        pcfg = dict(cfg, cells_per_rc=tune["cells_per_rc"], pair_precision=pp)
        reqs = [serve.Request(rid=k, ic=d, params=pcfg, arrival=float(a)) for k, (d, a) in enumerate(zip(ics, arr))]
        dev.sync()
        t0 = time.perf_counter()
        res = serve.BatchServer(dev, eos, pcfg, slots=slots, max_events=len(ics), slices=tune["slices"],
                                window=window).run(reqs)
        wall = time.perf_counter() - t0
        lat = np.array([r.latency for r in res]) * 1e3
        out[name] = dict(events_per_s=round(len(res) / wall, 4), wall_s=round(wall, 3),
                         p50_ms=round(float(np.percentile(lat, 50)), 1), p99_ms=round(float(np.percentile(lat, 99)), 1),
                         batches=len({r.batch for r in res}),
                         speedup_vs_plain=round(len(res) / wall / plain["events_per_s"], 3))
    return out


def milne_workload(half_width=3.0, center_sd=1.0, seed=0xCCA4E001):
    """3+1D lumpy event in Milne coordinates: 0.1 fm transverse lattice over
    [-L, L]^2, eta_s in [-4, 4] at 0.1, 20 blobs (centres N(0, center_sd))
    times a longitudinal Gaussian plateau (flat over |eta_s| <= 1, width
    0.5), points above 0.15 GeV/fm^3; table EoS, viscous BSQ as the lumpy
    2+1D event; isotropic kernel hT = 0.3 in (x, y, eta_s); two cells per
    cutoff radius and six x cells per cell (measured fastest: 4.64 ms per
    step against 5.07 with cubic cells and 5.61 with three cells per
    radius)."""
    from ccake_amd import config, ic
    d = ic.milne_lumpy(seed=seed, spacing=0.1, half_width=half_width, eta_spacing=0.1, eta_max=4.0,
                       center_sd=center_sd, e_cutoff_gev=0.15)
    cfg = config.lumpy_params(dim=3, hEta=0.3, cells_per_rc=2, x_subcells=6)
    return d, cfg


def milne_run(dev, eos, steps=20, warmup=2):
    """One 3+1D event: timed RK2 steps (particle-steps/s, per-stage times,
    the gradient kernel's roofline) and the whole event from initialize to
    complete freeze-out."""
    import ccake_amd
    from ccake_amd import config
    d, cfg = milne_workload()
    T_FO = 150.0 / config.HBARC_MEVFM
    cfg = dict(cfg, e_freeze=eos.efreeze(T_FO))
    n = len(d["x"])
    ev = ccake_amd.Event(dev, eos, cfg, n)
    ccake_amd.load_event(ev, d)
    ev.initialize()
    ev.run_steps(cfg["dt"], warmup)
    dev.sync()
    dev.profile(True)
    dev.profile_reset()
    ev.pair_total(reset=True)
    dev.sync()
    t0 = time.perf_counter()
    done = ev.run_steps(cfg["dt"], steps)
    dev.sync()
    sec = time.perf_counter() - t0
    dev.profile(False)
    prof = {}
    for k in ccake_amd.KERNELS:
        if k != "CCAKE_NUM_KERNELS":
            ms, cnt = dev.profile_read(k)
            prof[k[len("CCAKE_K_"):]] = (ms, cnt)
    pairs = ev.pair_total() / max(prof["SMOOTH"][1], 1)
    ev.close()
    gms = prof["GRAD"][0] / max(prof["GRAD"][1], 1)
    sms = prof["SMOOTH"][0] / max(prof["SMOOTH"][1], 1)
    fg = FLOP_PER_PAIR_GRAD3[bool(cfg["using_shear"])]
    ach = fg * pairs / (gms * 1e-3) / 1e12
    out = dict(workload="3+1D lumpy viscous BSQ event in Milne coordinates (tau, x, y, eta_s): 0.1 fm lattice over "
                        "[-3, 3]^2, eta_s in [-4, 4] at 0.1, Gaussian plateau (flat |eta_s| <= 1, width 0.5), "
                        "e > 0.15 GeV/fm^3, table EoS, isotropic 3D kernel hT = 0.3",
               particles=n, steps=done, ms_per_step=round(sec / max(done, 1) * 1e3, 4),
               particle_steps_per_s=round(n * done / sec, 1), pairs_per_stage=int(pairs),
               neighbours_per_particle=round(pairs / n, 1),
               stage_ms={k: round(v[0] / max(v[1], 1), 4) for k, v in prof.items() if v[1]},
               roofline=dict(kernel="k3_grad_slice", bound="valu", achieved=round(ach, 3), peak=PEAK_FP64_TFLOPS,
                             unit="TFLOP/s", frac=round(ach / PEAK_FP64_TFLOPS, 4),
                             work=f"{fg} flop/pair x {int(pairs)} pairs per launch", avg_launch_ms=round(gms, 4),
                             smoothing_frac=round(FLOP_PER_PAIR_SMOOTH3 * pairs / (sms * 1e-3) / 1e12
                                                  / PEAK_FP64_TFLOPS, 4)))
    out["event_to_freeze_out"] = whole_event(dev, eos, d, cfg)
    return out


def milne_stream(dev, eos, k=6, rate=50.0, slots=2, seed=0x3D5):
    """Served 3+1D stream: k Milne lumpy events with lognormal particle
    counts N in [100k, 2M] (median 450k; the transverse half-width
    L = 3 fm sqrt(N / 126k), as the 126k-particle event has L = 3 fm; blob
    centres N(0, L/3), 20 (L / 3 fm)^2 blobs, so that the blobs' density and
    the particle count per area stay those of that event), arriving as a Poisson stream at `rate` events/s,
    each run to complete freeze-out; served `slots` events at a time (one
    stream each, admitted as slots free up) against one at a time.  The
    served results must equal the sequential ones bit for bit."""
    import numpy as np
    from ccake_amd import config, ic, serve
    rng = np.random.default_rng(seed)
    T_FO = 150.0 / config.HBARC_MEVFM
    arr = serve.poisson_arrivals(k, rate, seed=seed)
    reqs = []
    for rid in range(k):
        npart = 0.0
        while not 100e3 <= npart <= 2e6:
            npart = float(np.exp(rng.normal(np.log(450e3), 0.6)))
        hw = round(3.0 * np.sqrt(npart / 126e3) * 10) / 10
        d = ic.milne_lumpy(seed=0xCCA4E001 + rid, spacing=0.1, half_width=hw, eta_spacing=0.1, eta_max=4.0,
                           center_sd=hw / 3.0, n_blobs=int(round(20 * (hw / 3.0) ** 2)), e_cutoff_gev=0.15)
        cfg = config.lumpy_params(dim=3, hEta=0.3, cells_per_rc=2, x_subcells=6, e_freeze=eos.efreeze(T_FO))
        reqs.append(serve.Request(rid=rid, ic=d, params=cfg, arrival=float(arr[rid]), max_tau=30.0,
                                  fields=("CCAKE_F_T", "CCAKE_F_E", "CCAKE_F_UX", "CCAKE_F_UY", "CCAKE_F_UETA")))
    n = [len(r.ic["x"]) for r in reqs]
    out = dict(workload=f"{k} 3+1D Milne lumpy events (0.1 fm and 0.1 in eta_s, eta_s in [-4, 4], lognormal "
                        f"particle counts in [100k, 2M], median 450k), Poisson arrivals at {rate:g} events/s, each "
                        "run to freeze-out",
               particles_min=min(n), particles_max=max(n), particles_median=int(np.median(n)),
               particles_total=int(sum(n)))
    runs = {}
    for name, sl in (("sequential", 1), ("served", slots)):
        dev.sync()
        t0 = time.perf_counter()
        res = serve.Server(dev, eos, slots=sl).run(reqs)
        wall = time.perf_counter() - t0
        runs[name] = res
        out[name] = dict(slots=sl, **{kk: round(v, 4) if isinstance(v, float) else v
                                      for kk, v in serve.summarize(res, wall).items()})
    out["results_equal"] = bool(all(np.array_equal(a.fields[f], b.fields[f]) for a, b in
                                    zip(runs["sequential"], runs["served"]) for f in a.fields))
    out["speedup"] = round(out["served"]["events_per_s"] / out["sequential"]["events_per_s"], 3)
    return out


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
        print(f"bench.py measures one GPU: this library has no multi-GPU path (asked for --gpus {args.gpus}). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    import ccake_amd
    from ccake_amd import config

    d, cfg, table, wl = workload(args.workload)
    if args.diffusion:
        cfg = dict(cfg, using_diffusion=1, kappa_over_T2=[0.1, 0.0, 0.0, 0.0, 0.1, 0.0, 0.0, 0.0, 0.1],
                   diffusion_regulator=1)
        wl = dict(wl, workload=wl["workload"] + ", BSQ diffusion (kappa/T^2 = 0.1 I, regulator on)")
    cfg = dict(cfg, cells_per_rc=args.cells_per_rc, x_subcells=args.x_subcells)
    wl["cells_per_rc"] = args.cells_per_rc
    wl["x_subcells"] = args.x_subcells
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
    t_init = time.perf_counter()
    ev.initialize()
    t_init = time.perf_counter() - t_init
    print(f"bench: initialize took {t_init * 1e3:.1f} ms", file=sys.stderr, flush=True)
    if args.warmup:
        ev.run_steps(cfg["dt"], args.warmup)
    dev.sync()

    dev.profile(True)
    dev.profile_reset()
    ev.pair_total(reset=True)
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
    stages = 2 * steps
    pairs_per_launch = ev.pair_total() / max(prof["SMOOTH"][1], 1)
    print("bench: per-stage ms: " + " ".join(
        f"{k}={v[0] / max(v[1], 1):.3f}x{v[1]}" for k, v in prof.items() if v[1]), file=sys.stderr, flush=True)
    print(f"bench: {pairs_per_launch:.0f} directed pairs per stage, {sec / steps * 1e3:.3f} ms per step",
          file=sys.stderr, flush=True)

    import numpy as np
    charges = any(np.any(d[k] != 0) for k in ("rhoB", "rhoS", "rhoQ")) or bool(cfg["using_diffusion"])
    bulk = not (cfg["zeta_mode"] == config.ZETA_CONSTANT and cfg["constant_zeta_over_s"] == 0.0
                and not cfg["using_shear"])
    roofline = make_roofline(prof, pairs_per_launch, n, cfg, args.workload, charges, bulk, ev.nslices())
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
    ev.close()
    fp64_rows = []
    if not args.no_event and cfg["particlization"]:
        print("bench: whole event from initialize to freeze-out", file=sys.stderr, flush=True)
        out["event_to_freeze_out"] = whole_event(dev, eos, d, cfg, rows=fp64_rows)
        print(f"bench: event {json.dumps(out['event_to_freeze_out'])}", file=sys.stderr, flush=True)
    if not args.no_mixed:
        print("bench: mixed-precision pair sums", file=sys.stderr, flush=True)
        out["mixed_precision"] = mixed_precision(dev, eos, d, cfg, args.warmup, steps, sec / steps * 1e3,
                                                 out.get("event_to_freeze_out"), fp64_rows)
        print(f"bench: mixed {json.dumps(out['mixed_precision'])}", file=sys.stderr, flush=True)
    if not args.no_batch and args.workload == "lumpy":
        print(f"bench: batch of {args.batch_events} lognormal lumpy events", file=sys.stderr, flush=True)
        out["batching"] = batched_run(dev, eos, cfg, args.batch_events)
        print(f"bench: batching {json.dumps(out['batching'])}", file=sys.stderr, flush=True)
    if not args.no_1m and args.workload == "lumpy":
        out["lattice_1m"] = lattice_1m(dev, eos, cfg)
        print(f"bench: 1m {json.dumps(out['lattice_1m'])}", file=sys.stderr, flush=True)
    if args.large_batch and args.workload == "lumpy":
        out["large_batch"] = large_batch(dev, eos, cfg, args.large_batch)
        print(f"bench: large batch {json.dumps(out['large_batch'])}", file=sys.stderr, flush=True)
    if not args.no_milne and table is not None:
        print("bench: 3+1D Milne event", file=sys.stderr, flush=True)
        out["milne_3d"] = milne_run(dev, eos)
        print(f"bench: milne {json.dumps(out['milne_3d'])}", file=sys.stderr, flush=True)
        if not args.no_milne_stream:
            print("bench: served 3+1D stream", file=sys.stderr, flush=True)
            out["milne_3d"]["served_stream"] = milne_stream(dev, eos)
            print(f"bench: milne stream {json.dumps(out['milne_3d']['served_stream'])}", file=sys.stderr, flush=True)
    eos.close()
    if not args.no_serve:
        print(f"bench: serving {args.serve_events} events, {args.serve_slots} slots", file=sys.stderr, flush=True)
        out["serving"] = serving_run(dev, args.serve_events, args.serve_slots)
        print(f"bench: serving {json.dumps(out['serving'])}", file=sys.stderr, flush=True)
    dev.close()
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
