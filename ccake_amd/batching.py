"""Many events per launch: the batched form of the path (ccake_batch_create),
its autotuner, and the lognormal batch workload.

A batch steps every event together, so each stage launches each kernel
once for all events: the launch gaps are paid once per batch, the
per-particle kernels see all particles, and small events no longer leave
most of the chip idle.  Each event evolves as it does alone (same pair sets;
fields equal to rounding), and stops counting toward the batch's work only
when the whole batch ends: an event that freezes out early is finished (its
freeze-out surface is complete) but its particles keep being stepped until
the last event of the batch freezes out, as the reference keeps stepping
frozen particles of a running event.
"""
from __future__ import annotations

import time

import numpy as np

from . import Batch, concat_ics, ic, load_event


def lognormal_sizes(k, seed=0xC3, median=80_000, lo=20_000, hi=300_000, sigma=0.5):
    """k particle counts ~ LogNormal(ln median, sigma) truncated to [lo, hi]
    by redrawing."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < k:
        v = float(np.exp(rng.normal(np.log(median), sigma)))
        if lo <= v <= hi:
            out.append(v)
    return out


def lumpy_batch_ics(k, seed=0xC3, spacing=0.02, e_cutoff_gev=0.15, **size_kw):
    """Initial conditions of k lumpy events: event j has seed 0xCCA4E001 + j
    and the square half-width L_j = spacing / 2 (sqrt(N_j) - 1) that gives
    N_j lattice points for the lognormal N_j (0.01 (sqrt(N_j) - 1) on the
    0.02 fm lattice); points at or below the energy cutoff are dropped."""
    sizes = lognormal_sizes(k, seed=seed, **size_kw)
    ics = []
    for j, nj in enumerate(sizes):
        half = round(spacing / 2.0 * (np.sqrt(nj) - 1.0) / spacing) * spacing
        ics.append(ic.lumpy(seed=0xCCA4E001 + j, spacing=spacing, half_width=half, e_cutoff_gev=e_cutoff_gev))
    return ics


def make_batch(dev, eos, cfg, ics):
    b = Batch(dev, eos, cfg, [len(d["x"]) for d in ics])
    load_event(b, concat_ics(ics))
    return b


def autotune(dev, eos, cfg, ics, cells_per_rc=(2, 3, 4), slices=(4, 8, 12, 16), steps=2):
    """Cell size and work per wave for a batch: for each cells-per-radius
    candidate a batch of the events is initialized and stepped twice, then
    every slice count (pair-kernel waves per tile) is timed over `steps`
    steps; the fastest pair is returned with the table of timings
    (ms per step).  The tuning batches are discarded."""
    table = {}
    best = None
    for cpr in cells_per_rc:
        b = make_batch(dev, eos, dict(cfg, cells_per_rc=cpr), ics)
        b.initialize()
        b.run_steps(cfg["dt"], 2)
        cap = b.nslices()
        for s in slices:
            if s > cap:
                continue
            b.nslices(s)
            b.run_steps(cfg["dt"], 1)
            dev.sync()
            t0 = time.perf_counter()
            b.run_steps(cfg["dt"], steps)
            dev.sync()
            ms = (time.perf_counter() - t0) / steps * 1e3
            table[f"{cpr},{s}"] = round(ms, 3)
            if best is None or ms < best[2]:
                best = (cpr, s, ms)
        b.close()
    return dict(cells_per_rc=best[0], slices=best[1], ms_per_step=round(best[2], 3), table=table)


def run_batch(dev, eos, cfg, ics, slices=None, max_tau=30.0, progress=None, every=50):
    """The events in one batch until every event has frozen out.  Returns
    the wall time, each event's completion time (the step after which its
    continue flag turned false) and the batch's step count.  `progress`
    (optional) is called as progress(steps, events_done) every `every`
    batch steps."""
    b = make_batch(dev, eos, cfg, ics)
    if slices:
        b.nslices(slices)
    dev.sync()
    t0 = time.perf_counter()
    b.initialize()
    done = [None] * len(ics)
    steps = 0
    while True:
        keep = b.batch_continue_evolution(max_tau=max_tau)
        now = time.perf_counter() - t0
        for e in range(len(ics)):
            if not keep[e] and done[e] is None:
                done[e] = now
        if not keep.any():
            break
        b.advance_timestep(cfg["dt"], 2)
        steps += 1
        if progress is not None and steps % every == 0:
            progress(steps, sum(t is not None for t in done))
    wall = time.perf_counter() - t0
    rows = b.surface_by_event()
    b.close()
    return dict(wall_s=wall, done_s=done, steps=steps, rows=rows)


def run_alone(dev, eos, cfg, ics, max_tau=30.0):
    """The plain path: each event alone, one after the other, until it has
    frozen out.  Completion times are cumulative (every event arrived at
    time 0)."""
    from . import Event
    done = []
    rows = []
    steps = []
    t0 = time.perf_counter()
    for d in ics:
        ev = Event(dev, eos, cfg, len(d["x"]))
        load_event(ev, d)
        ev.initialize()
        k = 0
        while ev.continue_evolution(max_tau=max_tau):
            ev.advance_timestep(cfg["dt"], 2)
            k += 1
        dev.sync()
        done.append(time.perf_counter() - t0)
        rows.append(ev.freeze_out_surface())
        steps.append(k)
        ev.close()
    return dict(wall_s=time.perf_counter() - t0, done_s=done, steps=steps, rows=rows)


def latency_summary(run, n_particles):
    done = np.asarray(run["done_s"])
    k = len(done)
    return dict(events=k, events_per_s=round(k / run["wall_s"], 4), wall_s=round(run["wall_s"], 3),
                p50_ms=round(float(np.percentile(done, 50)) * 1e3, 1),
                p99_ms=round(float(np.percentile(done, 99)) * 1e3, 1),
                particles=int(n_particles))
