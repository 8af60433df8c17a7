"""Continuous-batching request loop for event-by-event hydro runs.

Requests are initial conditions of uneven particle counts that arrive over
time.  Up to `slots` events are resident on the device at once, each on its
own HIP stream, so their kernels overlap; a request is admitted as soon as a
slot is free and the device's HBM budget has room for it (event creation is
refused beyond the budget, and the request then waits for a running event to
release its memory).  An admitted event stays resident until its end
condition -- every particle frozen out or tau past max_tau
(SPHWorkstation::continue_evolution, src/sph_workstation.cpp:1992-2024), or
a step cap -- and then its fields are downloaded and its memory released.
Latency is measured from a request's arrival to its downloaded result.

The reference runs one event per process and has no request loop; this
module is the serving layer on top of the drop-in path.  Each event's
result does not depend on what else runs beside it: the events share no
device state, and every kernel is deterministic.
"""
from __future__ import annotations

import dataclasses
import threading
import time

import numpy as np

import ccake_amd


@dataclasses.dataclass
class Request:
    rid: int
    ic: dict                     # x, y, e, ux, uy, rhoB, rhoS, rhoQ (load_event)
    params: dict                 # ccake_params fields (config.*_params)
    arrival: float = 0.0         # seconds after the start of the run
    max_steps: int = 0           # step cap; 0: none (run to freeze-out or max_tau)
    max_tau: float = 50.0
    fields: tuple = ("CCAKE_F_T", "CCAKE_F_UX", "CCAKE_F_UY", "CCAKE_F_E")


@dataclasses.dataclass
class Result:
    rid: int
    n: int
    steps: int
    t: float
    fields: dict
    arrival: float               # seconds after the start of the run
    admitted: float
    finished: float

    @property
    def latency(self):
        return self.finished - self.arrival


def run_event(dev, eos, req):
    """One request on one event: create, load, initialize, evolve, download.
    Returns (steps, t, fields)."""
    n = len(req.ic["x"])
    ev = ccake_amd.Event(dev, eos, req.params, n)
    try:
        ccake_amd.load_event(ev, req.ic)
        ev.initialize()
        steps = 0
        dt = req.params["dt"]
        if req.params.get("particlization", 0):
            while ev.continue_evolution(req.max_tau) and (req.max_steps <= 0 or steps < req.max_steps):
                ev.advance_timestep(dt, 2)
                steps += 1
        else:
            cap = req.max_steps
            if cap <= 0:
                # tau grid of the event: t0 + k dt up to max_tau
                cap = int(np.floor((req.max_tau - ev.t) / dt + 1e-9))
            steps = ev.run_steps(dt, cap)
        out = {f: ev.download(f) for f in req.fields}
        return steps, ev.t, out
    finally:
        ev.close()


class Server:
    """Admits requests into at most `slots` resident events (HBM budget
    permitting) and runs them concurrently on the device."""

    def __init__(self, dev, eos, slots=4):
        if slots < 1:
            raise ValueError("slots must be >= 1")
        self.dev, self.eos, self.slots = dev, eos, int(slots)

    def run(self, requests):
        """Serves `requests` (arrival times relative to now) and returns the
        results ordered by request id.  A request that cannot fit the budget
        even on an idle device raises."""
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        lock = threading.Condition()
        state = dict(next=0, live=0, error=None)
        results = {}
        t_start = time.perf_counter()

        def now():
            return time.perf_counter() - t_start

        def worker():
            while True:
                with lock:
                    while True:
                        if state["error"] is not None or state["next"] >= len(pending):
                            return
                        req = pending[state["next"]]
                        wait = req.arrival - now()
                        if wait <= 0:
                            state["next"] += 1
                            state["live"] += 1
                            break
                        lock.wait(timeout=wait)
                admitted = None
                try:
                    while True:
                        try:
                            admitted = now()
                            steps, t, out = run_event(self.dev, self.eos, req)
                            break
                        except ccake_amd.CcakeError as e:
                            if ccake_amd.STATUS.get(e.status) != "CCAKE_ERR_OOM":
                                raise
                            with lock:
                                # over the budget: wait for another event to finish
                                if state["live"] <= 1:
                                    raise
                                lock.wait(timeout=0.05)
                    res = Result(req.rid, len(req.ic["x"]), steps, t, out, req.arrival, admitted, now())
                    with lock:
                        results[req.rid] = res
                except BaseException as e:   # noqa: BLE001 -- handed to the caller
                    with lock:
                        if state["error"] is None:
                            state["error"] = e
                finally:
                    with lock:
                        state["live"] -= 1
                        lock.notify_all()

        threads = [threading.Thread(target=worker, daemon=True) for _ in range(self.slots)]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        if state["error"] is not None:
            raise state["error"]
        return [results[k] for k in sorted(results)]


def poisson_arrivals(k, rate, seed=0x50155):
    """Arrival times (seconds from the start) of k requests of a Poisson
    stream of `rate` events per second: cumulative seeded exponential gaps."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=k))


@dataclasses.dataclass
class SurfaceResult:
    rid: int
    n: int
    steps: int                   # batch steps until the event's continue flag turned false
    rows: np.ndarray             # the event's freeze-out hypersurface (its own particle ids)
    arrival: float
    admitted: float
    finished: float
    batch: int                   # which batch (in order of formation) ran the event

    @property
    def latency(self):
        return self.finished - self.arrival


class BatchServer:
    """Continuous batching of run-to-freeze-out requests.

    `slots` workers each run one batch (ccake_batch_create) at a time on the
    batch's own stream, so up to `slots` batches overlap on the device.  A
    free worker takes every request that has arrived by then -- at most
    `max_events` of them and `max_particles` particles, and as many as the
    device's HBM budget admits -- into a new batch, initialises it and steps
    it, with each event stopping when SPHWorkstation::continue_evolution
    (src/sph_workstation.cpp:1992-2024) does: the step after which all its
    particles are frozen out (or tau passes max_tau).  An event's
    hypersurface is complete at that step and is returned right away, while
    the batch goes on with its other events; requests that arrive meanwhile
    wait for the next free worker, which waits up to `window` seconds after
    the first of them arrived for more.  Every request shares `params` (the
    workload's settings, particlization on).  An event's rows equal the same
    event's run alone to rounding (test_serve.py)."""

    def __init__(self, dev, eos, params, slots=2, max_events=64, max_particles=4_000_000, max_tau=30.0,
                 slices=None, window=0.0, on_result=None):
        if slots < 1 or max_events < 1:
            raise ValueError("slots and max_events must be >= 1")
        if not params.get("particlization", 0):
            raise ValueError("run-to-freeze-out requests need particlization on")
        self.dev, self.eos, self.params = dev, eos, params
        self.slots, self.max_events, self.max_particles = int(slots), int(max_events), int(max_particles)
        self.max_tau = float(max_tau)
        self.slices = slices
        # on_result(result): called (under the server's lock) as each event's
        # result is ready
        self.on_result = on_result
        # a free worker waits up to `window` seconds after the first waiting
        # request's arrival for more arrivals before it forms the batch
        self.window = float(window)

    def _run_batch(self, reqs, admitted, now, bid, publish):
        ics = [r.ic for r in reqs]
        b = ccake_amd.Batch(self.dev, self.eos, self.params, [len(d["x"]) for d in ics])
        try:
            if self.slices:
                b.nslices(min(self.slices, b.nslices()))    # at most the count the batch was sized for
            ccake_amd.load_event(b, ccake_amd.concat_ics(ics))
            b.initialize()
            done = [False] * len(reqs)
            steps = 0
            dt = self.params["dt"]
            while True:
                keep = b.batch_continue_evolution(max_tau=self.max_tau)
                newly = [e for e in range(len(reqs)) if not keep[e] and not done[e]]
                if newly:
                    per_event = b.surface_by_event()
                    t_now = now()
                    for e in newly:
                        done[e] = True
                        publish(SurfaceResult(reqs[e].rid, len(ics[e]["x"]), steps, per_event[e], reqs[e].arrival,
                                              admitted, t_now, bid))
                if not keep.any():
                    break
                b.advance_timestep(dt, 2)
                steps += 1
        finally:
            b.close()

    def run(self, requests):
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        lock = threading.Condition()
        state = dict(next=0, error=None, batches=0, live=0)
        results = {}
        t_start = time.perf_counter()

        def now():
            return time.perf_counter() - t_start

        def publish(res):
            with lock:
                results[res.rid] = res
                if self.on_result is not None:
                    self.on_result(res)

        def take():
            """The requests that have arrived, up to the batch limits; waits
            for the first arrival; None when every request is taken."""
            with lock:
                while True:
                    if state["error"] is not None or state["next"] >= len(pending):
                        return None
                    first = pending[state["next"]].arrival
                    last = pending[min(state["next"] + self.max_events, len(pending)) - 1].arrival
                    wait = min(first + self.window, last) - now()
                    if wait <= 0:
                        break
                    lock.wait(timeout=wait)
                out, npart = [], 0
                t = now()
                while (state["next"] < len(pending) and len(out) < self.max_events
                       and pending[state["next"]].arrival <= t):
                    r = pending[state["next"]]
                    if out and npart + len(r.ic["x"]) > self.max_particles:
                        break
                    out.append(r)
                    npart += len(r.ic["x"])
                    state["next"] += 1
                state["batches"] += 1
                state["live"] += 1
                return out, state["batches"] - 1

        def give_back(reqs):
            """Return the tail of a batch the HBM budget refused to the queue."""
            with lock:
                state["next"] -= len(reqs)
                for k, r in enumerate(reqs):
                    pending[state["next"] + k] = r

        def worker():
            while True:
                got = take()
                if got is None:
                    return
                reqs, bid = got
                try:
                    while True:
                        try:
                            self._run_batch(reqs, now(), now, bid, publish)
                            break
                        except ccake_amd.CcakeError as e:
                            if ccake_amd.STATUS.get(e.status) != "CCAKE_ERR_OOM":
                                raise
                            with lock:
                                others = state["live"] > 1
                            if len(reqs) > 1:          # halve the batch, queue the rest
                                half = len(reqs) // 2
                                give_back(reqs[half:])
                                reqs = reqs[:half]
                            elif others:               # wait for another batch to free memory
                                with lock:
                                    lock.wait(timeout=0.05)
                            else:
                                raise
                except BaseException as e:   # noqa: BLE001 -- handed to the caller
                    with lock:
                        if state["error"] is None:
                            state["error"] = e
                finally:
                    with lock:
                        state["live"] -= 1
                        lock.notify_all()

        threads = [threading.Thread(target=worker, daemon=True) for _ in range(self.slots)]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        if state["error"] is not None:
            raise state["error"]
        return [results[k] for k in sorted(results)]


def summarize(results, wall):
    """events/s, particle-steps/s and latency percentiles of a served run."""
    lat = np.array([r.latency for r in results]) * 1e3
    ps = sum(r.n * r.steps for r in results)
    return dict(events=len(results), events_per_s=len(results) / wall, particle_steps_per_s=ps / wall,
                p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                wall_s=wall)
