"""Request loop: admission by slots and HBM budget, arrival order, results
identical alone and in the mix."""
import threading
import time

import numpy as np
import pytest

import ccake_amd
from ccake_amd import config, ic, serve


def gubser_request(rid, half_width, spacing=0.1, steps=10, arrival=0.0):
    d = ic.gubser_ideal(spacing=spacing, half_width=half_width)
    return serve.Request(rid=rid, ic=d, params=config.gubser_ideal_params(), arrival=arrival, max_steps=steps,
                         fields=("CCAKE_F_E", "CCAKE_F_UX", "CCAKE_F_UY", "CCAKE_F_X", "CCAKE_F_S"))


# ------------------------------------------------------------------ CPU: admission logic
class FakeBudget:
    """Stands in for run_event: each request holds `n` units of a budget
    while it runs and is refused (CCAKE_ERR_OOM) when they do not fit."""

    OOM = {v: k for k, v in ccake_amd.STATUS.items()}["CCAKE_ERR_OOM"]

    def __init__(self, budget):
        self.budget, self.used, self.peak, self.active, self.peak_active = budget, 0, 0, 0, 0
        self.lock = threading.Lock()
        self.order = []

    def __call__(self, dev, eos, req):
        n = len(req.ic["x"])
        with self.lock:
            if self.used + n > self.budget:
                raise ccake_amd.CcakeError(self.OOM, "ccake_event_create")
            self.used += n
            self.active += 1
            self.peak = max(self.peak, self.used)
            self.peak_active = max(self.peak_active, self.active)
            self.order.append(req.rid)
        time.sleep(0.02)
        with self.lock:
            self.used -= n
            self.active -= 1
        return req.max_steps, 1.0, {"n": np.full(3, n)}


def fake_requests(sizes, arrivals=None):
    arrivals = arrivals or [0.0] * len(sizes)
    return [serve.Request(rid=k, ic={"x": np.zeros(s)}, params={}, arrival=a, max_steps=k)
            for k, (s, a) in enumerate(zip(sizes, arrivals))]


def test_admission_respects_slots_and_budget(monkeypatch):
    fake = FakeBudget(budget=10)
    monkeypatch.setattr(serve, "run_event", fake)
    reqs = fake_requests([4, 4, 4, 4, 3, 6, 2, 5])
    res = serve.Server(None, None, slots=3).run(reqs)
    assert [r.rid for r in res] == list(range(8))
    assert all(r.steps == r.rid for r in res)
    assert fake.peak <= 10 and fake.peak_active <= 3
    assert fake.peak_active >= 2            # work overlapped
    assert all(r.finished >= r.admitted >= r.arrival for r in res)


def test_arrivals_are_not_admitted_early(monkeypatch):
    fake = FakeBudget(budget=100)
    monkeypatch.setattr(serve, "run_event", fake)
    res = serve.Server(None, None, slots=4).run(fake_requests([1, 1, 1], arrivals=[0.0, 0.08, 0.04]))
    for r in res:
        assert r.admitted >= r.arrival - 1e-3
    assert fake.order == [0, 2, 1]


def test_request_larger_than_budget_raises(monkeypatch):
    fake = FakeBudget(budget=5)
    monkeypatch.setattr(serve, "run_event", fake)
    with pytest.raises(ccake_amd.CcakeError):
        serve.Server(None, None, slots=2).run(fake_requests([2, 9]))


def test_summary_percentiles():
    rs = [serve.Result(k, 100, 10, 1.0, {}, 0.0, 0.0, 0.001 * (k + 1)) for k in range(100)]
    s = serve.summarize(rs, wall=0.5)
    assert s["events"] == 100 and abs(s["events_per_s"] - 200) < 1e-9
    assert abs(s["p50_ms"] - 50.5) < 1e-9 and abs(s["p99_ms"] - 99.01) < 1e-9
    assert s["particle_steps_per_s"] == 100 * 100 * 10 / 0.5


# ------------------------------------------------------------------ GPU: alone vs in the mix
@pytest.mark.gpu
def test_served_results_equal_alone(dev):
    """Five Gubser events of uneven size (1,681 to 2,601 particles), ten
    steps each: served four at a time with staggered arrivals, every field
    equals (to the bit) the same event run alone."""
    eos = ccake_amd.Eos(dev, 0)
    widths = [2.0, 2.5, 1.5, 2.2, 1.8]
    reqs = [gubser_request(k, w, arrival=0.002 * k) for k, w in enumerate(widths)]
    alone = [serve.run_event(dev, eos, r) for r in reqs]
    mixed = serve.Server(dev, eos, slots=4).run(reqs)
    for (steps, t, out), r in zip(alone, mixed):
        assert r.steps == steps == 10 and r.t == t
        for f, a in out.items():
            assert np.array_equal(a, r.fields[f]), (r.rid, f)
    eos.close()


@pytest.mark.gpu
def test_served_under_hbm_budget():
    """A budget that holds two of the events at a time: the loop still
    serves all of them, never holds more than the budget, and the results
    equal the unconstrained ones."""
    probe = ccake_amd.Device(0)
    eos = ccake_amd.Eos(probe, 0)
    reqs = [gubser_request(k, w, steps=5) for k, w in enumerate([2.0, 2.0, 2.0, 2.0])]
    base_used = probe.memory()[0]
    ev = ccake_amd.Event(probe, eos, reqs[0].params, len(reqs[0].ic["x"]))
    per_event = probe.memory()[0] - base_used
    ev.close()
    ref = [serve.run_event(probe, eos, r) for r in reqs]
    eos.close()
    probe.close()

    dev = ccake_amd.Device(0, hbm_budget=int(2.5 * per_event))
    eos = ccake_amd.Eos(dev, 0)
    res = serve.Server(dev, eos, slots=4).run(reqs)
    assert len(res) == 4
    for (steps, t, out), r in zip(ref, res):
        for f, a in out.items():
            assert np.array_equal(a, r.fields[f]), (r.rid, f)
    assert dev.memory()[0] == 0
    eos.close()
    dev.close()


# ------------------------------------------------------------------ continuous batching
def test_poisson_arrivals_rate_and_order():
    t = serve.poisson_arrivals(20000, 50.0, seed=3)
    assert np.all(np.diff(t) > 0)
    assert abs(20000 / t[-1] - 50.0) < 1.5          # the mean rate, within a few standard errors
    assert np.array_equal(t[:5], serve.poisson_arrivals(5, 50.0, seed=3))


def test_batch_server_forms_batches_from_arrived_requests(monkeypatch):
    """Stand-in batches: a free worker takes every request that has arrived
    (within max_events), later arrivals go to the next batch, and no request
    is taken before it arrives."""
    taken = []

    def fake_run_batch(self, reqs, admitted, now, bid, publish):
        taken.append((bid, [r.rid for r in reqs], admitted, [r.arrival for r in reqs]))
        time.sleep(0.05)
        for r in reqs:
            publish(serve.SurfaceResult(r.rid, 1, 3, np.zeros((0, 42)), r.arrival, admitted, now(), bid))

    monkeypatch.setattr(serve.BatchServer, "_run_batch", fake_run_batch)
    arr = [0.0, 0.001, 0.002, 0.08, 0.081, 0.2]
    reqs = [serve.Request(rid=k, ic={"x": np.zeros(10)}, params={}, arrival=a) for k, a in enumerate(arr)]
    srv = serve.BatchServer(None, None, dict(particlization=1, dt=0.02), slots=1, max_events=2)
    res = srv.run(reqs)
    assert [r.rid for r in res] == list(range(6))
    sizes = [len(t[1]) for t in taken]
    assert max(sizes) <= 2 and sum(sizes) == 6
    for bid, rids, admitted, arrivals in taken:
        assert all(a <= admitted + 1e-3 for a in arrivals)
    assert all(r.latency >= 0 for r in res)


@pytest.mark.gpu
def test_batch_server_results_equal_alone():
    """Three lumpy events (0.1 fm lattice, 8,077, 3,345 and 5,746 particles)
    arriving as a Poisson stream, served by two workers forming batches of
    at most two events and run to freeze-out: every event's hypersurface has
    the ids and steps of its run alone and columns within 1e-9 of them."""
    from ccake_amd import eos_tables
    t_fo = 150.0 / config.HBARC_MEVFM
    dev = ccake_amd.Device(0)
    tab, dims, mins, maxs = eos_tables.tanh_conformal_table()
    eos = ccake_amd.Eos(dev, config.EOS_TABLE, t_fo, dims, mins, maxs, tab)
    cfg = dict(config.lumpy_params(), particlization=1, e_freeze=eos.efreeze(t_fo))
    ics = [ic.lumpy(seed=0xCCA4E001 + k, spacing=0.1, half_width=hw, e_cutoff_gev=0.15)
           for k, hw in enumerate((5.0, 3.0, 4.0))]
    arr = serve.poisson_arrivals(3, 20.0, seed=11)
    reqs = [serve.Request(rid=k, ic=d, params=cfg, arrival=float(a)) for k, (d, a) in enumerate(zip(ics, arr))]
    res = serve.BatchServer(dev, eos, cfg, slots=2, max_events=2).run(reqs)
    assert [r.rid for r in res] == [0, 1, 2]
    for r, d in zip(res, ics):
        ev = ccake_amd.Event(dev, eos, cfg, len(d["x"]))
        ccake_amd.load_event(ev, d)
        ev.initialize()
        k = 0
        while ev.continue_evolution(max_tau=30.0):
            ev.advance_timestep(cfg["dt"], 2)
            k += 1
        ra = ev.freeze_out_surface()
        ev.close()
        assert r.steps == k, (r.rid, r.steps, k)
        assert r.rows.shape == ra.shape and np.array_equal(r.rows[:, :2], ra[:, :2]), r.rid
        for c in range(2, ra.shape[1]):
            s = float(np.max(np.abs(ra[:, c]))) or 1.0
            assert float(np.max(np.abs(r.rows[:, c] - ra[:, c]))) / s <= 1e-9, (r.rid, c)
        assert r.finished >= r.admitted >= r.arrival - 1e-3
    eos.close()
    dev.close()
