"""Batches (ccake_batch_create): many events stepped together, each kernel
launched once per stage for all of them.  Each event must evolve as it does
alone: the same pair sets (btrack equal), fields equal to rounding (the
grouping of an event's pair sums into slices depends on the batch's tile
count), the same freeze-out rows.  Tolerances are written in each test."""
import numpy as np
import pytest

import ccake_amd
from ccake_amd import config, eos_tables, ic

T_FO = 150.0 / config.HBARC_MEVFM
F = ccake_amd.FIELDS
FIELDS = ["X", "Y", "UX", "UY", "SIGMA_LAB", "SIGMA", "S", "E", "T", "RHOB", "RHOS", "RHOQ", "BULK", "EXT_S",
          "EXT_B", "GRADP0", "GRADE0", "GRADV0", "DUDT0", "SHV0+5"]


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


def three_events(spacing):
    return [ic.lumpy(seed=0xCCA4E001 + k, spacing=spacing, half_width=hw, e_cutoff_gev=0.15)
            for k, hw in enumerate((5.0, 3.0, 4.0))]


def make(dev, eos, cfg, ics):
    b = ccake_amd.Batch(dev, eos, cfg, [len(d["x"]) for d in ics])
    ccake_amd.load_event(b, ccake_amd.concat_ics(ics))
    b.initialize()
    alone = []
    for d in ics:
        ev = ccake_amd.Event(dev, eos, cfg, len(d["x"]))
        ccake_amd.load_event(ev, d)
        ev.initialize()
        alone.append(ev)
    return b, alone


def whole(a, b):
    return float(np.max(np.abs(a - b)) / (np.max(np.abs(b)) or 1.0))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_batch_equals_alone(dev, table_eos, precision):
    """Three lumpy viscous BSQ events of 8,077, 3,345 and 5,746 particles on
    the table EoS (0.1 fm lattice, different seeds and widths): after
    initialization and 20 steps in one batch every event's fields are within
    1e-12 (whole-field relative; FP64 pair sums; measured 4e-14) or 1e-6
    (mixed precision: the slices' FP32 blocks begin at other candidates, so
    the batch and the run alone differ by FP32 rounding, 2.5e-8 measured) of
    the same event run alone, with equal btrack and per-event entropy."""
    cfg = dict(config.lumpy_params(), pair_precision=precision)
    ics = three_events(0.1)
    b, alone = make(dev, table_eos, cfg, ics)
    assert b.n_events == 3 and list(np.diff(b.offsets)) == [len(d["x"]) for d in ics]
    b.run_steps(cfg["dt"], 20)
    for ev in alone:
        ev.run_steps(cfg["dt"], 20)
    tol = 1e-12 if precision == 0 else 1e-6
    s_batch = b.batch_reduce("CCAKE_R_ENTROPY")
    worst = 0.0
    for e, ev in enumerate(alone):
        sl = b.event_slice(e)
        assert np.array_equal(b.download_int("CCAKE_I_BTRACK")[sl], ev.download_int("CCAKE_I_BTRACK")), e
        assert np.array_equal(b.download_int("CCAKE_I_EOS_ID")[sl], ev.download_int("CCAKE_I_EOS_ID")), e
        for lab in FIELDS:
            err = whole(b.download(fid(lab))[sl], ev.download(fid(lab)))
            worst = max(worst, err)
            assert err <= tol, (e, lab, err)
        assert abs(s_batch[e] / ev.reduce("CCAKE_R_ENTROPY") - 1.0) < (1e-13 if precision == 0 else 1e-9)
    print(f"precision {precision}: worst field difference batch vs alone {worst:.2e}", flush=True)
    b.close()
    for ev in alone:
        ev.close()


@pytest.mark.gpu
def test_batch_runs_every_event_to_freeze_out(dev, table_eos):
    """The three events with particlization on, run in one batch until every
    event has frozen out (continue_evolution of the batch: true while any
    event continues): each event's freeze-out rows have the ids and steps of
    its run alone and their columns are within 1e-9 of them (whole-column
    relative); the per-event continue flags turn false at the step the run
    alone ends.  An event ends in the batch when its flag turns false (the
    later batch steps leave it as it is), so its final fields equal those
    of its run alone, which ended at that step, within 1e-9."""
    cfg = dict(config.lumpy_params(), particlization=1, e_freeze=table_eos.efreeze(T_FO))
    ics = three_events(0.1)
    b, alone = make(dev, table_eos, cfg, ics)
    steps_alone = []
    for ev in alone:
        k = 0
        while ev.continue_evolution(max_tau=30.0):
            ev.advance_timestep(cfg["dt"], 2)
            k += 1
        steps_alone.append(k)
    k = 0
    ended = [None] * 3
    while b.continue_evolution(max_tau=30.0):
        keep = b.batch_continue_evolution(max_tau=30.0)
        for e in range(3):
            if not keep[e] and ended[e] is None:
                ended[e] = k
        b.advance_timestep(cfg["dt"], 2)
        k += 1
    for e in range(3):
        if ended[e] is None:
            ended[e] = k
    print("steps alone", steps_alone, "batch: events ended at", ended, "batch steps", k, flush=True)
    assert ended == steps_alone and k == max(steps_alone)
    rows = b.surface_by_event()
    for e, ev in enumerate(alone):
        ra = ev.freeze_out_surface()
        rb = rows[e]
        assert rb.shape == ra.shape and np.array_equal(rb[:, :2], ra[:, :2]), e
        for c in range(2, ra.shape[1]):
            assert whole(rb[:, c], ra[:, c]) <= 1e-9, (e, c, whole(rb[:, c], ra[:, c]))
        sl = b.event_slice(e)
        for lab in ("X", "Y", "UX", "UY", "SIGMA", "E", "T", "RHOB", "EXT_S", "BULK"):
            f = ccake_amd.FIELDS["CCAKE_F_" + lab]
            assert whole(b.download(f)[sl], ev.download(f)) <= 1e-9, (e, lab)
    b.close()
    for ev in alone:
        ev.close()


@pytest.mark.gpu
def test_batch_refuses_bad_input(dev, table_eos):
    """An event without particles or more than 1024 events are refused."""
    cfg = config.lumpy_params()
    with pytest.raises(ccake_amd.CcakeError):
        ccake_amd.Batch(dev, table_eos, cfg, [10, 0, 5])
    with pytest.raises(ccake_amd.CcakeError):
        ccake_amd.Batch(dev, table_eos, cfg, [1] * 1025)
