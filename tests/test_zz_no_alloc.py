"""The allocation-free entry point ccake_export_pairs_noalloc (no
counterpart in the reference; for callers that cannot allocate): through
the C-ABI, its pair set equals ccake_export_pairs' on two calls in a row,
the second on a smaller event, and sizes beyond its static scratch array
are refused with a clear error."""
import numpy as np
import pytest

import ccake_amd
from ccake_amd import config, ic

pytestmark = pytest.mark.gpu


def lattice_event(dev, eos, spacing):
    cfg = config.gubser_ideal_params()
    d = ic.gubser_ideal(spacing=spacing)
    ev = ccake_amd.Event(dev, eos, cfg, len(d["x"]))
    ccake_amd.load_event(ev, d)
    ev.build_neighbors()
    return ev


def test_noalloc_export_equals_export_twice_smaller_second(dev):
    eos = ccake_amd.Eos(dev, 0)
    for spacing in (0.1, 0.2):                    # 10,201 then 2,601 particles
        ev = lattice_event(dev, eos, spacing)
        gi, gj = ev.pairs()
        ni, nj = ev.pairs_noalloc()
        print(f"spacing {spacing}: {len(gi)} pairs", flush=True)
        assert len(gi) > 0 and np.array_equal(gi, ni) and np.array_equal(gj, nj)
        ev.close()
    eos.close()


def test_noalloc_export_refuses_more_pairs_than_its_scratch(dev):
    """6,000 particles inside one cutoff radius: 35,994,000 directed pairs,
    more than the 15,204,320 the scratch array holds."""
    eos = ccake_amd.Eos(dev, 0)
    rng = np.random.default_rng(3)
    n = 6000
    x, y = rng.uniform(-0.2, 0.2, n), rng.uniform(-0.2, 0.2, n)
    ev = ccake_amd.Event(dev, eos, config.base_params(hT=0.3), n)
    ev.upload("CCAKE_F_X", x); ev.upload("CCAKE_F_Y", y)
    ev.build_neighbors()
    with pytest.raises(ccake_amd.CcakeError) as ex:
        ev.pairs_noalloc(cap=40_000_000)
    assert ex.value.status == ccake_amd.STATUS_CAPACITY and "scratch" in str(ex.value)
    ev.close(); eos.close()
