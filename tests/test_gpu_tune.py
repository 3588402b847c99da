"""GPU tests of the tile-order autotuner (glasma_tune_step, glasma_batch_tune): tuning leaves the fields and
the clock untouched, picks one of the orders that fit the lattice, and a tuned context evolves bit-identically
to an untuned one (the order changes the time of a step, never its result)."""
import numpy as np
import pytest

from tests.helpers import random_fields, upload

pytestmark = pytest.mark.gpu

KEYS = ("U", "E", "Jp", "Jm", "JpS", "JmS")


@pytest.mark.parametrize("N", [(256, 32, 4), (128, 64, 3), (48, 8, 5)])
def test_tune_keeps_fields_and_results(N):
    from soerenpragya3dglasma_amd import Glasma
    f = random_fields(N, seed=5)
    with Glasma(N, (1.0, 1.0, 1.0), 0.25) as g, Glasma(N, (1.0, 1.0, 1.0), 0.25) as ref:
        upload(g, f)
        upload(ref, f)
        before, t0 = g.get_fields(), g.time()
        res = g.tune_step(reps=1)
        after = g.get_fields()
        for k in KEYS:
            assert np.array_equal(before[k], after[k]), k
        assert g.time() == t0
        assert res["chosen"] in res["candidates_ms"]
        assert all(ms > 0 for ms in res["candidates_ms"].values())
        order, tuned = g.tile_order()
        assert tuned and order == res["chosen"]
        if N[0] % 64 == 0:  # column bands fit: several orders are timed
            assert any(k.startswith("bands") for k in res["candidates_ms"])
        g.Update(3)
        ref.Update(3)
        a, b = g.get_fields(), ref.get_fields()
        for k in KEYS:
            assert np.array_equal(a[k], b[k]), k


def test_batch_tune_keeps_results():
    from soerenpragya3dglasma_amd import Glasma, GlasmaBatch
    specs = [((256, 32, 4), (1.0, 1.0, 1.0), 0.25), ((16, 8, 4), (1.0, 1.0, 1.0), 0.2),
             ((256, 32, 4), (1.0, 1.0, 1.0), 0.25), ((48, 8, 5), (1.0, 0.9, 0.3), 0.1)]
    inputs = [random_fields(sp[0], seed=70 + i) for i, sp in enumerate(specs)]
    with GlasmaBatch(specs) as b:
        for ev, f in zip(b.events, inputs):
            upload(ev, f)
        b.update(2)
        b.tune(reps=1)
        assert b.events[0].tile_order() == b.events[2].tile_order()  # same lattice, same order
        b.update(3)
        got = [ev.get_fields() for ev in b.events]
    for sp, f, d in zip(specs, inputs, got):
        with Glasma(*sp) as g:
            upload(g, f)
            g.Update(5)
            ref = g.get_fields()
        for k in KEYS:
            assert np.array_equal(d[k], ref[k]), (sp, k)


def test_copy_ceiling_leaves_fields_unchanged():
    """glasma_copy_ceiling times a copy of the step's compulsory traffic into scratch; the fields stay as they were."""
    from soerenpragya3dglasma_amd import Glasma
    N = (64, 8, 6)
    with Glasma(N, (1.0, 1.0, 1.0), 0.25) as g:
        upload(g, random_fields(N, seed=9))
        before = g.checksum()
        ms = g.copy_ceiling(3)
        assert ms > 0.0
        assert g.checksum() == before
