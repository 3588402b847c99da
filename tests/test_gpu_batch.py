"""GPU tests of the event ensemble (glasma_batch_*): every event of a batched run is bit-identical to the
same event advanced alone; graph replays agree with eager runs, with eager work between replays."""
import numpy as np
import pytest

from tests.helpers import random_fields, upload

pytestmark = pytest.mark.gpu

SPECS = [((8, 8, 16), (1.0, 1.0, 1.0), 0.25), ((6, 5, 7), (1.0, 0.9, 0.3), 0.1), ((16, 8, 4), (1.0, 1.0, 1.0), 0.2),
         ((12, 12, 6), (0.8, 0.8, 0.5), 0.1, 1), ((64, 8, 5), (1.0, 1.0, 1.0), 0.25)]
KEYS = ("U", "E", "Jp", "Jm", "JpS", "JmS")


def _alone(spec, f, nsteps):
    from soerenpragya3dglasma_amd import Glasma
    with Glasma(*spec) as g:
        upload(g, f)
        g.Update(nsteps)
        return g.get_fields(), g.get_uold(), g.time()


def test_batched_events_bit_identical_to_single_runs():
    from soerenpragya3dglasma_amd import GlasmaBatch
    inputs = [random_fields(sp[0], seed=20 + i, bc=sp[3] if len(sp) > 3 else 0) for i, sp in enumerate(SPECS)]
    with GlasmaBatch(SPECS) as b:
        for ev, f in zip(b.events, inputs):
            upload(ev, f)
        b.update(7)
        got = [(ev.get_fields(), ev.get_uold(), ev.time()) for ev in b.events]
    for sp, f, (d, uold, t) in zip(SPECS, inputs, got):
        ref, ref_uold, ref_t = _alone(sp, f, 7)
        for k in KEYS:
            assert np.array_equal(d[k], ref[k]), (sp, k, np.abs(d[k] - ref[k]).max())
        assert np.array_equal(uold, ref_uold)
        assert t == ref_t


def test_graph_replays_with_eager_work_between():
    from soerenpragya3dglasma_amd import GlasmaBatch
    inputs = [random_fields(sp[0], seed=40 + i, bc=sp[3] if len(sp) > 3 else 0) for i, sp in enumerate(SPECS)]
    refs = [[_alone(sp, f, n)[0] for n in (3, 6, 9)] for sp, f in zip(SPECS, inputs)]
    with GlasmaBatch(SPECS) as b:
        for ev, f in zip(b.events, inputs):
            upload(ev, f)
        for r in range(3):
            b.update(3)  # the same graph, replayed
            for e, ev in enumerate(b.events):
                ev.bulk_raw()  # eager device work between replays
                d = ev.get_fields()
                for k in KEYS:
                    assert np.array_equal(d[k], refs[e][r][k]), (r, e, k)


def test_events_advanced_alone_in_between():
    """An odd single-event update flips that event's buffer parity; the next batched call still agrees."""
    from soerenpragya3dglasma_amd import GlasmaBatch
    specs = SPECS[:3]
    inputs = [random_fields(sp[0], seed=60 + i) for i, sp in enumerate(specs)]
    with GlasmaBatch(specs) as b:
        for ev, f in zip(b.events, inputs):
            upload(ev, f)
        b.update(2)
        b.events[1].Update(1)
        b.update(2)
        b.update_serial(1)
        got = [ev.get_fields() for ev in b.events]
    for i, (sp, f) in enumerate(zip(specs, inputs)):
        ref = _alone(sp, f, 5 + (1 if i == 1 else 0))[0]
        for k in KEYS:
            assert np.array_equal(got[i][k], ref[k]), (i, k)
