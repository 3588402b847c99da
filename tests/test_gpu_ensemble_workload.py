"""GPU test of the ensemble at its own workload: 64 MV collision events of uneven sizes, 200 steps.

The events are drawn as the benchmark draws them (master seed 12345, numpy RandomState): N_perp uniform in
{32, 48, 64, 96} (square), Nz uniform in {128, 192, 256, 384}, g^2 mu uniform in [0.3, 0.8], charge seeds 2k+1
(left) and 2k+2 (right), a = 1, dt = a/4, open z, collision initial states built on the device
(CollidingObjects::SetWilsonLines + SingleSheet, Simulation.cpp:327-378), tile orders autotuned per lattice.

* Every event after 200 batched steps (one launch per step over all events, replayed from a HIP graph) is bit
  for bit the event after 200 serial per-event steps: the device checksums (glasma_checksum) of U, UOld, E and
  the four currents agree, and so do the clocks.  The initial states of the two runs are checked to be
  bit-identical first.
* One event per transverse size (the first with the smallest Nz) is compared after its first 20 batched steps
  with the oracle's Dynamics::SUNc::Update (SC/DYNAMICS/SUNcDynamics.cpp:353-378) run from the same initial
  state: fields within 1e-11 absolute, the bulk sums within 1e-10 relative.
"""
import time

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import fields_maxdiff

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]


def _draw(nevents=64):
    rng = np.random.RandomState(12345)
    specs, qs = [], []
    for _ in range(nevents):
        nperp = int(rng.choice([32, 48, 64, 96]))
        nz = int(rng.choice([128, 192, 256, 384]))
        qs.append(float(rng.uniform(0.3, 0.8)))
        specs.append(((nperp, nperp, nz), (1.0, 1.0, 1.0), 0.25))
    return specs, qs


def _initial_states(b, specs, qs):
    from soerenpragya3dglasma_amd import Glasma
    for k, (ev, sp, q) in enumerate(zip(b.events, specs, qs)):
        N = sp[0]
        xl, _ = Glasma.draw_charges(2 * k + 1, N[0], N[1], q, q)
        _, xr = Glasma.draw_charges(2 * k + 2, N[0], N[1], q, q)
        mv = Glasma.mv_params(lam=1.65, meff=0.1, sigma_l=4.0, sigma_r=4.0,
                              center_l=int(0.85 * N[2] / 2 - 1), center_r=int(1.15 * N[2] / 2 - 1))
        ev.mv_initial_state(0, mv, xl, xr)


def test_ensemble_200_steps_bitwise_serial_and_oracle():
    from soerenpragya3dglasma_amd import GlasmaBatch
    specs, qs = _draw()
    picks = {}
    for k, sp in enumerate(specs):
        n = sp[0][0]
        if n not in picks or sp[0][2] < specs[picks[n]][0][2]:
            picks[n] = k
    print(f"[ensemble] {len(specs)} events, {sum(np.prod(s[0]) for s in specs)} sites; oracle events "
          f"{sorted(picks.values())}", flush=True)
    t0 = time.perf_counter()
    with GlasmaBatch(specs) as b:
        _initial_states(b, specs, qs)
        b.tune(reps=1)
        init = [ev.checksum() for ev in b.events]
        starts = {}
        for k in picks.values():
            ev = b.events[k]
            starts[k] = (ev.get_fields(), ev.get_uold())
        b.update(20)
        at20 = {k: (b.events[k].get_fields(), b.events[k].bulk_raw()) for k in picks.values()}
        b.update(180)
        batched = [(ev.checksum(), ev.time()) for ev in b.events]
        print(f"[ensemble] batched run done in {time.perf_counter() - t0:.1f} s", flush=True)
        _initial_states(b, specs, qs)
        again = [ev.checksum() for ev in b.events]
        b.update_serial(200)
        serial = [(ev.checksum(), ev.time()) for ev in b.events]
        print(f"[ensemble] serial run done in {time.perf_counter() - t0:.1f} s", flush=True)
    assert again == init, "the device initial states are not reproducible"
    bad = [(k, specs[k][0]) for k in range(len(specs)) if batched[k] != serial[k]]
    assert not bad, f"events whose batched run differs from the serial run: {bad}"
    assert all(c[0] != i[0] for c, i in zip(batched, init)), "an event did not advance"

    for n, k in sorted(picks.items()):
        N = specs[k][0]
        d0, uold = starts[k]
        f = O.Fields(N)
        f.U, f.E, f.Jp, f.Jm, f.JpS, f.JmS = d0["U"], d0["E"], d0["Jp"], d0["Jm"], d0["JpS"], d0["JmS"]
        O.update(f, (1.0, 1.0, 1.0), 0.25, 20)
        d20, bulk = at20[k]
        diff = fields_maxdiff(d20, f)
        ref = O.bulk(f, (1.0, 1.0, 1.0))
        rel = float(np.abs(bulk - ref).max() / np.abs(ref).max())
        print(f"[ensemble] event {k} {N}: 20 steps vs oracle, field diff {max(diff.values()):.2e}, bulk {rel:.2e}",
              flush=True)
        assert max(diff.values()) < 1e-11, (k, N, diff)
        assert rel < 1e-10, (k, N, rel)
