"""Size-independent properties at the largest workload's full size: an FP64 MV collision on 256 x 256 x 1024
(open z, a = 1, dt = a/4, charge seeds 1 and 2, longitudinal width 8), built on the device, then 40 leapfrog
steps with the tuned step kernel.  The oracle cannot run this size in seconds, so the checks are properties
that hold at any size:

* the Gauss law: the initial state satisfies it to round-off and the leapfrog conserves it, so every layer's
  residual stays at round-off (layer 0 reads the edge copy below it and is reported separately);
* unitarity of the links stays at round-off;
* before the nuclei meet, each sheet moves unchanged, so the total energy changes only by the leapfrog's
  O(dt^2) dispersion error (5.6e-4 relative over 40 steps; bounded at 1e-3);
* the per-layer EMT energy profile and the bulk T00 describe the same state (their sums differ only by the
  clover average of the links at t - dt/2, bounded at 1e-3 relative).
"""
import pytest

pytestmark = pytest.mark.gpu


def test_collision_256x256x1024_conserves_gauss_law_and_unitarity():
    from soerenpragya3dglasma_amd import Glasma
    N = (256, 256, 1024)
    a, dt = (1.0, 1.0, 1.0), 0.25
    xl, _ = Glasma.draw_charges(1, N[0], N[1], 0.5, 0.5)
    _, xr = Glasma.draw_charges(2, N[0], N[1], 0.5, 0.5)
    mv = Glasma.mv_params(lam=1.65, meff=0.1, sigma_l=8.0, sigma_r=8.0, center_l=int(0.85 * N[2] / 2 - 1),
                          center_r=int(1.15 * N[2] / 2 - 1))
    with Glasma(N, a, dt) as g:
        g.mv_initial_state(0, mv, xl, xr)
        pz0, gmax0, _ = g.gauss()
        T0 = g.bulk()["T00"]
        g.tune_step(reps=1)
        g.Update(40)
        pz1, gmax1, _ = g.gauss()
        umax, urms = g.unitarity()
        T1 = g.bulk()["T00"]
        emt = g.emt_profile()
    scale = max(pz0.max(), 1e-300)
    print(f"gauss max {gmax0:.3e} -> {gmax1:.3e}; per-z max {pz0.max():.3e} -> {pz1[1:].max():.3e} "
          f"(layer 0 {pz1[0]:.3e}); unitarity {umax:.3e}; T00 {T0:.12e} -> {T1:.12e}; EMT sum "
          f"{emt[:, 0].sum() * N[0] * N[1]:.12e}")
    # measured: max |G| 7e-18 -> 9e-15, interior per-layer <G^2> 2e-36 -> 3e-31, unitarity 2e-14
    assert gmax0 < 1e-15 and gmax1 < 1e-13
    assert pz1[1:].max() < max(100 * scale, 1e-28)
    assert umax < 1e-13
    assert T0 > 0 and abs(T1 - T0) < 1e-3 * T0
    assert abs(emt[:, 0].sum() * N[0] * N[1] - T1) < 1e-3 * T1  # measured 1e-4
