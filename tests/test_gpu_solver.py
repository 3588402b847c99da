"""GPU test of glasma_solver, the C++ command-line driver on top of the C ABI (the reference's collision /
left / right drivers, Simulation.cpp).  A short collision run writes the reference's text outputs; its energy
record agrees with the same run driven through the Python API (the same library: to the 15 printed digits, up
to the driver's own rounding of the T from the raw sums) and,
at t = 0 and after 20 steps, with the oracle (1e-10 relative)."""
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

SOLVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "soerenpragya3dglasma_amd", "_lib",
                      "glasma_solver")


def _run(tmp_path, *args):
    out = tmp_path / "out"
    r = subprocess.run([SOLVER, "-o", str(out)] + [str(a) for a in args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    return out


def test_collision_run_outputs_and_energy(tmp_path):
    from soerenpragya3dglasma_amd import Glasma
    from soerenpragya3dglasma_amd.engine import bulk_from_raw
    N, a, dt, seed = (8, 8, 16), (1.0, 1.0, 1.0), 0.25, 3
    out = _run(tmp_path, "-Nx", 8, "-Ny", 8, "-Nz", 16, "-a", 1, "-dT", dt, "-SEED", seed, "-steps", 100)
    names = sorted(os.listdir(out))
    for f in ("EnergyID3.txt", "GausCheck3.txt", "InfoID3.txt", "EMTDynT0ID3.txt", "EMTDynT25ID3.txt", "GVDynT0ID3.txt"):
        assert f in names, names
    energy = [l.split() for l in open(out / "EnergyID3.txt").read().splitlines()]
    assert len(energy) == 1 + 5  # t = 0, then every 20 steps
    assert [int(r[0]) for r in energy[1:]] == [20, 40, 60, 80, 100]
    assert len(open(out / "GausCheck3.txt").read().splitlines()) == 1 + 5
    emt = np.loadtxt(out / "EMTDynT25ID3.txt")
    assert emt.shape == (16, 11)

    # the same run through the Python API
    qsl = 0.125 / math.sqrt(a[0] * a[1])
    cl, cr = int(0.85 * N[2] / 2 - 1), int(1.15 * N[2] / 2 - 1)
    xl, xr = Glasma.draw_charges(seed, N[0], N[1], qsl, qsl)
    lam = 5 * 0.33  # Parameters::Lambda = 5 Qs0A, as the driver forms it (one ulp above the literal 1.65)
    mv = Glasma.mv_params(lam=lam, meff=0.33, sigma_l=16.0, sigma_r=16.0, center_l=cl, center_r=cr)
    with Glasma(N, a, dt) as g:
        g.mv_initial_state(0, mv, xl, xr)
        T = bulk_from_raw(g.bulk_raw(), a)
        # TXX, TYY are differences of terms of the size of T00: rounding is relative to T00
        assert [float(v) for v in energy[0]] == pytest.approx([0.0, T["T00"], T["TXX"], T["TYY"], T["TZZ"], T["SOURCE"]],
                                                               rel=1e-13, abs=1e-13 * T["T00"])
        for row in energy[1:]:
            g.Update(20)
            T = bulk_from_raw(g.bulk_raw(), a)
            assert float(row[2]) == pytest.approx(T["T00"], rel=1e-14)
        emt_api = g.emt_profile()
    ref = np.asarray(emt_api).reshape(16, 10)
    rel = np.abs(emt[:, 1:] - ref) / np.maximum(np.abs(ref), 1e-300)
    k = np.unravel_index(np.argmax(rel), rel.shape)
    assert rel.max() < 1e-13, (k, emt[:, 1:][k], ref[k], rel.max())

    # the oracle: the reference's initial state and 20 steps
    f = O.mv_initial_state(0, N, a, dt, *O.draw_charges(seed, N[0], N[1], qsl, qsl), lam=lam, meff=0.33,
                           sigma_l=16.0, sigma_r=16.0, center_l=cl, center_r=cr)
    T0 = O.bulk_T(O.bulk(f, a), a)
    assert float(energy[0][1]) == pytest.approx(T0["T00"], rel=1e-10)
    O.update(f, a, dt, 20)
    T20 = O.bulk_T(O.bulk(f, a), a)
    assert float(energy[1][2]) == pytest.approx(T20["T00"], rel=1e-10)


@pytest.mark.parametrize("driver", ["left", "right"])
def test_single_nucleus_drivers(tmp_path, driver):
    out = _run(tmp_path, "-N", 8, "-Nz", 12, "-a", 1, "-dT", 0.25, "-SEED", 5, "-steps", 40, "-driver", driver)
    energy = open(out / "EnergyID5.txt").read().splitlines()
    assert len(energy) == 3
    assert all(math.isfinite(float(v)) for v in energy[-1].split())


def test_bad_arguments_exit_nonzero(tmp_path):
    r = subprocess.run([SOLVER, "-driver", "sideways", "-o", str(tmp_path / "o")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 2 and "driver" in r.stderr
