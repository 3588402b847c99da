"""GPU parity of the physical-charge (TMD / GBW) initial conditions (glasma_tmd_initial_state;
SC/TMDNUCLEUS/*.cpp, SC/SimulationmitPC.cpp:315-370) against the oracle, from the same host draws
(glasma_tmd_draw, bit-identical to the oracle's).  Tolerance as for the MV initial state: fields 1e-11."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import fields_maxdiff

pytestmark = pytest.mark.gpu

AX = 0.33 * 0.197 / 0.5


def _params(A, centers, meff=0.33):
    from soerenpragya3dglasma_amd import Glasma
    afm = (AX, AX, AX / (0.33 * 512))
    return (Glasma.tmd_params(0.33, meff, 500 / 0.197, afm, A, *centers),
            O.TmdParams(0.33, meff, 500 / 0.197, afm, A, centers))


@pytest.mark.parametrize("driver,N,a,A,centers", [(0, (8, 8, 32), (1.0, 1.0, 1.0), 16, (8, 23)),
                                                  (0, (12, 10, 24), (1.0, 0.9, 0.8), 197, (6, 17)),
                                                  (1, (8, 6, 20), (1.0, 1.0, 1.0), 54, (9, 0)),
                                                  (2, (8, 6, 20), (1.0, 1.0, 1.0), 208, (0, 11))])
def test_tmd_initial_state_matches_oracle(driver, N, a, A, centers):
    from soerenpragya3dglasma_amd import Glasma
    dt = 0.25
    p, po = _params(A, centers)
    nl, nr, pl, pr = Glasma.tmd_draw(9, N, A)
    with Glasma(N, a, dt) as g:
        g.tmd_initial_state(driver, p, nl, nr, pl, pr)
        d = g.get_fields()
        uold = g.get_uold()
        _, gmax, _ = g.gauss()
    f, _, _ = O.tmd_initial_state(driver, N, a, dt, po, nl, nr, pl, pr)
    diff = fields_maxdiff(d, f)
    diff["UOld"] = float(np.abs(uold - f.UOld).max())
    print(driver, N, A, diff, gmax)
    assert max(diff.values()) < 1e-11, diff
    assert np.abs(d["E"]).max() > 1e-6  # a non-trivial state
    assert gmax < 1e-12


def test_tmd_argument_errors():
    from soerenpragya3dglasma_amd import Glasma, GlasmaError
    N = (4, 4, 8)
    p, _ = _params(16, (2, 5))
    nl, nr, pl, pr = Glasma.tmd_draw(1, N, 16)
    with Glasma(N, (1.0, 1.0, 1.0), 0.25) as g:
        p.com_energy = -1.0
        with pytest.raises(GlasmaError, match="physical-charge"):
            g.tmd_initial_state(0, p, nl, nr, pl, pr)
        p.com_energy = 500 / 0.197
        with pytest.raises(GlasmaError, match="driver"):
            g.tmd_initial_state(3, p, nl, nr, pl, pr)


def test_driver_tmd_run(tmp_path):
    """glasma_solver -driver tmd: the physical-charge collision's initial energy equals the API's, then 100 steps
    with the collision cadence."""
    import os
    import subprocess
    from soerenpragya3dglasma_amd import Glasma
    solver = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "soerenpragya3dglasma_amd",
                          "_lib", "glasma_solver")
    N, seed = (8, 8, 32), 4
    out = tmp_path / "out"
    r = subprocess.run([solver, "-o", str(out), "-Nx", "8", "-Ny", "8", "-Nz", "32", "-a", "1", "-dT", "0.25",
                        "-SEED", str(seed), "-driver", "tmd", "-A", "16", "-centerL", "8", "-centerR", "23",
                        "-steps", "100"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    energy = [l.split() for l in open(out / f"EnergyID{seed}.txt").read().splitlines()]
    assert len(energy) == 6 and [int(e[0]) for e in energy[1:]] == [20, 40, 60, 80, 100]
    assert os.path.exists(out / f"EMTDynT25ID{seed}.txt")
    qs = 0.33
    afm = qs * 0.197 / 0.5
    p = Glasma.tmd_params(qs, 0.33, 500 / 0.197, (afm, afm, afm / (qs * 512)), 16, 8, 23)
    nl, nr, pl, pr = Glasma.tmd_draw(seed, N, 16)
    with Glasma(N, (1.0, 1.0, 1.0), 0.25) as g:
        g.tmd_initial_state(0, p, nl, nr, pl, pr)
        T0 = g.bulk()["T00"]
    assert float(energy[0][1]) == pytest.approx(T0, rel=1e-13)
    r = subprocess.run([solver, "-o", str(tmp_path / "bad"), "-N", "4", "-Nz", "8", "-driver", "tmd", "-A", "12"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Woods-Saxon" in r.stderr
