"""GPU tests of the configuration files (glasma_save/load_configuration: the reference's text format,
IO/SaveConfiguration.cpp and IO/LoadConfiguration.cpp; glasma_save/load_state: exact binary restarts) and of
the command-line driver's restart paths."""
import os
import subprocess

import numpy as np
import pytest

from tests.helpers import random_fields, upload

pytestmark = pytest.mark.gpu

KEYS = ("U", "E", "Jp", "Jm", "JpS", "JmS")
SOLVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "soerenpragya3dglasma_amd", "_lib",
                      "glasma_solver")


def _glasma(N, bc=0, dt=0.2):
    from soerenpragya3dglasma_amd import Glasma
    return Glasma(N, (1.0, 0.9, 0.8), dt, bc)


def test_binary_state_restart_is_bit_identical(tmp_path):
    N = (8, 6, 10)
    f = random_fields(N, seed=4)
    path = str(tmp_path / "s.bin")
    with _glasma(N) as g, _glasma(N) as h:
        upload(g, f)
        g.Update(7)
        g.save_state(path)
        h.load_state(path)
        assert h.time() == g.time()
        g.Update(5)
        h.Update(5)
        a, b = g.get_fields(), h.get_fields()
        for k in KEYS:
            assert np.array_equal(a[k], b[k]), k
        assert np.array_equal(g.get_uold(), h.get_uold())
        assert g.time() == h.time()


@pytest.mark.parametrize("bc", [0, 1])
def test_text_configuration_round_trip(tmp_path, bc):
    N = (5, 4, 6)
    P = N[0] * N[1]
    f = random_fields(N, seed=8, bc=bc)
    u, e, j = (str(tmp_path / n) for n in ("U.txt", "E.txt", "J.txt"))
    with _glasma(N, bc) as g, _glasma(N, bc) as h:
        upload(g, f)
        g.Update(2)
        g.save_configuration(u, e, j)
        layers = N[2] + 2 if bc == 0 else N[2]  # open z: the ghost layers are saved too
        U_lines = open(u).read().splitlines()
        assert len(U_lines) == 3 * P * layers and len(open(e).read().splitlines()) == 3 * P * layers
        assert len(open(j).read().splitlines()) == 8 * P * (N[2] + 2)
        first = U_lines[0].split()
        assert first[:4] == ["0", "0", "-1" if bc == 0 else "0", "0"] and len(first) == 4 + 18
        h.load_configuration(u, e, j, t=1.6)
        assert h.time() == (1.6, 8)  # round(t / dt)
        a, b = g.get_fields(), h.get_fields()
        for k in KEYS:
            scale = np.abs(a[k]).max()
            assert np.abs(a[k] - b[k]).max() <= 1e-14 * scale, k  # 15 printed digits


def test_each_current_record_sets_its_own_colour(tmp_path):
    """The reference's LoadCurrents wrote a line's values into colours a..2; each record sets colour a only."""
    N = (3, 2, 2)
    j = tmp_path / "J.txt"
    with open(j, "w") as out:
        for z in range(-1, N[2] + 1):
            for y in range(N[1]):
                for x in range(N[0]):
                    for a in range(8):
                        v = 1000 * z + 100 * y + 10 * x + a
                        out.write(f"{x} {y} {z} {a} {v} {v + 0.25} {v + 0.5} {v + 0.75} \n")
    with _glasma(N) as g:
        g.load_configuration(jfile=str(j), t=0.0)
        d = g.get_fields()
    for z in range(N[2]):
        for y in range(N[1]):
            for x in range(N[0]):
                v = 1000 * z + 100 * y + 10 * x + np.arange(8)
                assert np.array_equal(d["JpS"][z + 1, y, x], v)
                assert np.array_equal(d["Jm"][z + 1, y, x], v + 0.75)


def test_load_errors_are_reported(tmp_path):
    from soerenpragya3dglasma_amd import GlasmaError
    N = (4, 3, 3)
    u = tmp_path / "U.txt"
    with _glasma(N) as g:
        g.save_configuration(ufile=str(u))
        lines = open(u).read().splitlines()
        (tmp_path / "short.txt").write_text("\n".join(lines[:-3]) + "\n")
        with pytest.raises(GlasmaError, match="records"):
            g.load_configuration(ufile=str(tmp_path / "short.txt"))
        (tmp_path / "swap.txt").write_text("\n".join([lines[1], lines[0]] + lines[2:]) + "\n")
        with pytest.raises(GlasmaError, match="positions do not match"):
            g.load_configuration(ufile=str(tmp_path / "swap.txt"))
        with pytest.raises(GlasmaError, match="cannot open"):
            g.load_configuration(efile=str(tmp_path / "missing.txt"))
        g.save_state(str(tmp_path / "s.bin"))
    with _glasma((4, 3, 4)) as h:
        with pytest.raises(GlasmaError, match="another lattice"):
            h.load_state(str(tmp_path / "s.bin"))


def _solver(*args):
    r = subprocess.run([SOLVER] + [str(a) for a in args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]


def test_driver_restart_from_state_matches_one_run(tmp_path):
    base = ["-N", 8, "-Nz", 12, "-a", 1, "-dT", 0.25, "-SEED", 2]
    _solver(*base, "-steps", 80, "-o", tmp_path / "one")
    _solver(*base, "-steps", 40, "-o", tmp_path / "a", "-state_out", tmp_path / "s.bin")
    _solver(*base, "-steps", 80, "-o", tmp_path / "b", "-state_in", tmp_path / "s.bin")
    one = open(tmp_path / "one" / "EnergyID2.txt").read().splitlines()
    b = open(tmp_path / "b" / "EnergyID2.txt").read().splitlines()
    assert b[-2:] == one[-2:]  # steps 60 and 80, to the printed digit


def test_driver_loaded_configuration(tmp_path):
    from soerenpragya3dglasma_amd.engine import bulk_from_raw
    N = (6, 6, 8)
    f = random_fields(N, seed=12)
    u, e, j = (str(tmp_path / n) for n in ("U.txt", "E.txt", "J.txt"))
    with _glasma(N, dt=0.25) as g:
        upload(g, f)
        g.save_configuration(u, e, j)
        g.load_configuration(u, e, j, t=5.0)
        T = bulk_from_raw(g.bulk_raw(), (1.0, 0.9, 0.8))
    out = tmp_path / "run"
    _solver("-Nx", 6, "-Ny", 6, "-Nz", 8, "-ax", 1, "-ay", 0.9, "-az", 0.8, "-dT", 0.25, "-SEED", 1, "-Loaded", 1,
            "-i1", u, "-i2", e, "-i3", j, "-Time", 5.0, "-steps", 40, "-o", out)
    names = os.listdir(out)
    assert {"UDynT5ID1.txt", "EDynT5ID1.txt", "JDynT5ID1.txt"} <= set(names), names
    first = [float(v) for v in open(out / "EnergyID1.txt").read().splitlines()[0].split()]
    assert first[0] == 5.0 and first[1] == pytest.approx(T["T00"], rel=1e-13)
