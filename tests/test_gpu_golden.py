"""The HIP engine against the committed golden fixtures (tests/golden/*.npz): no oracle at run time."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def test_leapfrog_fixture_on_gpu():
    from soerenpragya3dglasma_amd import Glasma
    d = load("leapfrog_4x4x6.npz")
    N, a, dt, bc = tuple(int(n) for n in d["N"]), tuple(d["a"]), float(d["dt"]), int(d["bc"])
    with Glasma(N, a, dt, bc) as g:
        g.set_fields(d["in_U"], d["in_E"], d["in_Jp"], d["in_Jm"], d["in_JpS"], d["in_JmS"])
        g.Update(int(d["steps"]))
        out = g.get_fields()
        uold = g.get_uold()
        bulk = g.bulk_raw()
        emt = g.emt_profile()
    for k in ("U", "E", "Jp", "Jm", "JpS", "JmS"):
        assert np.abs(out[k] - d["out_" + k]).max() < 1e-12, k
    assert np.abs(uold - d["out_UOld"]).max() < 1e-12
    assert np.allclose(bulk, d["bulk"], rtol=1e-12)
    assert np.allclose(emt, d["emt"], rtol=1e-11, atol=1e-14)


def test_mv_fixture_on_gpu():
    from soerenpragya3dglasma_amd import Glasma
    d = load("mv_collision_6x6x8.npz")
    N, a, dt = tuple(int(n) for n in d["N"]), tuple(d["a"]), float(d["dt"])
    lam, meff, sl, sr, cl, cr = d["mv"]
    mv = Glasma.mv_params(lam, meff, sl, sr, int(cl), int(cr))
    with Glasma(N, a, dt) as g:
        g.mv_initial_state(0, mv, d["xi_l"], d["xi_r"])
        out = g.get_fields()
        uold = g.get_uold()
    for k in ("U", "E", "Jp", "Jm", "JpS", "JmS"):
        assert np.abs(out[k][1:-1] - d["out_" + k][1:-1]).max() < 1e-11, k
    assert np.abs(uold[1:-1] - d["out_UOld"][1:-1]).max() < 1e-11


def test_coulomb_fixture_on_gpu():
    from soerenpragya3dglasma_amd import Glasma
    d = load("coulomb_8x6x1.npz")
    N, a = tuple(int(n) for n in d["N"]), tuple(d["a"])
    with Glasma(N, a, 0.1) as g:
        g.set_fields(U=d["in_U"], E=d["in_E"])
        st = g.coulomb_gauge(1e-12, 200, 2.0)
        UG, EG, G = g.get_gauge_fixed()
    assert abs(st[0] - d["stats"][0]) <= 1
    assert np.abs(UG[1:-1] - d["UG"][1:-1]).max() < 1e-10
    assert np.abs(EG[1:-1] - d["EG"][1:-1]).max() < 1e-10
