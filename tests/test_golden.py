"""The oracle reproduces the committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py).  CPU only; the GPU counterpart is tests/test_gpu_golden.py."""
import os

import numpy as np

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIELDS = ("U", "UOld", "E", "Jp", "Jm", "JpS", "JmS")


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def fields_from(d, prefix, N):
    f = O.Fields(tuple(int(n) for n in N))
    for k in FIELDS:
        getattr(f, k)[...] = d[prefix + k]
    return f


def test_leapfrog_fixture():
    d = load("leapfrog_4x4x6.npz")
    N, a, dt, bc = d["N"], tuple(d["a"]), float(d["dt"]), int(d["bc"])
    f = fields_from(d, "in_", N)
    O.update(f, a, dt, int(d["steps"]), bc)
    for k in FIELDS:
        assert np.abs(getattr(f, k) - d["out_" + k]).max() < 1e-13, k
    assert np.allclose(O.bulk(f, a, bc), d["bulk"], rtol=1e-13)
    assert np.allclose(O.emt(f, a, bc), d["emt"], rtol=1e-12, atol=1e-15)


def test_mv_fixture():
    d = load("mv_collision_6x6x8.npz")
    N, a, dt = tuple(int(n) for n in d["N"]), tuple(d["a"]), float(d["dt"])
    lam, meff, sl, sr, cl, cr = d["mv"]
    xl, xr = O.draw_charges(1, N[0], N[1], 0.5, 0.5)
    assert np.array_equal(xl, d["xi_l"]) and np.array_equal(xr, d["xi_r"])
    f = O.mv_initial_state(0, N, a, dt, xl, xr, lam=lam, meff=meff, sigma_l=sl, sigma_r=sr, center_l=int(cl),
                           center_r=int(cr))
    for k in FIELDS:
        assert np.abs(getattr(f, k) - d["out_" + k]).max() < 1e-13, k


def test_coulomb_fixture():
    d = load("coulomb_8x6x1.npz")
    N, a = tuple(int(n) for n in d["N"]), tuple(d["a"])
    f = O.Fields(N)
    f.U[...] = d["in_U"]
    f.E[...] = d["in_E"]
    G = np.zeros_like(d["G"])
    G[..., 0, 0] = G[..., 1, 1] = G[..., 2, 2] = 1
    UG, EG, st = O.coulomb_gauge(f, a, G, 1e-12, 200, 2.0)
    assert st[0] == d["stats"][0]
    assert np.abs(UG[1:-1] - d["UG"][1:-1]).max() < 1e-12
    assert np.abs(G[1:-1] - d["G"][1:-1]).max() < 1e-12
