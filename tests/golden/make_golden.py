"""Writes the golden fixtures of this directory from the oracle (oracle/cpu_ref.cpp).

The reference ships no tests, golden vectors or fixtures, and cannot be built here (FFTW and GSL are
absent), so the fixtures are the oracle's outputs on small seeded inputs, frozen: they pin the oracle
against later edits (tests/test_golden.py) and give the GPU tests inputs and expected outputs that do
not need the oracle at run time (tests/test_gpu_golden.py).

    python tests/golden/make_golden.py [--with-collision]
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402
from tests.helpers import random_fields  # noqa: E402

FIELDS = ("U", "UOld", "E", "Jp", "Jm", "JpS", "JmS")


def leapfrog():
    N, a, dt, bc, steps = (4, 4, 6), (1.0, 0.9, 0.7), 0.1, 0, 3
    f = random_fields(N, seed=101, bc=bc)
    out = {"N": np.array(N), "a": np.array(a), "dt": dt, "bc": bc, "steps": steps}
    out.update({"in_" + k: getattr(f, k).copy() for k in FIELDS})
    O.update(f, a, dt, steps, bc)
    out.update({"out_" + k: getattr(f, k).copy() for k in FIELDS})
    out["bulk"] = O.bulk(f, a, bc)
    pz, gm, gr = O.gauss(f, a, bc)
    out["gauss_per_z"], out["gauss_max_rms"] = pz, np.array([gm, gr])
    out["emt"] = O.emt(f, a, bc)
    np.savez_compressed(os.path.join(HERE, "leapfrog_4x4x6.npz"), **out)


def mv_collision():
    N, a, dt = (6, 6, 8), (1.0, 1.0, 1.0), 0.25
    mv = dict(lam=1.65, meff=0.1, sigma_l=2.0, sigma_r=2.0, center_l=int(0.85 * 8 / 2 - 1), center_r=int(1.15 * 8 / 2 - 1))
    xl, xr = O.draw_charges(1, N[0], N[1], 0.5, 0.5)
    f = O.mv_initial_state(0, N, a, dt, xl, xr, **mv)
    out = {"N": np.array(N), "a": np.array(a), "dt": dt, "xi_l": xl, "xi_r": xr,
           "mv": np.array([mv["lam"], mv["meff"], mv["sigma_l"], mv["sigma_r"], mv["center_l"], mv["center_r"]])}
    out.update({"out_" + k: getattr(f, k).copy() for k in FIELDS})
    np.savez_compressed(os.path.join(HERE, "mv_collision_6x6x8.npz"), **out)


def coulomb():
    N, a = (8, 6, 1), (1.0, 1.0, 1.0)
    f = random_fields(N, seed=102, link_scale=0.3)
    G = np.zeros((N[2] + 2, N[1], N[0], 3, 3), complex)
    G[..., 0, 0] = G[..., 1, 1] = G[..., 2, 2] = 1
    out = {"N": np.array(N), "a": np.array(a), "in_U": f.U.copy(), "in_E": f.E.copy()}
    UG, EG, st = O.coulomb_gauge(f, a, G, 1e-12, 200, 2.0)
    out.update({"UG": UG, "EG": EG, "G": G, "stats": st})
    np.savez_compressed(os.path.join(HERE, "coulomb_8x6x1.npz"), **out)


def collision_config(steps=(0, 20, 100)):
    """The two-nucleus MV collision workload at its full size: 64 x 64 x 256, a = 1, dt = a/4, charges of
    g^2 mu = 0.5 from seeds 1 (left) and 2 (right), longitudinal width 6, centres 64 apart; the profiles that
    must agree to 1e-10 over the first 100 steps (bulk sums, the energy-momentum tensor per z, the Gauss law
    per z).  Takes several minutes on the CPU, so tests/test_golden.py does not recompute it; the oracle's
    parts are pinned at small sizes there and in tests/test_oracle_*.py."""
    N, a, dt = (64, 64, 256), (1.0, 1.0, 1.0), 0.25
    mv = dict(lam=1.65, meff=0.1, sigma_l=6.0, sigma_r=6.0, center_l=128 - 32, center_r=128 + 32)
    xl, _ = O.draw_charges(1, N[0], N[1], 0.5, 0.5)
    _, xr = O.draw_charges(2, N[0], N[1], 0.5, 0.5)
    f = O.mv_initial_state(0, N, a, dt, xl, xr, **mv)
    out = {"N": np.array(N), "a": np.array(a), "dt": dt, "steps": np.array(steps),
           "mv": np.array([mv["lam"], mv["meff"], mv["sigma_l"], mv["sigma_r"], mv["center_l"], mv["center_r"]]),
           "seeds": np.array([1, 2]), "qs": 0.5}
    done = 0
    for n in steps:
        O.update(f, a, dt, n - done)
        done = n
        out[f"bulk_{n}"] = O.bulk(f, a)
        out[f"emt_{n}"] = O.emt(f, a)
        pz, gm, gr = O.gauss(f, a)
        out[f"gauss_per_z_{n}"], out[f"gauss_max_rms_{n}"] = pz, np.array([gm, gr])
    np.savez_compressed(os.path.join(HERE, "collision_64x64x256.npz"), **out)


if __name__ == "__main__":
    leapfrog()
    mv_collision()
    coulomb()
    if "--with-collision" in sys.argv:
        collision_config()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))
