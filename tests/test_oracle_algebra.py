"""Pins of the oracle's per-site SU(3) / SU(2) algebra (oracle/cpu_ref.cpp) against numpy / scipy, and of
the engine's per-site device code (su3.h, compiled for the host as libglasma_hostcheck.so) against the
oracle.  CPU only.

What each check can catch: products (transposed / conjugated operand), ReTrIGenU (generator order or
sign, the 1/sqrt3 of t^8), MatrixIExp (sign of the exponent, normalisation of t^a = lambda^a/2),
MatrixILog (branch, round-off near the identity), the adjoint rotations (V vs V^dagger), GetAdjoint
(index order).
"""
import ctypes

import numpy as np
import pytest
import scipy.linalg as sl

from oracle import oracle as O
from tests.helpers import T, herm, random_su3

dp = ctypes.POINTER(ctypes.c_double)


def flat(M):
    """3x3 matrix [i, j] -> reference storage (column-major i + 3j)."""
    return np.ascontiguousarray(np.asarray(M).T).reshape(9)


def unflat(v):
    return np.asarray(v).reshape(3, 3).T


def alg(M):
    """Components A^a = 2 tr(t^a M) of a Hermitian traceless matrix."""
    return np.array([2 * np.trace(T[a] @ M).real for a in range(8)])


@pytest.fixture(scope="module")
def mats():
    rng = np.random.default_rng(17)
    V = random_su3(rng, (40,), 1.0)
    W = random_su3(rng, (40,), 2.0)
    Z = rng.normal(size=(40, 3, 3)) + 1j * rng.normal(size=(40, 3, 3))
    al = rng.normal(size=(40, 8))
    return V, W, Z, al


def test_products(mats):
    V, W, Z, _ = mats
    for A, B in zip(V, Z):
        for mode, ref in ((0, A @ B), (1, A @ B.conj().T), (2, A.conj().T @ B), (3, A.conj().T @ B.conj().T)):
            C = np.zeros(9, complex)
            O.lib().ref_su3_mul(O.P(flat(A)), O.P(flat(B)), O.P(C), mode)
            assert np.abs(unflat(C) - ref).max() < 1e-14


def test_traces_and_norms(mats):
    _, _, Z, _ = mats
    for M in Z:
        r = O.su3_call("ref_su3_retrigenu", flat(M))
        ref = np.array([np.trace(1j * T[a] @ M).real for a in range(8)])
        assert np.abs(r - ref).max() < 1e-14
        f = O.lib().ref_su3_retridminusu(O.P(flat(M)))
        assert abs(f - np.trace(np.eye(3) - M).real) < 1e-14
        u = O.lib().ref_su3_unitarity_norm(O.P(flat(M)))
        assert abs(u - np.linalg.norm(M @ M.conj().T - np.eye(3))) < 1e-12 * max(1, abs(u))


@pytest.mark.parametrize("c", [1.0, -0.37, 2.5])
def test_expi_against_scipy(mats, c):
    _, _, _, al = mats
    for v in al:
        X = unflat(O.su3_expi(c, v))
        ref = sl.expm(1j * c * herm(v))
        assert np.abs(X - ref).max() < 2e-14
        assert np.abs(X @ X.conj().T - np.eye(3)).max() < 2e-15
        assert abs(np.linalg.det(X) - 1) < 2e-15


@pytest.mark.parametrize("scale", [1e-7, 1e-4, 0.1, 1.0, 2.5])
def test_ilog_round_trip_and_scipy(scale):
    rng = np.random.default_rng(int(scale * 1e7) % 1000 + 3)
    for M in random_su3(rng, (60,), scale):
        e = O.su3_ilog(-0.25, flat(M))
        assert np.abs(unflat(O.su3_expi(-0.25, e)) - M).max() < 3e-15
        # branch: the principal phases of M, with one moved by 2 pi when they sum to +-2 pi -- the traceless
        # lift of smallest norm
        ev = np.sort(np.linalg.eigvalsh(-0.25 * herm(e)))
        ph = np.angle(np.linalg.eigvals(M))
        lifts = [ph + 2 * np.pi * np.eye(3)[k] * s for k in range(3) for s in (-1, 0, 1)]
        lifts = [l for l in lifts if abs(l.sum()) < 1e-9]
        best = min(lifts, key=lambda l: np.sum(l ** 2))
        assert np.abs(ev - np.sort(best)).max() < 1e-9
        if scale <= 1.0:
            ref = alg(sl.logm(M) / (1j * -0.25))
            assert np.abs(e - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())


def test_ilog_near_identity_with_opposite_phases():
    """Eigenphases +phi, ~0, -phi near the identity: cos is degenerate, the logarithm is not."""
    rng = np.random.default_rng(8)
    Q = random_su3(rng, (1,), 1.0)[0]
    ph = np.array([1.7547e-5, 8.69e-8, -1.7547e-5 - 8.69e-8])
    M = Q @ np.diag(np.exp(1j * ph)) @ Q.conj().T
    e = O.su3_ilog(1.0, flat(M))
    assert np.abs(np.sort(np.linalg.eigvalsh(herm(e))) - np.sort(ph)).max() < 1e-15


def test_adjoint_rotations(mats):
    V, _, _, al = mats
    for M, v in zip(V, al):
        A = herm(v)
        ad = O.su3_call("ref_su3_adj_mul", flat(M), v)
        iad = O.su3_call("ref_su3_inv_adj_mul", flat(M), v)
        assert np.abs(ad - alg(M @ A @ M.conj().T)).max() < 1e-14
        assert np.abs(iad - alg(M.conj().T @ A @ M)).max() < 1e-14
        adj = O.su3_call("ref_su3_get_adjoint", flat(M), out_shape=64).reshape(8, 8).T  # AdjIndex(a, b) = a + 8b
        ref = np.array([[2 * np.trace(T[a] @ M @ T[b] @ M.conj().T).real for b in range(8)] for a in range(8)])
        assert np.abs(adj - ref).max() < 1e-14
        assert np.abs(adj @ v - ad).max() < 1e-13


def su2_matrix(q):
    """SU(2) element u3 + i u_k sigma_{k+1} (SC/SUNc/SU2/GroupOperations.cpp:179-191) as a 2x2 matrix."""
    u0, u1, u2, u3 = q
    return np.array([[u3 + 1j * u2, 1j * u0 + u1], [1j * u0 - u1, u3 - 1j * u2]])


def test_su2_exp_and_log():
    rng = np.random.default_rng(5)
    sig = [np.array([[0, 1], [1, 0]]), np.array([[0, -1j], [1j, 0]]), np.diag([1.0, -1.0])]
    for _ in range(30):
        v = rng.normal(size=3)
        q = np.zeros(4)
        O.lib().ref_su2_expi(0.8, O.P(np.ascontiguousarray(v)), O.P(q))
        ref = sl.expm(1j * 0.8 * sum(v[k] * sig[k] / 2 for k in range(3)))
        assert np.abs(su2_matrix(q) - ref).max() < 1e-14
        w = np.zeros(3)
        O.lib().ref_su2_ilog(0.8, O.P(q), O.P(w))
        if np.linalg.norm(v) * 0.8 / 2 < np.pi:
            assert np.abs(w - v).max() < 1e-12


# ------------------------------------------------------------------------------------------------
# the engine's per-site device code, built for the host, against the oracle
# ------------------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def hc():
    from tests.helpers import hostcheck_lib
    return hostcheck_lib()


def test_device_algebra_matches_oracle(hc, mats):
    V, W, _, al = mats
    for M, v in zip(np.concatenate([V, W]), np.concatenate([al, al])):
        Ms = flat(M)
        X = np.zeros(9, complex)
        hc.hc_su3_expi(ctypes.c_double(0.9), v.ctypes.data_as(dp), O.P(X))
        assert np.abs(X - O.su3_expi(0.9, v)).max() < 2e-15
        Xm = np.zeros(9, complex)
        hc.hc_su3_expi_m1(ctypes.c_double(0.9), v.ctypes.data_as(dp), O.P(Xm))
        assert np.abs(Xm + flat(np.eye(3)) - O.su3_expi(0.9, v)).max() < 2e-15
        e = np.zeros(8)
        hc.hc_su3_ilog(ctypes.c_double(0.9), O.P(Ms), e.ctypes.data_as(dp))
        assert np.abs(e - O.su3_ilog(0.9, Ms)).max() < 1e-11
        for name in ("adj_mul", "inv_adj_mul"):
            o = np.zeros(8)
            getattr(hc, "hc_su3_" + name)(O.P(Ms), v.ctypes.data_as(dp), o.ctypes.data_as(dp))
            assert np.abs(o - O.su3_call("ref_su3_" + name, Ms, v)).max() < 1e-14
        r = np.zeros(8)
        hc.hc_su3_retrigenu(O.P(Ms), r.ctypes.data_as(dp))
        assert np.abs(r - O.su3_call("ref_su3_retrigenu", Ms)).max() < 1e-15
