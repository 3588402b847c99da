"""Pins of the oracle's leapfrog (Dynamics::SUNc::Update, SC/DYNAMICS/SUNcDynamics.cpp:353-378) by what
the mathematics fixes, and of the engine's per-site step code (built for the host) against the oracle.
CPU only.

* gauge covariance (periodic z): stepping a gauge-transformed configuration equals transforming the
  stepped one -- catches a wrong link orientation or a missing parallel transport;
* Gauss-law conservation: a configuration satisfying the Gauss law keeps it at round-off;
* free-field dispersion: a small transverse plane wave oscillates with the leapfrog dispersion
  sin(w dt/2) = (dt/a) sin(k a/2) -- pins the overall normalisation and sign of the E update and of the
  link exponent;
* J advection: with U = 1 and dt = a_z, J+ / J+stat move one site up and J- / J-stat one site down per
  step, exactly;
* SU(2) in SU(3): an SU(3) run whose fields lie in the lambda_1..3 subalgebra equals the SU(2) run of the
  reference's own SU(2) formulas (SC/SUNc/SU2/*.cpp), to round-off;
* energy: with J = 0 the energy drift shrinks with the time step.
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from soerenpragya3dglasma_amd import layout as LY
from tests.helpers import T, hostcheck_lib, random_fields, random_su3

A1 = (1.0, 1.0, 1.0)


def random_gauge(N, seed, bc=1):
    rng = np.random.default_rng(seed)
    G = np.zeros((N[2] + 2, N[1], N[0], 3, 3), complex)
    G[1:-1] = np.swapaxes(random_su3(rng, (N[2], N[1], N[0]), 1.5), -1, -2)  # column-major storage
    if bc:
        G[0], G[-1] = G[-2], G[1]
    else:
        G[0], G[-1] = G[1], G[-2]
    return G


def transformed(f, a, G, bc=1):
    g = f.copy()
    O.gauge_transform(g, a, G, bc)
    return g


@pytest.mark.parametrize("a", [A1, (1.0, 0.8, 0.5)])
def test_gauge_covariance_periodic(a):
    N = (4, 3, 5)
    f = random_fields(N, seed=1, bc=1)
    G = random_gauge(N, 2)
    g = transformed(f, a, G)
    O.update(f, a, 0.1, 2, bc=1)
    O.update(g, a, 0.1, 2, bc=1)
    ft = transformed(f, a, G)
    for k in ("U", "E", "Jp", "Jm", "JpS", "JmS"):
        assert np.abs(getattr(g, k)[1:-1] - getattr(ft, k)[1:-1]).max() < 1e-13, k


def test_observables_gauge_invariant():
    N = (4, 4, 4)
    a = (1.0, 0.9, 0.7)
    f = random_fields(N, seed=3, bc=1)
    f.UOld[:] = random_fields(N, seed=4, bc=1).U
    G = random_gauge(N, 5)
    g = transformed(f, a, G)
    g.UOld[:] = f.UOld
    O.lib().ref_gauge_transform(O.ivec(N), O.dvec(a), 1, O.P(G), O.P(g.UOld), O.P(np.zeros_like(g.E)),
                                *[O.P(np.zeros_like(g.Jp)) for _ in range(4)])
    assert np.allclose(O.bulk(g, a, 1), O.bulk(f, a, 1), rtol=1e-13)
    assert np.allclose(O.gauss(g, a, 1)[0], O.gauss(f, a, 1)[0], rtol=1e-12)
    assert np.allclose(O.emt(g, a, 1), O.emt(f, a, 1), rtol=1e-11, atol=1e-14)


@pytest.mark.parametrize("a,bc", [(A1, 0), ((1.0, 0.9, 0.6), 0), (A1, 1)])
def test_gauss_law_conserved_from_pure_links(a, bc):
    """E = J = 0 with random links satisfies the Gauss law; the dynamics keeps it in every layer except,
    with open z boundaries, z = 0: there the law's backward difference reads the ghost copy
    U_z(-1) = U_z(0), E_z(-1) = E_z(0), which is not a link of the dynamics (reference behaviour)."""
    N = (4, 4, 6)
    f = random_fields(N, seed=6, e_scale=0.0, j_scale=0.0, bc=bc)
    O.update(f, a, 0.05, 30, bc)
    per_z, _, _ = O.gauss(f, a, bc)
    assert np.abs(f.E).max() > 1e-2
    assert per_z[1 if bc == 0 else 0:].max() < 1e-26


def test_gauss_law_conserved_from_mv_initial_state():
    N = (8, 8, 16)
    xl, xr = O.draw_charges(1, N[0], N[1], 0.5, 0.5)
    f = O.mv_initial_state(1, N, A1, 0.25, xl, xr, lam=1.65, meff=0.1, sigma_l=2.0, sigma_r=2.0,
                           center_l=5, center_r=8)
    """The MV sheet satisfies the Gauss law by construction (J-stat is its charge); the dynamics keeps it in
    every layer but the open boundary z = 0 (see above)."""
    assert O.gauss(f, A1)[0].max() < 1e-26
    O.update(f, A1, 0.25, 20)
    assert O.gauss(f, A1)[0][1:].max() < 1e-26


def test_free_field_dispersion():
    """U_x(z) = exp(i eps cos(k z) t^3), E = 0 at t = -dt/2: A_n = A_0 cos(w (n + 1/2) dt) / cos(w dt / 2)."""
    N = (2, 2, 32)
    dt, eps = 0.1, 1e-6
    kz = 2 * np.pi * 3 / N[2]
    f = O.Fields(N)
    z = np.arange(N[2])
    for zi in z:
        al = np.zeros(8)
        al[2] = eps * np.cos(kz * zi)
        f.U[zi + 1, :, :, 0] = O.su3_expi(1.0, al).reshape(3, 3)
    f.sync(1)
    w = 2.0 / dt * np.arcsin(dt * np.sin(kz / 2))
    n = 37
    O.update(f, A1, dt, n, bc=1)
    A = np.array([O.su3_ilog(1.0, f.U[zi + 1, 0, 0, 0].reshape(9))[2] for zi in z])
    ref = eps * np.cos(kz * z) * np.cos(w * (n + 0.5) * dt) / np.cos(w * dt / 2)
    assert np.abs(A - ref).max() < 1e-5 * eps


def test_j_advection_exact_shift():
    N = (2, 2, 8)
    rng = np.random.default_rng(9)
    f = O.Fields(N)
    fp = rng.normal(size=(N[2], 8))
    fm = np.roll(fp, 1, axis=0)  # fm(z) = fp(z-1): then J- = J+ everywhere, E_z gets no source, U stays 1
    for zi in range(N[2]):
        f.JpS[zi + 1] = fp[zi]
        f.Jp[zi + 1] = fp[zi]                  # J+(z) = rho+(z): a right mover
        f.JmS[zi + 1] = fm[zi]
        f.Jm[zi + 1] = fm[(zi + 1) % N[2]]     # J-(z) = rho-(z+1): a left mover
    f.sync(1)
    O.update(f, A1, 1.0, 1, bc=1)
    for zi in range(N[2]):
        assert np.abs(f.JpS[zi + 1, 0, 0] - fp[(zi - 1) % N[2]]).max() < 1e-14
        assert np.abs(f.Jp[zi + 1, 0, 0] - fp[(zi - 1) % N[2]]).max() < 1e-14
        assert np.abs(f.JmS[zi + 1, 0, 0] - fm[(zi + 1) % N[2]]).max() < 1e-14
        assert np.abs(f.Jm[zi + 1, 0, 0] - fm[(zi + 2) % N[2]]).max() < 1e-14


def _su2_to_su3(q):
    """quaternion (u0,u1,u2,u3) -> SU(3) storage (column-major) with the SU(2) block upper-left."""
    u0, u1, u2, u3 = np.moveaxis(q, -1, 0)
    M = np.zeros(q.shape[:-1] + (3, 3), complex)
    M[..., 0, 0] = u3 + 1j * u2
    M[..., 0, 1] = 1j * u0 + u1
    M[..., 1, 0] = 1j * u0 - u1
    M[..., 1, 1] = u3 - 1j * u2
    M[..., 2, 2] = 1.0
    return np.swapaxes(M, -1, -2)


@pytest.mark.parametrize("bc", [0, 1])
def test_su2_embedding(bc):
    N = (4, 3, 5)
    a = (1.0, 0.9, 0.7)
    rng = np.random.default_rng(10)
    f2 = O.Fields(N, nc=2)
    for idx in np.ndindex(N[2], N[1], N[0], 3):
        v = np.ascontiguousarray(rng.normal(size=3) * 0.7)
        q = np.zeros(4)
        O.lib().ref_su2_expi(1.0, O.P(v), O.P(q))
        f2.U[(idx[0] + 1,) + idx[1:]] = q
    f2.E[1:-1] = rng.normal(size=f2.E[1:-1].shape) * 0.3
    for J in (f2.Jp, f2.Jm, f2.JpS, f2.JmS):
        J[1:-1] = rng.normal(size=J[1:-1].shape) * 0.2
    f2.sync(bc)
    f3 = O.Fields(N)
    f3.U[...] = _su2_to_su3(f2.U)
    f3.E[..., :3] = f2.E
    for k in ("Jp", "Jm", "JpS", "JmS"):
        getattr(f3, k)[..., :3] = getattr(f2, k)
    O.update(f2, a, 0.1, 3, bc)
    O.update(f3, a, 0.1, 3, bc)
    assert np.abs(f3.U - _su2_to_su3(f2.U)).max() < 1e-14
    assert np.abs(f3.E[..., :3] - f2.E).max() < 1e-14 and np.abs(f3.E[..., 3:]).max() < 1e-15
    for k in ("Jp", "Jm", "JpS", "JmS"):
        assert np.abs(getattr(f3, k)[..., :3] - getattr(f2, k)).max() < 1e-14
    b2, b3 = O.bulk(f2, a, bc), O.bulk(f3, a, bc)
    assert np.allclose(b2, b3, rtol=1e-13, atol=1e-14)


def test_energy_drift_shrinks_with_dt():
    """J = 0: the largest excursion of T00 over a fixed time interval falls as the step shrinks."""
    N = (4, 4, 4)
    a = A1
    drift = []
    for dt in (0.1, 0.05):
        f = random_fields(N, seed=12, link_scale=0.3, e_scale=0.2, j_scale=0.0, bc=1)
        e0 = O.bulk_T(O.bulk(f, a, 1), a)["T00"]
        dev = 0.0
        for _ in range(int(round(1.0 / dt))):
            O.update(f, a, dt, 1, bc=1)
            dev = max(dev, abs(O.bulk_T(O.bulk(f, a, 1), a)["T00"] - e0))
        drift.append(dev / e0)
    assert drift[1] < 0.6 * drift[0], drift


# ------------------------------------------------------------------------------------------------
# the engine's per-site step code (site.h), built for the host, against the oracle
# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("N,bc", [((6, 5, 7), 0), ((8, 4, 6), 1), ((16, 8, 1), 0), ((3, 3, 2), 0)])
def test_device_step_code_matches_oracle(N, bc):
    hc = hostcheck_lib()
    dp = ctypes.POINTER(ctypes.c_double)
    a, dt = (1.0, 0.9, 0.7), 0.1
    f = random_fields(N, seed=2, bc=bc)
    S = N[0] * N[1] * N[2]
    Us = LY.tile(LY.u_aos_to_soa(f.U))
    Es = LY.tile(LY.v_aos_to_soa(f.E, 24))
    Js = LY.tile(np.concatenate([LY.v_aos_to_soa(J, 8) for J in (f.Jp, f.Jm, f.JpS, f.JmS)]))
    Un, Jn, UO = np.zeros_like(Us), np.zeros_like(Js), np.zeros_like(Us)
    p = lambda x: x.ctypes.data_as(dp)  # noqa: E731
    hc.hc_step((ctypes.c_int * 3)(*N), (ctypes.c_double * 3)(*a), bc, dt, p(Us), p(Un), p(Es), p(Js), p(Jn), p(UO))
    g = f.copy()
    O.update(g, a, dt, 1, bc)
    assert np.abs(LY.u_soa_to_aos(LY.untile(Un, 54, S), N, bc) - g.U).max() < 1e-13
    assert np.abs(LY.v_soa_to_aos(LY.untile(Es, 24, S), N, (3, 8), bc) - g.E).max() < 1e-13
    Jg = LY.untile(Jn, 32, S)
    for k, J in enumerate((g.Jp, g.Jm, g.JpS, g.JmS)):
        assert np.abs(LY.v_soa_to_aos(Jg[8 * k:8 * k + 8], N, (8,), bc) - J).max() < 1e-13
    assert np.abs(LY.u_soa_to_aos(LY.untile(UO, 54, S), N, bc) - g.UOld).max() < 1e-13
