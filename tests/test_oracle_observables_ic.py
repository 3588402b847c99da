"""Pins of the oracle's observables, MV initial conditions and Coulomb gauge fixing.  CPU only.

Observables (SC/OBSERVABLES/*.cpp):
* bulk sums against a numpy brute force on a tiny lattice (plaquettes as explicit matrix products);
* the clover-averaged energy-momentum profile (EnergyMomentumTensor::CreateMap) against the plaquette
  bulk sums: with UOld = U and periodic z, the transverse averages summed over z equal the bulk T00,
  TXX, TYY, TZZ exactly (every plaquette sits in four clovers, every E link in two averages);
* closed forms with U = 1 and uniform E (T0i = 0, Tij = q_i q_j e_i . e_j);
* unitarity and Gauss law against numpy.
* the output-file profiles (IO::SaveFields, IO::SaveTransverselyIntegratedCurrents,
  EnergyMomentumTensor::CreateMapBfCollision) against numpy plaquettes and J sums, and their totals against
  the bulk sums.
MV initial conditions (SC/COLLIDINGOBJECTS/CollidingObjects.cpp, RevSingleNucleus.cpp,
SUNcDynamics.cpp:275-350):
* Poisson solve of a single transverse mode: A = rho / p^2 with the lattice p^2;
* smoothing of a single mode: rho exp(-p^2/2 Lambda^2) p^2/(p^2 + m^2);
* EvolveRhoViaFourier of a single z mode: cos(k z -+ w dt) (pure phase);
* the Wilson line of one transverse site against scipy.linalg.expm of the Poisson solution;
* the sheet satisfies the Gauss law exactly;
* the sheet from a given charge density (the oracle side of the gradient checks) equals the MV left mover
  for the density SetChargeDensities forms;
* the numpy slab energy of the gradient checks, over all layers, equals the oracle's bulk T00.
Coulomb gauge fixing (CoulombGaugeFixing.cpp, FourierAcceleration.cpp): converges for Nz = 1, keeps
gauge-invariant observables, recovers a gauge-rotated copy.
"""
import numpy as np
import pytest
import scipy.linalg as sl

from oracle import oracle as O
from tests.helpers import T, herm, random_fields, random_su3

A1 = (1.0, 1.0, 1.0)


def mat(U, z, y, x, mu, N):
    """link [i, j] at (x, y, z) with periodic x, y and the ghost layers in z (z in [-1, Nz])."""
    return U[z + 1, y % N[1], x % N[0], mu].T


def test_bulk_brute_force():
    N = (3, 2, 4)
    a = (1.0, 0.9, 0.7)
    f = random_fields(N, seed=4)
    out = O.bulk(f, a)
    B = np.zeros(3)
    E2 = (f.E[1:-1] ** 2).sum(axis=(0, 1, 2, 4))
    EJ = (f.E[1:-1, ..., 2, :] * (f.Jp[1:-1] - f.Jm[1:-1])).sum() / np.sqrt(2)
    for z in range(N[2]):
        for y in range(N[1]):
            for x in range(N[0]):
                U = lambda dx, dy, dz, mu: mat(f.U, z + dz, y + dy, x + dx, mu, N)  # noqa: E731
                D = lambda M: M.conj().T  # noqa: E731
                Pxy = U(0, 0, 0, 0) @ U(1, 0, 0, 1) @ D(U(0, 1, 0, 0)) @ D(U(0, 0, 0, 1))
                Pyz = U(0, 0, 0, 1) @ U(0, 1, 0, 2) @ D(U(0, 0, 1, 1)) @ D(U(0, 0, 0, 2))
                Pzx = U(0, 0, 0, 2) @ U(0, 0, 1, 0) @ D(U(1, 0, 0, 2)) @ D(U(0, 0, 0, 0))
                B += [4 * np.trace(np.eye(3) - P).real for P in (Pyz, Pzx, Pxy)]
    assert np.allclose(out[:3], E2, rtol=1e-14)
    assert np.allclose(out[3:6], B, rtol=1e-13)
    assert abs(out[8] - EJ) < 1e-13 * abs(EJ) and out[6] == 0 and out[7] == 0


def test_unitarity_and_gauss_brute_force():
    N = (3, 3, 3)
    a = (1.0, 0.8, 0.6)
    f = random_fields(N, seed=5)
    f.U[2, 1, 1, 0] *= 1.001  # one non-unitary link
    f.sync(0)
    um, ur = O.unitarity(f, a)
    norms = np.array([np.linalg.norm(M.T @ M.T.conj().T - np.eye(3)) for M in f.U[1:-1].reshape(-1, 3, 3)])
    assert abs(um - norms.max()) < 1e-15 and abs(ur - np.sqrt((norms ** 2).mean())) < 1e-15
    per_z, gmax, grms = O.gauss(f, a)
    norm = 1.0 / np.prod(a)
    G = np.zeros((N[2], N[1], N[0], 8))
    for z in range(N[2]):
        for y in range(N[1]):
            for x in range(N[0]):
                acc = np.zeros(8)
                for mu, d in enumerate(((1, 0, 0), (0, 1, 0), (0, 0, 1))):
                    xm, ym, zm = x - d[0], y - d[1], z - d[2]
                    Um = mat(f.U, zm, ym, xm, mu, N)
                    Em = f.E[zm + 1, ym % N[1], xm % N[0], mu]
                    rot = Um.conj().T @ herm(Em) @ Um
                    acc += f.E[z + 1, y, x, mu] - np.array([2 * np.trace(T[k] @ rot).real for k in range(8)])
                G[z, y, x] = norm * (acc - (f.JpS[z + 1, y, x] + f.JmS[z + 1, y, x]) / np.sqrt(2))
    assert np.allclose(per_z, (G ** 2).sum(axis=(1, 2, 3)) / (N[0] * N[1]), rtol=1e-12)
    assert abs(gmax - np.abs(G).max()) < 1e-14 and abs(grms - np.sqrt((G ** 2).mean())) < 1e-14


@pytest.mark.parametrize("a", [A1, (1.0, 0.9, 0.7)])
def test_emt_sums_equal_bulk(a):
    N = (4, 3, 5)
    f = random_fields(N, seed=6, bc=1)
    f.UOld[:] = f.U
    emt = O.emt(f, a, bc=1)
    bulk = O.bulk_T(O.bulk(f, a, bc=1), a)
    tot = emt.sum(axis=0) * N[0] * N[1]
    for k, name in enumerate(("T00", "TXX", "TYY", "TZZ")):
        assert abs(tot[k] - bulk[name]) < 1e-12 * abs(bulk["T00"]), name


def test_emt_uniform_electric_field():
    N = (3, 3, 4)
    a = (1.0, 0.9, 0.7)
    f = O.Fields(N)
    rng = np.random.default_rng(7)
    e = rng.normal(size=(3, 8))
    f.E[...] = e
    emt = O.emt(f, a)
    q = np.array(a) / np.prod(a)
    es = (e ** 2).sum(axis=1) * q ** 2
    ref = [0.5 * es.sum(), 0.5 * (es[1] + es[2] - es[0]), 0.5 * (es[0] + es[2] - es[1]), 0.5 * (es[0] + es[1] - es[2]),
           0, 0, 0, q[0] * q[1] * e[0] @ e[1], q[1] * q[2] * e[1] @ e[2], q[2] * q[0] * e[2] @ e[0]]
    assert np.allclose(emt, np.tile(ref, (N[2], 1)), rtol=1e-13, atol=1e-15)


def _emt_brute_force(f, a):
    """EnergyMomentumTensor::CreateMap (SC/OBSERVABLES/EnergyMomentumTensor.cpp:92-170) in numpy, one site at a
    time, written from the reference's macros rather than from the oracle:
    * B: the four plaquettes around the site in each plane, on UOld, averaged (COMPUTE_ELEMENTARY /
      NEIGHBORING / EXTENDED_PLAQUETTES, SC/CALC/Plaquettes.cpp:83-97; AvgMatrix, SC/SUNc/AdvancedOperations.cpp:35);
      B_mu^a = 2 Re tr(i t^a avg), B_mu^2 = 4 Re tr(1 - avg) (SC/CALC/AvgFieldStrength.cpp:68-82);
    * E: the vector is the mean of E_mu(x) and U_mu(x-mu)^dag E_mu(x-mu) U_mu(x-mu) with the CURRENT links (the
      macro's U is CreateMap's U), E^2 the mean of the two squares (SC/CALC/AvgFieldStrength.cpp:86-105);
    * T^{mu nu} with q_mu = a_mu / a^3, per-site terms divided by Nx Ny and summed per z."""
    Nx, Ny, Nz = f.N
    N = f.N
    q = np.array(a) / np.prod(a)
    D = lambda M: M.conj().T  # noqa: E731
    out = np.zeros((Nz, 10))
    for z in range(Nz):
        for y in range(Ny):
            for x in range(Nx):
                V = lambda dx, dy, dz, mu: mat(f.UOld, z + dz, y + dy, x + dx, mu, N)  # noqa: E731
                W = lambda dx, dy, dz, mu: mat(f.U, z + dz, y + dy, x + dx, mu, N)  # noqa: E731
                Ev = lambda dx, dy, dz, mu: f.E[z + dz + 1, (y + dy) % Ny, (x + dx) % Nx, mu]  # noqa: E731
                # plane (mu, nu): x-y, y-z, z-x; UUDD = V1 V2 V3^dag V4^dag, UDDU = V1 V2^dag V3^dag V4, ...
                Uxy = V(0, 0, 0, 0) @ V(1, 0, 0, 1) @ D(V(0, 1, 0, 0)) @ D(V(0, 0, 0, 1))
                Uyz = V(0, 0, 0, 1) @ V(0, 1, 0, 2) @ D(V(0, 0, 1, 1)) @ D(V(0, 0, 0, 2))
                Uzx = V(0, 0, 0, 2) @ V(0, 0, 1, 0) @ D(V(1, 0, 0, 2)) @ D(V(0, 0, 0, 0))
                UMxy = V(0, 0, 0, 1) @ D(V(-1, 1, 0, 0)) @ D(V(-1, 0, 0, 1)) @ V(-1, 0, 0, 0)
                UMyz = V(0, 0, 0, 2) @ D(V(0, -1, 1, 1)) @ D(V(0, -1, 0, 2)) @ V(0, -1, 0, 1)
                UMzx = V(0, 0, 0, 0) @ D(V(1, 0, -1, 2)) @ D(V(0, 0, -1, 0)) @ V(0, 0, -1, 2)
                UxMy = D(V(0, -1, 0, 1)) @ V(0, -1, 0, 0) @ V(1, -1, 0, 1) @ D(V(0, 0, 0, 0))
                UyMz = D(V(0, 0, -1, 2)) @ V(0, 0, -1, 1) @ V(0, 1, -1, 2) @ D(V(0, 0, 0, 1))
                UzMx = D(V(-1, 0, 0, 0)) @ V(-1, 0, 0, 2) @ V(-1, 0, 1, 0) @ D(V(0, 0, 0, 2))
                UMxMy = D(V(-1, 0, 0, 0)) @ D(V(-1, -1, 0, 1)) @ V(-1, -1, 0, 0) @ V(0, -1, 0, 1)
                UMyMz = D(V(0, -1, 0, 1)) @ D(V(0, -1, -1, 2)) @ V(0, -1, -1, 1) @ V(0, 0, -1, 2)
                UMzMx = D(V(0, 0, -1, 2)) @ D(V(-1, 0, -1, 0)) @ V(-1, 0, -1, 2) @ V(-1, 0, 0, 0)
                avg = [0.25 * (Uyz + UMyz + UyMz + UMyMz), 0.25 * (Uzx + UMzx + UzMx + UMzMx),
                       0.25 * (Uxy + UMxy + UxMy + UMxMy)]
                B = [np.array([2.0 * np.trace(1j * T[c] @ M).real for c in range(8)]) for M in avg]
                B2 = [4.0 * (3.0 - np.trace(M).real) for M in avg]
                Evec, E2 = [], []
                for mu, d in enumerate(((1, 0, 0), (0, 1, 0), (0, 0, 1))):
                    Um = W(-d[0], -d[1], -d[2], mu)
                    em = Ev(-d[0], -d[1], -d[2], mu)
                    rot = D(Um) @ herm(em) @ Um
                    down = np.array([2.0 * np.trace(T[c] @ rot).real for c in range(8)])
                    Evec.append(0.5 * (Ev(0, 0, 0, mu) + down))
                    E2.append(0.5 * ((Ev(0, 0, 0, mu) ** 2).sum() + (em ** 2).sum()))
                e2 = [q[m] ** 2 * E2[m] for m in range(3)]
                b2 = [q[m] ** 2 * B2[m] for m in range(3)]
                t = [0.5 * (sum(e2) + sum(b2)),
                     0.5 * ((e2[1] + e2[2] + b2[1] + b2[2]) - (e2[0] + b2[0])),
                     0.5 * ((e2[0] + e2[2] + b2[0] + b2[2]) - (e2[1] + b2[1])),
                     0.5 * ((e2[0] + e2[1] + b2[0] + b2[1]) - (e2[2] + b2[2])),
                     q[1] * q[2] * Evec[1] @ B[2] - q[2] * q[1] * Evec[2] @ B[1],
                     q[2] * q[0] * Evec[2] @ B[0] - q[0] * q[2] * Evec[0] @ B[2],
                     q[0] * q[1] * Evec[0] @ B[1] - q[1] * q[0] * Evec[1] @ B[0],
                     q[0] * q[1] * Evec[0] @ Evec[1] + q[0] * q[1] * B[0] @ B[1],
                     q[1] * q[2] * Evec[1] @ Evec[2] + q[1] * q[2] * B[1] @ B[2],
                     q[2] * q[0] * Evec[2] @ Evec[0] + q[2] * q[0] * B[2] @ B[0]]
                out[z] += np.array(t) / (Nx * Ny)
    return out


@pytest.mark.parametrize("N,a,bc", [((3, 4, 5), (1.0, 0.9, 0.7), 0), ((4, 3, 4), (1.0, 1.0, 1.0), 1),
                                    ((3, 3, 4), (0.8, 1.0, 0.5), 0)])
def test_emt_brute_force(N, a, bc):
    """All ten components per layer against the brute force, with UOld independent of U (so the clovers on UOld,
    the transport with the current U and the layer each clover belongs to are all exercised), open and periodic
    z, anisotropic spacings and Nx != Ny.  Each of these slips in the restatement moves a component by 0.7 % or
    more of the largest: clovers on U, transport with UOld, the sign of T0X, one plaquette of the z-x clover a
    layer off, E(x - z) read from the same layer."""
    f = random_fields(N, seed=21, bc=bc)
    old = random_fields(N, seed=22, bc=bc)
    f.UOld[:] = old.U
    f.sync(bc)
    ref = _emt_brute_force(f, a)
    got = O.emt(f, a, bc=bc)
    scale = np.abs(ref).max()
    assert np.abs(ref[:, 4:]).max() > 1e-2 * scale  # the off-diagonal components are not trivially small
    assert np.abs(got - ref).max() < 1e-13 * scale, np.abs(got - ref).max(axis=0) / scale


# ------------------------------------------------------------------------------------------------
# MV initial conditions
# ------------------------------------------------------------------------------------------------
def lattice_p2(n, N, a):
    return (2 - 2 * np.cos(2 * np.pi * n / N)) / a ** 2


def test_poisson_single_mode():
    N = (12, 8, 1)
    a = (1.0, 0.8, 1.0)
    x, y = np.meshgrid(np.arange(N[0]), np.arange(N[1]))
    rho = np.zeros((N[1], N[0], 8))
    rho[..., 3] = np.cos(2 * np.pi * (2 * x / N[0] + 1 * y / N[1]))
    out = np.ascontiguousarray(rho.copy())
    O.lib().ref_solve_laplace(O.ivec(N), O.dvec(a), O.P(out))
    p2 = lattice_p2(2, N[0], a[0]) + lattice_p2(1, N[1], a[1])
    assert np.abs(out - rho / p2).max() < 1e-14


def _wilson(N, a, xi, D, center=3, sigma=2.0, lam=1.65, meff=0.3, mult=1.0):
    Nz = N[2]
    rhoOld = np.zeros((Nz + 2, N[1], N[0], 8))
    rho = np.zeros_like(rhoOld)
    V = np.zeros((Nz + 2, N[1], N[0], 3, 3), complex)
    V[..., 0, 0] = V[..., 1, 1] = V[..., 2, 2] = 1
    O.lib().ref_set_wilson_lines(O.ivec(N), O.dvec(a), 0, lam, meff, sigma, mult, O.P(np.ascontiguousarray(xi)),
                                 D.encode(), center, O.P(rhoOld), O.P(rho), O.P(V))
    return rhoOld, rho, V


def test_charge_profile_and_smoothing_single_mode():
    N = (8, 6, 10)
    a = (1.0, 0.9, 1.0)
    x, y = np.meshgrid(np.arange(N[0]), np.arange(N[1]))
    xi = np.zeros((N[1], N[0], 8))
    xi[..., 5] = np.cos(2 * np.pi * (1 * x / N[0] + 2 * y / N[1]))
    lam, meff, sigma, center = 1.3, 0.4, 2.0, 4
    rhoOld, rho, _ = _wilson(N, a, xi, "R", center, sigma, lam, meff, mult=1.5)
    z = np.arange(N[2])
    T = np.exp(-0.5 * ((z - center) / sigma) ** 2) / np.sqrt(2 * np.pi * sigma ** 2)
    raw = 1.5 * np.sqrt(a[0] * a[1]) * T[:, None, None, None] * xi[None]
    assert np.abs(rhoOld[1:-1] - raw).max() < 1e-15
    p2 = lattice_p2(1, N[0], a[0]) + lattice_p2(2, N[1], a[1])
    filt = np.exp(-p2 / (2 * lam ** 2)) * p2 / (p2 + meff ** 2)
    assert np.abs(rho[1:-1] - raw * filt).max() < 1e-14


@pytest.mark.parametrize("D", ["L", "R"])
def test_wilson_line_against_expm(D):
    N = (6, 4, 5)
    a = A1
    xl, _ = O.draw_charges(4, N[0], N[1], 0.8, 0.8)
    _, rho, V = _wilson(N, a, xl, D)
    # Poisson solve of each slice by numpy FFT, then the path-ordered product with scipy expm
    W = np.eye(3, dtype=complex)
    zs = range(N[2]) if D == "R" else range(N[2] - 1, -1, -1)
    sgn = 1.0 if D == "R" else -1.0
    kx = lattice_p2(np.arange(N[0]), N[0], 1.0)
    ky = lattice_p2(np.arange(N[1]), N[1], 1.0)
    p2 = ky[:, None] + kx[None, :]
    p2[0, 0] = np.inf
    for z in zs:
        A = np.real(np.fft.ifft2(np.fft.fft2(rho[z + 1], axes=(0, 1)) / p2[..., None], axes=(0, 1)))
        W = sl.expm(1j * sgn / np.sqrt(2) * herm(A[1, 2])) @ W
        assert np.abs(V[z + 1, 1, 2].T - W).max() < 1e-13


@pytest.mark.parametrize("D,sgn", [("L", -1.0), ("R", 1.0)])
def test_evolve_rho_single_mode(D, sgn):
    N = (2, 2, 16)
    a = (1.0, 1.0, 0.5)
    dt = 0.1
    z = np.arange(N[2])
    k = 2 * np.pi * 3 / N[2]
    J = np.zeros((N[2] + 2, N[1], N[0], 8))
    J[1:-1, ..., 6] = np.cos(k * z)[:, None, None]
    out = np.zeros_like(J)
    O.lib().ref_evolve_rho(O.ivec(N), O.dvec(a), dt, D.encode(), O.P(J), O.P(out))
    w = np.sqrt(lattice_p2(3, N[2], a[2]))
    # EvolveRhoViaFourier: phase sgn * w * dt on the positive half of the spectrum, the opposite on the other
    ref = np.cos(k * z + sgn * w * dt)
    assert np.abs(out[1:-1, 0, 0, 6] - ref).max() < 1e-14


# ------------------------------------------------------------------------------------------------
# Coulomb gauge fixing (transverse: Nz = 1, where the reference algorithm converges)
# ------------------------------------------------------------------------------------------------
def _identity_G(N):
    G = np.zeros((N[2] + 2, N[1], N[0], 3, 3), complex)
    G[..., 0, 0] = G[..., 1, 1] = G[..., 2, 2] = 1
    return G


def test_coulomb_gauge_recovers_rotated_copy():
    N = (8, 6, 1)
    a = A1
    f = random_fields(N, seed=8, link_scale=0.3)
    UG, EG, st = O.coulomb_gauge(f.copy(), a, _identity_G(N), 1e-13, 300)
    assert st[1] <= 1e-13 and st[0] < 300
    # the gauge-fixed copy, rotated by a random g, is fixed again to a configuration with the same
    # gauge-invariant content and the Coulomb condition
    fc = f.copy()
    fc.U[1:-1] = UG[1:-1]
    fc.E[1:-1] = EG[1:-1]
    fc.sync(0)
    rng = np.random.default_rng(3)
    G = _identity_G(N)
    G[1:-1] = np.swapaxes(random_su3(rng, (N[2], N[1], N[0]), 0.4), -1, -2)
    G[0], G[-1] = G[1], G[-2]
    fr = fc.copy()
    O.gauge_transform(fr, a, G, 0)
    UG2, EG2, st2 = O.coulomb_gauge(fr.copy(), a, _identity_G(N), 1e-13, 300)
    assert st2[1] <= 1e-13
    f2 = fr.copy()
    f2.U[1:-1] = UG2[1:-1]
    f2.E[1:-1] = EG2[1:-1]
    f2.sync(0)
    assert np.allclose(O.bulk(f2, a)[[0, 1, 2, 5]], O.bulk(fc, a)[[0, 1, 2, 5]], rtol=1e-12)
    # the Coulomb-gauge copy is unique up to a global rotation: plaquette traces per site agree
    assert abs(st2[6] - st[6]) < 1e-10 * abs(st[6])


def test_left_sheet_from_charge_is_the_mv_left_mover():
    """The oracle's sheet from a given charge density equals its MV left mover when the density is the one
    SetChargeDensities forms from the same draws (mult * sqrt(a0 a1) * profile(z) * xi)."""
    N, a, dt, c, sig = (6, 5, 12), (1.0, 0.9, 0.8), 0.25, 4, 2.0
    xl, xr = O.draw_charges(3, N[0], N[1], 0.6, 0.6)
    f1 = O.mv_initial_state(1, N, a, dt, xl, xr, lam=1.65, meff=0.2, sigma_l=sig, sigma_r=sig, center_l=c, center_r=c)
    z = np.arange(N[2])
    prof = np.exp(-0.5 * ((z - c) / sig) ** 2) / np.sqrt(2 * np.pi * sig * sig)
    rho = np.sqrt(a[0] * a[1]) * prof[:, None, None, None] * np.asarray(xl).reshape(N[1], N[0], 8)[None]
    f2 = O.left_sheet_from_charge(N, a, dt, rho, lam=1.65, meff=0.2)
    for k in ("U", "UOld", "E", "Jp", "JpS"):
        assert np.abs(getattr(f1, k) - getattr(f2, k)).max() < 1e-13, k


def test_numpy_slab_energy_covering_the_lattice_is_bulk_T00():
    from tests.helpers import random_fields, slab_energy_np
    N, a = (5, 4, 6), (1.0, 0.9, 0.8)
    f = random_fields(N, seed=2)
    O.update(f, a, 0.1, 1)
    T00 = O.bulk_T(O.bulk(f, a), a)["T00"]
    assert abs(slab_energy_np(f, a, 3, 10) - T00) < 1e-12 * abs(T00)
    # a slab is the sum of its layers
    parts = sum(slab_energy_np(f, a, z, 1) for z in range(N[2]))
    assert abs(parts - T00) < 1e-12 * abs(T00)


# ------------------------------------------------------------------------------------------------
# Output-file observables: IO::SaveFields, IO::SaveTransverselyIntegratedCurrents,
# EnergyMomentumTensor::CreateMapBfCollision
# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("bc", [0, 1])
def test_field_profile_against_numpy_and_bulk(bc):
    from tests.helpers import site_fields_np
    N = (4, 3, 5)
    a = (1.0, 0.9, 0.7)
    f = random_fields(N, seed=21, bc=bc)
    prof = O.field_profile(f, a, bc)
    E2, B = site_fields_np(f)
    P = N[0] * N[1]
    ref = np.stack([0.5 * (B[..., 0] + B[..., 1]).sum(axis=(1, 2)) / P, 0.5 * B[..., 2].sum(axis=(1, 2)) / P,
                    0.5 * (E2[..., 0] + E2[..., 1]).sum(axis=(1, 2)) / P, 0.5 * E2[..., 2].sum(axis=(1, 2)) / P], axis=1)
    assert np.allclose(prof, ref, rtol=1e-13, atol=0)
    raw = O.bulk(f, a, bc)  # the layer sums add up to the bulk sums
    tot = 2 * P * prof.sum(axis=0)
    assert np.allclose(tot, [raw[3] + raw[4], raw[5], raw[0] + raw[1], raw[2]], rtol=1e-13)


def test_current_profile_rows_and_columns():
    N = (3, 4, 5)
    f = random_fields(N, seed=22)
    for k, A in enumerate((f.Jp, f.Jm, f.JpS, f.JmS)):  # distinct scales expose a swapped column
        A *= 1.0 + k
    f.Jm[0] *= 3.0  # the ghost layers enter as stored
    out = O.current_profile(f)
    P = N[0] * N[1]
    for r in range(N[2] + 2):
        ref = [0.5 * (A[r] ** 2).sum() / P for A in (f.Jm, f.JmS, f.Jp, f.JpS)]
        assert np.allclose(out[r], ref, rtol=1e-14), r


@pytest.mark.parametrize("Nz", [6, 7])
def test_transverse_map_against_numpy_and_bulk(Nz):
    from tests.helpers import site_fields_np
    N = (4, 3, Nz)
    a = (1.0, 0.9, 0.7)
    f = random_fields(N, seed=23)
    m = O.transverse_map(f, a)
    E2, B = site_fields_np(f)
    acube = np.prod(a)
    c = (np.array(a) / acube) ** 2
    dens = 0.5 * ((E2 + B) * c).sum(axis=-1) / Nz  # (Nz, Ny, Nx)
    zs = (Nz - 2) // 2  # the last left layer
    assert np.allclose(m[..., 0], dens[: zs + 1].sum(axis=0), rtol=1e-13, atol=0)
    assert np.allclose(m[..., 1], dens[zs + 1:].sum(axis=0), rtol=1e-13, atol=0)
    T = O.bulk_T(O.bulk(f, a), a)
    assert abs(m.sum() * Nz - T["T00"]) < 1e-12 * T["T00"]
