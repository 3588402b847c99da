"""Shared test helpers: seeded SU(3) configurations in the reference layout and comparison utilities."""
from __future__ import annotations

import numpy as np

from oracle import oracle as O
from soerenpragya3dglasma_amd import layout as L

# Gell-Mann matrices / 2 (the reference's t^a, SC/SUNc/SU3/AlgebraOperations.cpp:6-45)
_lam = np.zeros((8, 3, 3), complex)
_lam[0][0, 1] = _lam[0][1, 0] = 1
_lam[1][0, 1] = -1j
_lam[1][1, 0] = 1j
_lam[2][0, 0] = 1
_lam[2][1, 1] = -1
_lam[3][0, 2] = _lam[3][2, 0] = 1
_lam[4][0, 2] = -1j
_lam[4][2, 0] = 1j
_lam[5][1, 2] = _lam[5][2, 1] = 1
_lam[6][1, 2] = -1j
_lam[6][2, 1] = 1j
_lam[7] = np.diag([1, 1, -2]) / np.sqrt(3)
T = _lam / 2


def herm(alpha):
    return np.einsum("...a,aij->...ij", alpha, T)


def random_su3(rng, shape, scale):
    """exp(i alpha^a t^a) with Gaussian alpha of width `scale` (via eigendecomposition in numpy)."""
    al = rng.normal(size=tuple(shape) + (8,)) * scale
    H = herm(al)
    w, v = np.linalg.eigh(H)
    return np.einsum("...ij,...j,...kj->...ik", v, np.exp(1j * w), v.conj())


def random_fields(N, seed=0, link_scale=0.6, e_scale=0.3, j_scale=0.2, bc=0) -> O.Fields:
    rng = np.random.default_rng(seed)
    Nx, Ny, Nz = N
    f = O.Fields(N)
    M = random_su3(rng, (Nz, Ny, Nx, 3), link_scale)  # [..., i, j]
    f.U[1:-1] = np.swapaxes(M, -1, -2)  # storage [..., j, i]
    f.E[1:-1] = rng.normal(size=f.E[1:-1].shape) * e_scale
    for J in (f.Jp, f.Jm, f.JpS, f.JmS):
        J[1:-1] = rng.normal(size=J[1:-1].shape) * j_scale
    f.sync(bc)
    return f


def upload(g, f: O.Fields):
    g.set_fields(f.U, f.E, f.Jp, f.Jm, f.JpS, f.JmS)


def relerr(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def fields_maxdiff(d: dict, f: O.Fields) -> dict:
    return {
        "U": float(np.abs(d["U"] - f.U).max()),
        "E": float(np.abs(d["E"] - f.E).max()),
        "Jp": float(np.abs(d["Jp"] - f.Jp).max()),
        "Jm": float(np.abs(d["Jm"] - f.Jm).max()),
        "JpS": float(np.abs(d["JpS"] - f.JpS).max()),
        "JmS": float(np.abs(d["JmS"] - f.JmS).max()),
    }


_hc = None


def hostcheck_lib():
    """The engine's per-site code compiled for the host (libglasma_hostcheck.so), built on demand with g++."""
    global _hc
    if _hc is None:
        import ctypes
        import os
        from soerenpragya3dglasma_amd import build as B
        path = os.path.join(B.OUT, "libglasma_hostcheck.so")
        if not os.path.exists(path):
            B.build_hostcheck()
        lib = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        cd = ctypes.c_double
        for name, args in {"hc_su3_expi": [cd, dp, dp], "hc_su3_expi_m1": [cd, dp, dp], "hc_su3_ilog": [cd, dp, dp],
                           "hc_su3_adj_mul": [dp, dp, dp], "hc_su3_inv_adj_mul": [dp, dp, dp],
                           "hc_su3_retrigenu": [dp, dp],
                           "hc_step": [ctypes.POINTER(ctypes.c_int), dp, ctypes.c_int, cd, dp, dp, dp, dp, dp, dp]}.items():
            getattr(lib, name).argtypes = args
        _hc = lib
    return _hc


def slab_energy_np(f, a, zc, half):
    """Energy of the layers |z - zc| < half from host fields in the reference layout (numpy, independent of the
    engine): 1/2 sum over those sites of sum_mu c_mu (E_mu^a E_mu^a + 4 Re tr(1 - P_mu)), c_mu = (a_mu / a^3)^2,
    P_mu the plaquette in the plane (mu+1, mu+2) rooted at the site, z neighbours from the ghost layers."""
    Nx, Ny, Nz = f.N
    U = np.swapaxes(f.U, -1, -2)  # [z+1, y, x, mu, i, j]
    acube = a[0] * a[1] * a[2]
    c = [(a[m] / acube) ** 2 for m in range(3)]

    def at(mu, zg, d=None):
        """U_mu on layer zg, shifted by +1 in direction d (0: x, 1: y, 2: z)."""
        if d == 2:
            return U[zg + 1, :, :, mu]
        M = U[zg, :, :, mu]
        if d == 0:
            return np.roll(M, -1, axis=1)
        if d == 1:
            return np.roll(M, -1, axis=0)
        return M

    dag = lambda M: np.conj(np.swapaxes(M, -1, -2))
    tot = 0.0
    for z in range(Nz):
        if not abs(z - zc) < half:
            continue
        zg = z + 1
        for mu in range(3):
            j, k = (mu + 1) % 3, (mu + 2) % 3
            P = at(j, zg) @ at(k, zg, j) @ dag(at(j, zg, k)) @ dag(at(k, zg))
            retr = 3.0 - np.trace(P, axis1=-2, axis2=-1).real
            e2 = (f.E[zg, :, :, mu, :] ** 2).sum(axis=-1)
            tot += 0.5 * c[mu] * float((e2 + 4.0 * retr).sum())
    return tot


def site_fields_np(f):
    """Per-site squared fields from host fields in the reference layout (numpy, independent of the engine):
    (E2, B), each shaped (Nz, Ny, Nx, 3); E2[..., mu] = sum_a E_mu^a E_mu^a and B[..., mu] = 4 Re tr(1 - P_mu)
    with P_mu the elementary plaquette in the plane (mu+1, mu+2) rooted at the site (z neighbours from the
    ghost layers)."""
    Nx, Ny, Nz = f.N
    U = np.swapaxes(f.U, -1, -2)  # [z+1, y, x, mu, i, j]
    dag = lambda M: np.conj(np.swapaxes(M, -1, -2))  # noqa: E731

    def at(mu, d=None):
        if d == 2:
            return U[2:, :, :, mu]
        M = U[1:-1, :, :, mu]
        if d == 0:
            return np.roll(M, -1, axis=2)
        if d == 1:
            return np.roll(M, -1, axis=1)
        return M

    B = np.zeros((Nz, Ny, Nx, 3))
    for mu in range(3):
        j, k = (mu + 1) % 3, (mu + 2) % 3
        P = at(j) @ at(k, j) @ dag(at(j, k)) @ dag(at(k))
        B[..., mu] = 4.0 * (3.0 - np.trace(P, axis1=-2, axis2=-1).real)
    E2 = (f.E[1:-1] ** 2).sum(axis=-1)
    return E2, B
