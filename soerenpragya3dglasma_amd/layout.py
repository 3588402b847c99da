"""Conversions between the reference's array-of-structs field layout and the engine's device layout.

Reference layout (SC/LATTICE/GaugeLinks.cpp:27-36, ElectricFields.cpp:25-35, ChargeDensity.cpp:27-37), with the
two z ghost layers:

* U: complex128 (Nz+2, Ny, Nx, 3, 3, 3) indexed [z+1, y, x, mu, j, i] (flat = 9*(mu + 3*Index3D) + i + 3*j)
* E: float64 (Nz+2, Ny, Nx, 3, 8);  J: float64 (Nz+2, Ny, Nx, 8)

Device layout (no ghosts, s = x + Nx*(y + Ny*z)): tiles of 64 sites; planes grouped in pairs, pair q of a
tile stored contiguously with the pair members adjacent: plane p of site s at
((s // 64) * NP/2 + p // 2) * 128 + (s % 64) * 2 + p % 2, S padded to a multiple of 64.

* U: NP = 54: plane (mu*18 + 2k + c), k = i + 3j, c = 0 real / 1 imaginary
* E: NP = 24: plane (mu*8 + a);  J: NP = 32: plane (w*8 + a), w = J+, J-, J+stat, J-stat

The ``*_soa`` functions below produce plane-major arrays (NP, S); ``tile``/``untile`` convert those to and
from the tiled device layout.
"""
from __future__ import annotations

import numpy as np


def u_aos_to_soa(U: np.ndarray) -> np.ndarray:
    Nz2, Ny, Nx = U.shape[:3]
    inner = U[1:-1].reshape(-1, 3, 9)  # (S, mu, k)
    re = inner.real
    im = inner.imag
    out = np.empty((3, 9, 2, inner.shape[0]))
    out[:, :, 0, :] = re.transpose(1, 2, 0)
    out[:, :, 1, :] = im.transpose(1, 2, 0)
    return np.ascontiguousarray(out.reshape(54, -1))


def u_soa_to_aos(Us: np.ndarray, N, bc: int = 0) -> np.ndarray:
    Nx, Ny, Nz = N
    S = Nx * Ny * Nz
    t = Us.reshape(3, 9, 2, S)
    c = (t[:, :, 0, :] + 1j * t[:, :, 1, :]).transpose(2, 0, 1)  # (S, mu, k)
    U = np.empty((Nz + 2, Ny, Nx, 3, 3, 3), np.complex128)
    U[1:-1] = c.reshape(Nz, Ny, Nx, 3, 3, 3)
    sync_ghosts(U, bc)
    return U


def v_aos_to_soa(A: np.ndarray, rec: int) -> np.ndarray:
    """E (rec=24) or J (rec=8) from AoS with ghosts to (rec, S)."""
    inner = A[1:-1].reshape(-1, rec)
    return np.ascontiguousarray(inner.T)


def v_soa_to_aos(As: np.ndarray, N, shape_tail, bc: int = 0) -> np.ndarray:
    Nx, Ny, Nz = N
    out = np.empty((Nz + 2, Ny, Nx) + tuple(shape_tail))
    out[1:-1] = As.T.reshape((Nz, Ny, Nx) + tuple(shape_tail))
    sync_ghosts(out, bc)
    return out


def sync_ghosts(A: np.ndarray, bc: int = 0) -> None:
    """SynchronizeGhostCells: open copies the edge layers, periodic wraps (SC/LATTICE/GaugeLinks.cpp:59-73)."""
    if bc == 0:
        A[0] = A[1]
        A[-1] = A[-2]
    else:
        A[0] = A[-2]
        A[-1] = A[1]


def su3_flat(M: np.ndarray) -> np.ndarray:
    """3x3 matrix -> reference storage order (column-major, i + 3j)."""
    return np.ascontiguousarray(np.asarray(M).T).reshape(9)


def su3_unflat(v: np.ndarray) -> np.ndarray:
    return np.asarray(v).reshape(3, 3).T


TILE = 64


def tile(soa: np.ndarray) -> np.ndarray:
    """(NP, S) plane-major -> flat tiled device layout (S padded to a multiple of 64)."""
    NP, S = soa.shape
    Sp = (S + TILE - 1) // TILE * TILE
    pad = np.zeros((NP, Sp))
    pad[:, :S] = soa
    # (pair, member, tile, lane) -> (tile, pair, lane, member)
    t = pad.reshape(NP // 2, 2, Sp // TILE, TILE).transpose(2, 0, 3, 1)
    return np.ascontiguousarray(t).reshape(-1)


def untile(flat: np.ndarray, NP: int, S: int) -> np.ndarray:
    Sp = (S + TILE - 1) // TILE * TILE
    t = flat.reshape(Sp // TILE, NP // 2, TILE, 2).transpose(1, 3, 0, 2).reshape(NP, Sp)
    return np.ascontiguousarray(t[:, :S])
