"""The block -> tile orders of the step (site.h tile_site, host build): each visits every 64-site tile once.  CPU
only; the GPU tuner tests check that a tuned order evolves the fields bit for bit like the default one."""
import ctypes

import numpy as np
import pytest

from tests.helpers import hostcheck_lib


@pytest.mark.parametrize("N", [(256, 32, 4), (128, 64, 3), (96, 96, 8), (48, 48, 16), (32, 32, 8)])
def test_tile_orders_visit_every_tile_once(N):
    """Every block -> tile order the step and its autotuner use (site.h tile_site; the fields tile_map fills for a
    lattice) visits each 64-site tile exactly once: linear, plane bands, column bands of several shapes, z slabs."""
    hc = hostcheck_lib()
    hc.hc_tile_sites.argtypes = [ctypes.POINTER(ctypes.c_int)] + [ctypes.c_int] * 6 + [ctypes.c_long,
                                                                                      ctypes.POINTER(ctypes.c_long)]
    nx, ny, nz = N
    P, S = nx * ny, nx * ny * nz
    nt = (S + 63) // 64
    orders = [(0, 0, 0, 0, 0, 0)]
    if P % 512 == 0:
        orders.append((1, P // 512, 0, 0, 0, 0))
    if nt % 8 == 0:
        orders.append((3, nt // 8, 0, 0, 0, 0))
    if nx % 64 == 0:
        tpr = nx // 64
        for bw, bh in ((1, 16), (2, 8), (4, 4), (1, 8)):
            if tpr % bw == 0 and ny % bh == 0 and (tpr // bw) * (ny // bh) % 8 == 0:
                orders.append((2, 0, bw, bh, tpr // bw, bw * bh))
    n = (ctypes.c_int * 3)(*N)
    for mode, bpb, bw, bh, nbx, tpbp in orders:
        out = np.zeros(nt, dtype=np.int64)
        hc.hc_tile_sites(n, mode, bpb, bw, bh, nbx, tpbp, nt, out.ctypes.data_as(ctypes.POINTER(ctypes.c_long)))
        assert np.array_equal(np.sort(out), np.arange(nt) * 64), (N, mode, bw, bh)
