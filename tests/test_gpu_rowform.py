"""GPU parity of the step's forms against the oracle over the lattice shapes that select them.

The row-staged form (link operands on the tile's own rows read from an LDS copy) runs when row_tiles holds:
Nx a multiple of 64 (a tile is part of one row, halo sites at both ends, x wrapping across tiles), a divisor of 64
(several rows per tile, x wrapping inside the tile; the last tile is padded when the site count is not a multiple
of 64), or every 64-site tile spans at most two row segments (Nx >= 64 - gcd(64, Nx) with no partial last tile:
Nx = 48, 72, 80, 96, ...; four halo slots: the left of each segment's first site and the right of its last, with
x wrapping within each row).  Other Nx take the plain form.  Each case runs three steps with anisotropic spacings
and both z boundaries; fields and UOld agree with the oracle to 1e-12 absolute.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import fields_maxdiff, random_fields, upload

pytestmark = pytest.mark.gpu

SHAPES = [
    (64, 3, 5),    # one row per tile, nx = 64: the halo sites wrap onto the tile itself
    (128, 2, 3),   # one row per tile, two tiles per row
    (32, 6, 4),    # two rows per tile
    (16, 5, 3),    # four rows per tile, 240 sites: padded last tile
    (8, 8, 6),     # eight rows per tile
    (4, 3, 7),     # sixteen rows per tile, 84 sites: padded last tile
    (1, 64, 2),    # nx = 1: the x neighbour is the site itself
    (48, 4, 3),    # two row segments per tile (tile starts at x = 0, 16, 32)
    (96, 2, 3),    # two row segments (x = 0: one segment; x = 64: 32 + 32)
    (80, 3, 4),    # two segments or one (x0 = 0, 16: one segment)
    (72, 8, 3),    # gcd(64, 72) = 8: one and two segments
    (40, 8, 2),    # Nx < 64 - gcd: plain form
    (48, 5, 3),    # 720 sites, a partial last tile: plain form
]


@pytest.mark.parametrize("N", SHAPES)
@pytest.mark.parametrize("bc", [0, 1])
def test_forms_match_oracle(N, bc):
    from soerenpragya3dglasma_amd import Glasma
    a, dt = (1.0, 0.85, 0.6), 0.1
    f = random_fields(N, seed=sum(N) + 7 * bc, bc=bc)
    with Glasma(N, a, dt, bc) as g:
        upload(g, f)
        g.Update(3)
        d = g.get_fields()
        uold = g.get_uold()
    O.update(f, a, dt, 3, bc)
    diff = fields_maxdiff(d, f)
    print(N, bc, diff)
    assert max(diff.values()) < 1e-12, (N, bc, diff)
    assert np.abs(uold - f.UOld).max() < 1e-12
