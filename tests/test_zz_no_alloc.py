"""GPU test of glasma_bulk_noalloc, the entry point that allocates nothing (its block partials live in a fixed
device array of the library): called through the C ABI, it returns the very sums glasma_bulk returns (bit for
bit), on two calls in a row, the second on a smaller lattice; it leaves the context's own scratch unallocated;
and it refuses a lattice whose partials exceed the fixed array."""
import ctypes

import numpy as np
import pytest

from tests.helpers import random_fields, upload

pytestmark = pytest.mark.gpu


def _lib():
    from soerenpragya3dglasma_amd import _lib as L
    lib = L.lib()
    lib.glasma_bulk_noalloc.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
    lib.glasma_bulk_noalloc.restype = ctypes.c_int
    return lib


def _noalloc(lib, g):
    out = np.zeros(9)
    rc = lib.glasma_bulk_noalloc(g._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 0, lib.glasma_last_error(g._h)
    return out


def _memory(lib, g):
    b = ctypes.c_size_t()
    assert lib.glasma_memory_bytes(g._h, ctypes.byref(b)) == 0
    return b.value


def test_noalloc_bulk_matches_bulk_two_calls_second_smaller():
    from soerenpragya3dglasma_amd import Glasma
    lib = _lib()
    cases = [((32, 24, 40), (1.0, 0.9, 0.7), 11), ((8, 6, 5), (1.0, 1.0, 1.0), 12)]
    for N, a, seed in cases:
        with Glasma(N, a, 0.1) as g:
            upload(g, random_fields(N, seed=seed))
            g.Update(2)
            before = _memory(lib, g)
            got = _noalloc(lib, g)
            assert _memory(lib, g) == before  # the context's reduction buffer was not allocated
            ref = g.bulk_raw()
        print(N, got, ref)
        assert np.array_equal(got, ref), (N, got - ref)


def test_noalloc_bulk_after_other_observables_and_steps():
    """The array keeps the last call's contents; a later call on other fields still returns its own sums."""
    from soerenpragya3dglasma_amd import Glasma
    lib = _lib()
    N = (16, 16, 12)
    with Glasma(N, (1.0, 1.0, 1.0), 0.25) as g:
        upload(g, random_fields(N, seed=13))
        first = _noalloc(lib, g)
        g.emt_profile()
        g.Update(3)
        second = _noalloc(lib, g)
        assert np.array_equal(second, g.bulk_raw())
        assert not np.array_equal(first, second)


def test_noalloc_bulk_refuses_oversized_lattice():
    from soerenpragya3dglasma_amd import Glasma
    lib = _lib()
    N = (256, 256, 512)  # 512 * 257 * 9 + 512 * 9 + 9 doubles of partials, over the 2^20 of the fixed array
    with Glasma(N, (1.0, 1.0, 1.0), 0.25) as g:
        out = np.zeros(9)
        rc = lib.glasma_bulk_noalloc(g._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        assert rc == -1
        assert b"scratch" in lib.glasma_last_error(g._h)
