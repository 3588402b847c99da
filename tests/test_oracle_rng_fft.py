"""Pins of the oracle's RNG and FFT, and the engine's host-side charge drawing (libglasma
glasma_draw_charges) against them.  CPU only (glasma_draw_charges is host code).

* MT19937: the published known answer (10000th output of the default-seeded generator = 4123659995,
  Matsumoto & Nishimura's reference code / C++11 std::mt19937) and numpy's legacy RandomState, which
  seeds with the same init_genrand recursion and returns the raw 32-bit words for full-range uint32.
* GSL seeding: gsl_rng_set maps seed 0 to 4357 (documented default seed of GSL's mt19937).
* The Gaussian (gsl_ran_gaussian, Marsaglia polar method) is "parity unpinned": GSL is absent and the
  reference ships no draws.  Checked here only for its distribution and for the draw layout.
* FFT: numpy.fft for every dimension count, batch and both signs (FFTW conventions, unnormalised).
"""
import numpy as np
import pytest

from oracle import oracle as O


def test_mt19937_known_answer():
    assert O.mt_raw(5489, 10000, std_seeding=True)[-1] == 4123659995


@pytest.mark.parametrize("seed", [1, 2, 1234, 2 ** 31 + 5])
def test_mt19937_matches_numpy_legacy(seed):
    ref = np.random.RandomState(seed).randint(0, 2 ** 32, size=2000, dtype=np.uint32)
    assert np.array_equal(O.mt_raw(seed, 2000), ref)


def test_gsl_seed_zero_maps_to_4357():
    assert np.array_equal(O.mt_raw(0, 100), O.mt_raw(4357, 100))


def test_gaussian_distribution():
    x = O.gauss_draws(12345, 400000)
    assert abs(x.mean()) < 5e-3
    assert abs(x.var() - 1.0) < 1e-2
    # fourth moment of a unit normal is 3
    assert abs((x ** 4).mean() - 3.0) < 5e-2


def test_draw_charges_layout_and_engine_parity():
    """Left nucleus first, then right, each in (y, x, a) order from one stream; the engine's host RNG is
    bit-identical to the oracle's."""
    from soerenpragya3dglasma_amd import Glasma
    Nx, Ny = 5, 3
    xl, xr = O.draw_charges(7, Nx, Ny, 0.5, 0.25)
    g = O.gauss_draws(7, 2 * 8 * Nx * Ny)
    assert np.array_equal(xl.reshape(-1), 0.5 * g[: 8 * Nx * Ny])
    assert np.array_equal(xr.reshape(-1), 0.25 * g[8 * Nx * Ny:])
    for seed in (0, 1, 2, 12345):
        el, er = Glasma.draw_charges(seed, 16, 12, 0.5, 0.7)
        ol, orr = O.draw_charges(seed, 16, 12, 0.5, 0.7)
        assert np.array_equal(el, ol) and np.array_equal(er, orr)


@pytest.mark.parametrize("dims,batch", [((8,), 1), ((12,), 3), ((6, 10), 8), ((4, 6, 5), 8), ((9, 7, 3), 2), ((64,), 1)])
def test_oracle_fft_matches_numpy(dims, batch):
    rng = np.random.default_rng(sum(dims) + batch)
    shape = tuple(reversed(dims)) + (batch,)
    x = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    axes = tuple(range(len(dims)))
    f = O.fft_nd(x, dims, batch, -1)
    b = O.fft_nd(x, dims, batch, +1)
    assert np.abs(f - np.fft.fftn(x, axes=axes)).max() < 1e-12 * np.abs(f).max()
    assert np.abs(b - np.fft.ifftn(x, axes=axes) * np.prod(dims)).max() < 1e-12 * np.abs(b).max()
