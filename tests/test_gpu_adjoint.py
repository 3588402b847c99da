"""GPU tests of the reverse-mode derivatives (glasma_step_vjp, glasma_slab_energy*, the torch Function).

The reference has no gradient code, so the derivatives are checked against central finite differences of
the engine's own forward map: one step at eps = 1e-6 to 1e-7 relative; the whole charge -> sheet -> steps
-> energy map along 32 seeded random directions to 1e-6 relative with Richardson-extrapolated central
differences (eps = 1e-3 and 5e-4).  Plain central differences at eps = 1e-6 carry the forward map's own
round-off (about 1e-12 absolute in the energy, from the initial condition's path-ordered products and
FFTs), so they are held to 1e-4 only.  The slab energy itself is checked against the oracle's bulk T00
(a slab covering every layer).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import random_fields, upload

pytestmark = pytest.mark.gpu

KEYS = ("U", "E", "Jp", "Jm", "JpS", "JmS")


def _dot(a, b):
    return float(sum(np.vdot(a[k][1:-1].view(np.float64) if a[k].dtype == complex else a[k][1:-1],
                             b[k][1:-1].view(np.float64) if b[k].dtype == complex else b[k][1:-1]).real for k in KEYS))


@pytest.mark.parametrize("N,a,bc", [((6, 5, 7), (1.0, 0.9, 0.7), 0), ((4, 4, 4), (1.0, 1.0, 1.0), 1), ((5, 4, 1), (1.0, 1.0, 1.0), 0)])
def test_step_vjp_against_finite_differences(N, a, bc):
    from soerenpragya3dglasma_amd import Glasma
    f = random_fields(N, seed=3, bc=bc)
    rng = np.random.default_rng(4)
    base = {k: getattr(f, k).copy() for k in KEYS}
    w = {k: rng.normal(size=base[k].shape) + (1j * rng.normal(size=base[k].shape) if k == "U" else 0) for k in KEYS}
    v = {k: rng.normal(size=base[k].shape) + (1j * rng.normal(size=base[k].shape) if k == "U" else 0) for k in KEYS}
    with Glasma(N, a, 0.1, bc) as g:
        def L(state):
            g.set_fields(*[np.ascontiguousarray(state[k]) for k in KEYS])
            g.Update(1)
            return _dot(w, g.get_fields())
        eps = 1e-6
        fd = (L({k: base[k] + eps * v[k] for k in KEYS}) - L({k: base[k] - eps * v[k] for k in KEYS})) / (2 * eps)
        g.set_fields(*[np.ascontiguousarray(base[k]) for k in KEYS])
        vj = g.step_vjp(w)
    ad = _dot(vj, v)
    print(N, bc, fd, ad)
    assert abs(fd - ad) < 1e-7 * abs(fd)


def test_slab_energy_covering_the_lattice_is_bulk_T00():
    from soerenpragya3dglasma_amd import Glasma
    N, a = (6, 4, 8), (1.0, 0.9, 0.7)
    f = random_fields(N, seed=5)
    with Glasma(N, a, 0.1) as g:
        upload(g, f)
        e = g.slab_energy(4, 100)
        e_part = g.slab_energy(4, 2)
    ref = O.bulk_T(O.bulk(f, a), a)["T00"]
    assert abs(e - ref) < 1e-12 * ref
    assert 0 < e_part < e


def _left_mover_case(N=(8, 8, 16), seed=1):
    from soerenpragya3dglasma_amd import Glasma
    xl, _ = O.draw_charges(seed, N[0], N[1], 0.5, 0.5)
    z = np.arange(N[2])
    prof = np.exp(-0.5 * ((z - 6) / 2.0) ** 2) / np.sqrt(2 * np.pi * 4.0)
    rho = prof[:, None, None, None] * xl[None]
    mv = Glasma.mv_params(lam=1.65, meff=0.1, sigma_l=2.0, sigma_r=2.0, center_l=6, center_r=8)
    return rho, mv


def test_energy_from_charge_matches_mv_initial_state():
    """The sheet built from the charge SetChargeDensities would produce equals glasma_mv_initial_state's
    left mover, so the energies after the same steps agree."""
    from soerenpragya3dglasma_amd import Glasma
    N = (8, 8, 16)
    xl, xr = O.draw_charges(1, N[0], N[1], 0.5, 0.5)
    mv = Glasma.mv_params(lam=1.65, meff=0.1, sigma_l=2.0, sigma_r=2.0, center_l=6, center_r=8)
    z = np.arange(N[2])
    rho = (np.exp(-0.5 * ((z - 6) / 2.0) ** 2) / np.sqrt(2 * np.pi * 4.0))[:, None, None, None] * xl[None]
    with Glasma(N, (1.0, 1.0, 1.0), 0.25) as g:
        e1 = g.slab_energy_from_charge(mv, rho, 5, 8, 4)
        g.mv_initial_state(1, mv, xl, xr)
        g.Update(5)
        e2 = g.slab_energy(8, 4)
    assert abs(e1 - e2) < 1e-12 * abs(e2)


@pytest.mark.parametrize("nsteps,ckpt", [(6, 4), (5, 5), (0, 3)])
def test_charge_gradient_against_finite_differences(nsteps, ckpt):
    from soerenpragya3dglasma_amd import Glasma
    N = (8, 8, 16)
    rho, mv = _left_mover_case(N)
    rng = np.random.default_rng(7)
    with Glasma(N, (1.0, 1.0, 1.0), 0.25) as g:
        e0, grad = g.slab_energy_gradient(mv, rho, nsteps, ckpt, 9, 4)
        assert abs(e0 - g.slab_energy_from_charge(mv, rho, nsteps, 9, 4)) < 1e-13 * abs(e0)

        def cd(v, eps):
            return (g.slab_energy_from_charge(mv, rho + eps * v, nsteps, 9, 4) -
                    g.slab_energy_from_charge(mv, rho - eps * v, nsteps, 9, 4)) / (2 * eps)
        worst_r = worst_6 = 0.0
        for _ in range(32):
            v = rng.normal(size=rho.shape) * np.abs(rho).max()
            ad = float(np.sum(grad * v))
            rich = (4.0 * cd(v, 5e-4) - cd(v, 1e-3)) / 3.0
            worst_r = max(worst_r, abs(rich - ad) / abs(ad))
            worst_6 = max(worst_6, abs(cd(v, 1e-6) - ad) / abs(ad))
    print(nsteps, ckpt, e0, worst_r, worst_6)
    assert worst_r < 1e-6
    assert worst_6 < 1e-4


@pytest.mark.parametrize("nsteps", [6, 0])
def test_charge_gradient_against_oracle_differences(nsteps):
    """The reference has no gradient code, so the reference gradient is the derivative of the oracle's own map
    charge -> left-moving sheet -> nsteps leapfrog steps -> slab energy (the oracle's SetWilsonLines /
    SingleSheet / Update and a numpy slab energy, none of them engine code), taken by Richardson-extrapolated
    central differences along 8 seeded directions; the device's reverse-mode gradient agrees to 1e-6
    relative, and the two energies to 1e-10."""
    from soerenpragya3dglasma_amd import Glasma
    from tests.helpers import slab_energy_np
    N, a, dt, zc, half = (8, 8, 16), (1.0, 1.0, 1.0), 0.25, 9, 4
    rho, mv = _left_mover_case(N)

    def e_oracle(r):
        f = O.left_sheet_from_charge(N, a, dt, r, lam=mv.lambda_, meff=mv.m_eff)
        O.update(f, a, dt, nsteps)
        return slab_energy_np(f, a, zc, half)

    with Glasma(N, a, dt) as g:
        e_dev, grad = g.slab_energy_gradient(mv, rho, nsteps, 3, zc, half)
    e_ref = e_oracle(rho)
    assert abs(e_dev - e_ref) < 1e-10 * abs(e_ref), (e_dev, e_ref)
    rng = np.random.default_rng(11)
    worst = 0.0
    for _ in range(8):
        v = rng.normal(size=rho.shape) * np.abs(rho).max()

        def cd(eps):
            return (e_oracle(rho + eps * v) - e_oracle(rho - eps * v)) / (2 * eps)
        rich = (4.0 * cd(5e-4) - cd(1e-3)) / 3.0
        ad = float(np.sum(grad * v))
        worst = max(worst, abs(rich - ad) / abs(ad))
    print(nsteps, e_dev, e_ref, worst)
    assert worst < 1e-6


def test_torch_function_backward():
    torch = pytest.importorskip("torch")
    from soerenpragya3dglasma_amd import Glasma
    from soerenpragya3dglasma_amd.autograd import slab_energy
    N = (8, 8, 16)
    rho, mv = _left_mover_case(N)
    with Glasma(N, (1.0, 1.0, 1.0), 0.25) as g:
        r = torch.tensor(rho, dtype=torch.float64, requires_grad=True)
        e = slab_energy(r, g, mv, 4, 2, 9, 4)
        (3.0 * e).backward()
        _, grad = g.slab_energy_gradient(mv, rho, 4, 2, 9, 4)
    assert np.abs(r.grad.numpy() - 3.0 * grad).max() <= 1e-14 * np.abs(grad).max()


def test_gradient_work_arrays_reused_across_calls():
    """The slab-energy and gradient calls keep their work arrays in the context: a repeated gradient call of the
    same sizes allocates nothing and returns the same energy and gradient; a longer call after it (more segments,
    so more and larger arrays) agrees with the same call on a fresh context, and so does a shorter one after that.
    The energies agree bit for bit.  The gradients agree to 1e-12 of their largest entry: the final state's link
    cotangent is scattered with atomic adds, whose order, and so the last bit, varies from call to call; an array
    that kept a previous call's contents would show up far above that."""
    import ctypes
    from soerenpragya3dglasma_amd import Glasma, _lib
    N = (8, 8, 16)
    rho, mv = _left_mover_case(N)
    lib = _lib.lib()

    def held(g):
        b = ctypes.c_size_t()
        assert lib.glasma_memory_bytes(g._h, ctypes.byref(b)) == 0
        return b.value

    def fresh(nsteps, ckpt):
        with Glasma(N, (1.0, 1.0, 1.0), 0.25) as f:
            return f.slab_energy_gradient(mv, rho, nsteps, ckpt, 9, 4)

    with Glasma(N, (1.0, 1.0, 1.0), 0.25) as g:
        e1, g1 = g.slab_energy_gradient(mv, rho, 6, 3, 9, 4)
        m1 = held(g)
        e2, g2 = g.slab_energy_gradient(mv, rho, 6, 3, 9, 4)
        assert held(g) == m1
        assert e1 == e2 and np.abs(g1 - g2).max() <= 1e-12 * np.abs(g1).max()
        e3, g3 = g.slab_energy_gradient(mv, rho, 12, 3, 9, 4)
        e4, g4 = g.slab_energy_gradient(mv, rho, 4, 4, 9, 4)
    for (e, gr), ref in [((e3, g3), fresh(12, 3)), ((e4, g4), fresh(4, 4))]:
        print(e, ref[0], np.abs(gr - ref[1]).max() / np.abs(ref[1]).max())
        assert e == ref[0] and np.abs(gr - ref[1]).max() <= 1e-12 * np.abs(ref[1]).max()


@pytest.mark.parametrize("nsteps,ckpt", [(12, 3), (7, 3), (2, 1)])
def test_gradient_one_or_two_segment_sets(monkeypatch, nsteps, ckpt):
    """The gradient with the segment slots in two sets (a segment's recompute on a second stream, overlapping the
    sweep of the segment after it) equals the gradient with one set (each segment recomputed just before its sweep):
    energies bit for bit, gradients to 1e-12 of their largest entry (the final state's atomic scatter)."""
    from soerenpragya3dglasma_amd import Glasma
    N = (8, 8, 16)
    rho, mv = _left_mover_case(N)
    out = {}
    for sets in ("1", "2"):
        monkeypatch.setenv("GLASMA_GRAD_SETS", sets)
        with Glasma(N, (1.0, 1.0, 1.0), 0.25) as g:
            out[sets] = g.slab_energy_gradient(mv, rho, nsteps, ckpt, 9, 4)
    (e1, g1), (e2, g2) = out["1"], out["2"]
    assert e1 == e2 and np.abs(g1 - g2).max() <= 1e-12 * np.abs(g1).max()
