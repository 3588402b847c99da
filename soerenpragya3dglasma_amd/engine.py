"""Python host mirror of the reference's field and observable API, over the libglasma C ABI.

The reference keeps its state in global objects (SC/LATTICE/SUNcVariables.cpp:4-62) and acts on them with
namespace functions.  ``Glasma`` owns one device context and exposes the same operations under the
reference's names:

===============================================  =================================================
reference                                        here
===============================================  =================================================
Lattice::Init (LATTICE/SUNcVariables.cpp:71)     ``Glasma(N, a, dt)``
Dynamics::SUNc::Update (SUNcDynamics.cpp:353)    ``Glasma.Update(nsteps)``
Observables::Bulk::Update (BulkObservables:172)  ``Glasma.bulk()``
GaussLaw::UpdateViolation (GaussLawViolation:16) ``Glasma.gauss()``
Unitarity::UpdateViolation (UnitarityViol.:13)   ``Glasma.unitarity()``
EnergyMomentumTensor::CreateMap (EMT.cpp:92)     ``Glasma.emt_profile()``
CollidingObjects::SetWilsonLines + SingleSheet   ``Glasma.mv_initial_state(...)``
CoulombGaugeFixing::SetCoulombGauge (CGF.cpp:32) ``Glasma.coulomb_gauge(...)``
===============================================  =================================================

Host arrays follow the reference layout with z ghost layers (see ``layout.py``).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib
from ._lib import GlasmaError, MVParams, Params, check

OBS_NAMES = ("T00", "TXX", "TYY", "TZZ", "T0X", "T0Y", "T0Z", "TXY", "TYZ", "TZX")
_TUNE_MAX = 16  # GLASMA_TUNE_MAX of glasma.h


def _p(a: np.ndarray):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    if a.dtype == np.complex128:
        a = a.view(np.float64)
    assert a.dtype == np.float64
    return a.ctypes.data_as(_lib.dp)


def bulk_from_raw(raw, a) -> dict:
    """T00, TXX, TYY, TZZ, ELECTRIC, MAGNETIC, SOURCE from E^2, B^2, E.J sums (BulkObservables.cpp:22-62)."""
    a = np.asarray(a, float)
    acube = a.prod()
    c = (a / acube) ** 2
    cS = a / acube ** 2
    raw = np.asarray(raw)
    ES, BS, EJ = raw[0:3], raw[3:6], raw[6:9]
    t = c * ES + c * BS
    return {
        "T00": 0.5 * t.sum(), "TXX": 0.5 * (t[1] + t[2] - t[0]), "TYY": 0.5 * (t[0] + t[2] - t[1]),
        "TZZ": 0.5 * (t[0] + t[1] - t[2]), "ELECTRIC": 0.5 * (c * ES).sum(), "MAGNETIC": 0.5 * (c * BS).sum(),
        "SOURCE": (cS * EJ).sum(),
    }


class Glasma:
    """One SU(3) lattice on the current GPU."""

    def __init__(self, N, a=(1.0, 1.0, 1.0), dt: float = 0.25, bc: int = 0, _handle=None):
        self.N = tuple(int(n) for n in N)
        self.a = tuple(float(x) for x in a)
        self.dt = float(dt)
        self.bc = int(bc)
        p = Params()
        p.nc = 3
        for i in range(3):
            p.n[i] = self.N[i]
            p.a[i] = self.a[i]
        p.dt = self.dt
        p.bc_z = self.bc
        self.params = p
        self._owned = _handle is None
        if _handle is not None:  # an event of a GlasmaBatch: the batch owns the context
            self._h = _handle
            return
        h = ctypes.c_void_p()
        rc = _lib.lib().glasma_create(ctypes.byref(p), ctypes.byref(h))
        if rc != 0:
            raise GlasmaError(f"glasma_create failed ({rc}): {_lib.lib().glasma_last_error(None)}")
        self._h = h

    # -- lifetime -----------------------------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            if getattr(self, "_owned", True):
                _lib.lib().glasma_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _chk(self, rc, what):
        check(rc, self._h, what)

    # -- fields -------------------------------------------------------------------------------------
    def shapes(self):
        Nx, Ny, Nz = self.N
        return {
            "U": ((Nz + 2, Ny, Nx, 3, 3, 3), np.complex128),
            "E": ((Nz + 2, Ny, Nx, 3, 8), np.float64),
            "J": ((Nz + 2, Ny, Nx, 8), np.float64),
        }

    def set_fields(self, U=None, E=None, Jp=None, Jm=None, JpS=None, JmS=None):
        arrs = [U, E, Jp, Jm, JpS, JmS]
        arrs = [None if x is None else np.ascontiguousarray(x) for x in arrs]
        self._chk(_lib.lib().glasma_set_fields(self._h, *[_p(x) for x in arrs]), "glasma_set_fields")

    def get_fields(self) -> dict:
        sh = self.shapes()
        out = {"U": np.empty(*sh["U"]), "E": np.empty(*sh["E"])}
        for k in ("Jp", "Jm", "JpS", "JmS"):
            out[k] = np.empty(*sh["J"])
        self._chk(_lib.lib().glasma_get_fields(self._h, _p(out["U"]), _p(out["E"]), _p(out["Jp"]), _p(out["Jm"]),
                                               _p(out["JpS"]), _p(out["JmS"])), "glasma_get_fields")
        return out

    def get_uold(self) -> np.ndarray:
        U = np.empty(*self.shapes()["U"])
        self._chk(_lib.lib().glasma_get_uold(self._h, _p(U)), "glasma_get_uold")
        return U

    def set_uold(self, U):
        U = np.ascontiguousarray(U, np.complex128)
        self._chk(_lib.lib().glasma_set_uold(self._h, _p(U)), "glasma_set_uold")

    def time(self):
        t = ctypes.c_double()
        n = ctypes.c_long()
        self._chk(_lib.lib().glasma_get_time(self._h, ctypes.byref(t), ctypes.byref(n)), "glasma_get_time")
        return t.value, n.value

    def set_time(self, t: float, tsteps: int):
        self._chk(_lib.lib().glasma_set_time(self._h, float(t), int(tsteps)), "glasma_set_time")

    def fill_random(self, seed: int = 1, link_scale: float = 0.5, e_scale: float = 0.1, j_scale: float = 0.05):
        """Synthetic fields generated on the device (no counterpart in the reference)."""
        self._chk(_lib.lib().glasma_fill_random(self._h, int(seed), link_scale, e_scale, j_scale), "glasma_fill_random")

    # -- dynamics -----------------------------------------------------------------------------------
    def Update(self, nsteps: int = 1):
        """Dynamics::SUNc::Update, nsteps times (DYNAMICS/SUNcDynamics.cpp:353-378)."""
        self._chk(_lib.lib().glasma_update(self._h, int(nsteps)), "glasma_update")

    update = Update

    def update_async(self, nsteps: int):
        self._chk(_lib.lib().glasma_update_async(self._h, int(nsteps)), "glasma_update_async")

    def synchronize(self):
        self._chk(_lib.lib().glasma_synchronize(self._h), "glasma_synchronize")

    def last_step_kernel_ms(self) -> float:
        ms = ctypes.c_double()
        self._chk(_lib.lib().glasma_last_step_kernel_ms(self._h, ctypes.byref(ms)), "glasma_last_step_kernel_ms")
        return ms.value

    # -- configuration files (IO/SaveConfiguration.cpp, IO/LoadConfiguration.cpp) ---------------------
    @staticmethod
    def _path(p):
        return None if p is None else os.fsencode(p)

    def save_configuration(self, ufile=None, efile=None, jfile=None):
        """The reference's text configuration files (U, E and J; any may be None)."""
        self._chk(_lib.lib().glasma_save_configuration(self._h, self._path(ufile), self._path(efile),
                                                       self._path(jfile)), "glasma_save_configuration")

    def load_configuration(self, ufile=None, efile=None, jfile=None, t: float = 0.0):
        """Fields from the reference's text configuration files; the clock is set to t."""
        self._chk(_lib.lib().glasma_load_configuration(self._h, self._path(ufile), self._path(efile),
                                                       self._path(jfile), float(t)), "glasma_load_configuration")

    def save_state(self, path):
        """Exact binary state (fields, UOld, clock) for a bit-identical restart."""
        self._chk(_lib.lib().glasma_save_state(self._h, self._path(path)), "glasma_save_state")

    def load_state(self, path):
        self._chk(_lib.lib().glasma_load_state(self._h, self._path(path)), "glasma_load_state")

    # -- tile-order autotuner (no counterpart in the reference) --------------------------------------
    @staticmethod
    def _order(code: int) -> str:
        mode, bw, bh = code // 10000, (code // 100) % 100, code % 100
        if mode == 2:
            return f"bands {bw}x{bh}"
        return {1: "plane bands", 3: "z slabs"}.get(mode, "linear")

    def tune_step(self, reps: int = 3) -> dict:
        """Time the step under every block -> tile order that fits the lattice (on scratch outputs: the
        fields are unchanged) and keep the fastest for later steps.  Returns the candidates' ms per step
        and the chosen order."""
        n = ctypes.c_int()
        codes = np.zeros(_TUNE_MAX, dtype=np.int32)
        ms = np.zeros(_TUNE_MAX)
        best = ctypes.c_int()
        self._chk(_lib.lib().glasma_tune_step(self._h, int(reps), ctypes.byref(n),
                                              codes.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _p(ms),
                                              ctypes.byref(best)), "glasma_tune_step")
        cand = {self._order(int(c)): float(t) for c, t in zip(codes[:n.value], ms[:n.value])}
        return {"candidates_ms": cand, "chosen": self._order(int(codes[best.value]))}

    def copy_ceiling(self, reps: int = 5) -> float:
        """ms per launch of a copy of the step's compulsory traffic (880 + 880 B per site, the step's tile layout
        and grid): the bandwidth ceiling the step kernel is measured against (glasma_copy_ceiling)."""
        ms = ctypes.c_double()
        self._chk(_lib.lib().glasma_copy_ceiling(self._h, int(reps), ctypes.byref(ms)), "glasma_copy_ceiling")
        return ms.value

    def tile_order(self) -> tuple:
        """(order in use, whether it was tuned)."""
        code, tuned = ctypes.c_int(), ctypes.c_int()
        self._chk(_lib.lib().glasma_tile_order(self._h, ctypes.byref(code), ctypes.byref(tuned)), "glasma_tile_order")
        return self._order(code.value), bool(tuned.value)

    # -- observables --------------------------------------------------------------------------------
    def bulk_raw(self) -> np.ndarray:
        out = np.zeros(9)
        self._chk(_lib.lib().glasma_bulk(self._h, _p(out)), "glasma_bulk")
        return out

    def bulk(self) -> dict:
        return bulk_from_raw(self.bulk_raw(), self.a)

    def gauss(self):
        pz = np.zeros(self.N[2])
        mx = ctypes.c_double()
        rms = ctypes.c_double()
        self._chk(_lib.lib().glasma_gauss(self._h, _p(pz), ctypes.byref(mx), ctypes.byref(rms)), "glasma_gauss")
        return pz, mx.value, rms.value

    def unitarity(self):
        mx = ctypes.c_double()
        rms = ctypes.c_double()
        self._chk(_lib.lib().glasma_unitarity(self._h, ctypes.byref(mx), ctypes.byref(rms)), "glasma_unitarity")
        return mx.value, rms.value

    def checksum(self) -> tuple:
        """Bitwise checksums (U, UOld, E, currents) of the device state, see glasma_checksum."""
        out = (ctypes.c_uint64 * 4)()
        self._chk(_lib.lib().glasma_checksum(self._h, out), "glasma_checksum")
        return tuple(int(v) for v in out)

    def emt_profile(self) -> np.ndarray:
        out = np.zeros((self.N[2], 10))
        self._chk(_lib.lib().glasma_emt(self._h, _p(out)), "glasma_emt")
        return out

    def field_profile(self) -> np.ndarray:
        """IO::SaveFields columns per z (Nz, 4): 0.5(Bx^2+By^2), 0.5 Bz^2, 0.5(Ex^2+Ey^2), 0.5 Ez^2."""
        out = np.zeros((self.N[2], 4))
        self._chk(_lib.lib().glasma_field_profile(self._h, _p(out)), "glasma_field_profile")
        return out

    def current_profile(self) -> np.ndarray:
        """IO::SaveTransverselyIntegratedCurrents columns for z = -1..Nz (Nz + 2, 4): 0.5<Jm^2>,
        0.5<JmStat^2>, 0.5<Jp^2>, 0.5<JpStat^2>."""
        out = np.zeros((self.N[2] + 2, 4))
        self._chk(_lib.lib().glasma_current_profile(self._h, _p(out)), "glasma_current_profile")
        return out

    def transverse_energy_map(self) -> np.ndarray:
        """EnergyMomentumTensor::CreateMapBfCollision values (Ny, Nx, 2): T00Left, T00Right."""
        out = np.zeros((self.N[1], self.N[0], 2))
        self._chk(_lib.lib().glasma_transverse_energy_map(self._h, _p(out)), "glasma_transverse_energy_map")
        return out

    # -- initial conditions -------------------------------------------------------------------------
    @staticmethod
    def draw_charges(seed: int, Nx: int, Ny: int, qs_l: float, qs_r: float):
        xl = np.zeros((Ny, Nx, 8))
        xr = np.zeros((Ny, Nx, 8))
        check(_lib.lib().glasma_draw_charges(int(seed), Nx, Ny, qs_l, qs_r, _p(xl), _p(xr)), None, "glasma_draw_charges")
        return xl, xr

    @staticmethod
    def mv_params(lam, meff, sigma_l, sigma_r, center_l, center_r, mult_l=1.0, mult_r=1.0) -> MVParams:
        m = MVParams()
        m.lambda_, m.m_eff, m.sigma_l, m.sigma_r = lam, meff, sigma_l, sigma_r
        m.nucleus_l, m.nucleus_r = mult_l, mult_r
        m.center_l, m.center_r = int(center_l), int(center_r)
        return m

    def mv_initial_state(self, driver: int, mv: MVParams, xi_l, xi_r):
        """Simulation::Run's initial state (SC/Simulation.cpp:327-378); driver 0 collision, 1 left, 2 right."""
        xl = np.ascontiguousarray(xi_l, np.float64)
        xr = np.ascontiguousarray(xi_r, np.float64)
        self._chk(_lib.lib().glasma_mv_initial_state(self._h, int(driver), ctypes.byref(mv), _p(xl), _p(xr)),
                  "glasma_mv_initial_state")

    @staticmethod
    def tmd_params(qs0a, m_eff, com_energy, a_fm, atomic_number, center_l, center_r) -> "_lib.TmdParams":
        """Physical-charge parameters (SC/LATTICE/Parameters.cpp): Qs0A and MEff in lattice units, CoMEnergy in
        fm^-1, the spacings in fm, the atomic number and the nuclei's z centres."""
        p = _lib.TmdParams()
        p.qs0a, p.m_eff, p.com_energy = qs0a, m_eff, com_energy
        for i in range(3):
            p.a_fm[i] = a_fm[i]
        p.atomic_number, p.center_l, p.center_r = int(atomic_number), int(center_l), int(center_r)
        return p

    @staticmethod
    def tmd_draw(seed: int, N, atomic_number: int):
        """SimulationmitPC::Run's draws: (noise_l, noise_r) shaped (Nz, Ny, Nx, 8), nucleons (A, 3) in fm."""
        Nx, Ny, Nz = N
        nl = np.zeros((Nz, Ny, Nx, 8))
        nr = np.zeros_like(nl)
        pl = np.zeros((atomic_number, 3))
        pr = np.zeros_like(pl)
        check(_lib.lib().glasma_tmd_draw(int(seed), Nx, Ny, Nz, int(atomic_number), _p(nl), _p(nr), _p(pl), _p(pr)),
              None, "glasma_tmd_draw")
        return nl, nr, pl, pr

    def tmd_initial_state(self, driver: int, p, noise_l, noise_r, nucleons_l, nucleons_r):
        """SimulationmitPC::Run's initial state (physical-charge nuclei); driver 0 collision, 1 left, 2 right."""
        arrs = [np.ascontiguousarray(v, np.float64) for v in (noise_l, noise_r, nucleons_l, nucleons_r)]
        self._chk(_lib.lib().glasma_tmd_initial_state(self._h, int(driver), ctypes.byref(p), *[_p(v) for v in arrs]),
                  "glasma_tmd_initial_state")

    def set_wilson_lines(self, direction: str, mv: MVParams, xi):
        Nx, Ny, Nz = self.N
        V = np.empty((Nz + 2, Ny, Nx, 3, 3), np.complex128)
        rho = np.empty((Nz + 2, Ny, Nx, 8))
        x = np.ascontiguousarray(xi, np.float64)
        self._chk(_lib.lib().glasma_set_wilson_lines(self._h, direction.encode()[0:1], ctypes.byref(mv), _p(x),
                                                     _p(V), _p(rho)), "glasma_set_wilson_lines")
        return V, rho

    # -- reverse-mode derivatives ---------------------------------------------------------------------
    def slab_energy(self, zc: int, half: int) -> float:
        """Energy of the layers |z - zc| < half (the local form of the bulk T00)."""
        e = ctypes.c_double()
        self._chk(_lib.lib().glasma_slab_energy(self._h, int(zc), int(half), ctypes.byref(e)), "glasma_slab_energy")
        return e.value

    def step_vjp(self, cot: dict) -> dict:
        """Cotangents of the fields before one leapfrog step from those after it (keys U, E, Jp, Jm, JpS, JmS)."""
        d = {k: np.ascontiguousarray(cot[k]).copy() for k in ("U", "E", "Jp", "Jm", "JpS", "JmS")}
        self._chk(_lib.lib().glasma_step_vjp(self._h, *[_p(d[k]) for k in ("U", "E", "Jp", "Jm", "JpS", "JmS")]),
                  "glasma_step_vjp")
        return d

    def _rho(self, rho):
        Nx, Ny, Nz = self.N
        r = np.zeros((Nz + 2, Ny, Nx, 8))
        rho = np.asarray(rho, np.float64)
        if rho.shape == (Nz, Ny, Nx, 8):
            r[1:-1] = rho
        else:
            r[...] = rho
        return r

    def slab_energy_from_charge(self, mv: MVParams, rho, nsteps: int, zc: int, half: int) -> float:
        """Left-moving MV sheet from the charge density rho (Nz, Ny, Nx, 8), nsteps steps, slab energy."""
        e = ctypes.c_double()
        r = self._rho(rho)
        self._chk(_lib.lib().glasma_slab_energy_from_charge(self._h, ctypes.byref(mv), _p(r), int(nsteps), int(zc),
                                                            int(half), ctypes.byref(e)),
                  "glasma_slab_energy_from_charge")
        return e.value

    def slab_energy_gradient(self, mv: MVParams, rho, nsteps: int, ckpt: int, zc: int, half: int):
        """(energy, d energy / d rho) with rho and the gradient shaped (Nz, Ny, Nx, 8)."""
        e = ctypes.c_double()
        r = self._rho(rho)
        gr = np.zeros_like(r)
        self._chk(_lib.lib().glasma_slab_energy_gradient(self._h, ctypes.byref(mv), _p(r), int(nsteps), int(ckpt),
                                                         int(zc), int(half), ctypes.byref(e), _p(gr)),
                  "glasma_slab_energy_gradient")
        return e.value, gr[1:-1].copy()

    # -- gauge fixing -------------------------------------------------------------------------------
    def coulomb_gauge(self, tol: float = 1e-15, maxit: int = 1000, alpha: float = 2.0) -> np.ndarray:
        stats = np.zeros(7)
        self._chk(_lib.lib().glasma_coulomb_gauge(self._h, tol, int(maxit), alpha, _p(stats)), "glasma_coulomb_gauge")
        self._warn_cap(stats, tol, maxit)
        return stats

    GAUGE_ALGORITHMS = {"fourier": 0, "jacobi": 1, "losalamos": 2, "consistent": 3}

    def coulomb_gauge_algorithm(self, algorithm: str, tol: float = 1e-15, maxit: int = 1000,
                                param: float = None) -> np.ndarray:
        """Coulomb gauge fixing with the reference's "fourier" (param = alpha, 2), "jacobi" (param = step size,
        0.3) or "losalamos" (param = over-relaxation probability, 0) update, or the boundary-consistent Fourier
        update "consistent" (param = alpha, 2; no counterpart in the reference), which converges for Nz > 1.
        A RuntimeWarning is issued when the iteration cap is reached above a tolerance tol > 0."""
        algo = self.GAUGE_ALGORITHMS[algorithm]
        if param is None:
            param = (2.0, 0.3, 0.0, 2.0)[algo]
        stats = np.zeros(7)
        self._chk(_lib.lib().glasma_coulomb_gauge_algorithm(self._h, algo, tol, int(maxit), float(param), _p(stats)),
                  "glasma_coulomb_gauge_algorithm")
        self._warn_cap(stats, tol, maxit)
        return stats

    def _warn_cap(self, stats, tol, maxit):
        if tol > 0 and stats[0] >= maxit and stats[1] > tol:
            import warnings
            warnings.warn(_lib.lib().glasma_last_error(self._h).decode(), RuntimeWarning, stacklevel=3)

    def gauss_restore(self, transverse_only: bool = False, tol: float = 1e-14, maxit: int = 30000) -> np.ndarray:
        """GaussLawRestoration::Restore (or the 2D variant): relax E toward the Gauss law in place; returns
        (iterations, max |G|, RMS G)."""
        stats = np.zeros(3)
        self._chk(_lib.lib().glasma_gauss_restore(self._h, int(bool(transverse_only)), tol, int(maxit), _p(stats)),
                  "glasma_gauss_restore")
        return stats

    def gauge_reset(self):
        self._chk(_lib.lib().glasma_gauge_reset(self._h), "glasma_gauge_reset")

    def get_gauge_fixed(self):
        sh = self.shapes()
        Nx, Ny, Nz = self.N
        UG = np.empty(*sh["U"])
        EG = np.empty(*sh["E"])
        G = np.empty((Nz + 2, Ny, Nx, 3, 3), np.complex128)
        self._chk(_lib.lib().glasma_get_gauge_fixed(self._h, _p(UG), _p(EG), _p(G)), "glasma_get_gauge_fixed")
        return UG, EG, G

    def set_gauge_transformation(self, G):
        G = np.ascontiguousarray(G, np.complex128)
        self._chk(_lib.lib().glasma_set_gauge_transformation(self._h, _p(G)), "glasma_set_gauge_transformation")


class GlasmaBatch:
    """An ensemble of independent events advanced together (one launch per step over all events, the
    steps replayed from a HIP graph).  ``events[e]`` is a ``Glasma`` view of event e for fields, initial
    conditions and observables."""

    def __init__(self, specs):
        """specs: list of (N, a, dt) or (N, a, dt, bc)."""
        n = len(specs)
        arr = (Params * n)()
        for i, sp in enumerate(specs):
            N, a, dt = sp[:3]
            bc = sp[3] if len(sp) > 3 else 0
            arr[i].nc = 3
            for k in range(3):
                arr[i].n[k] = int(N[k])
                arr[i].a[k] = float(a[k])
            arr[i].dt = float(dt)
            arr[i].bc_z = int(bc)
        h = ctypes.c_void_p()
        rc = _lib.lib().glasma_batch_create(arr, n, ctypes.byref(h))
        if rc != 0:
            raise GlasmaError(f"glasma_batch_create failed ({rc})")
        self._h = h
        self.events = []
        for i, sp in enumerate(specs):
            c = ctypes.c_void_p()
            self._chk(_lib.lib().glasma_batch_context(self._h, i, ctypes.byref(c)), "glasma_batch_context")
            self.events.append(Glasma(sp[0], sp[1], sp[2], sp[3] if len(sp) > 3 else 0, _handle=c))

    def _chk(self, rc, what):
        if rc != 0:
            msg = _lib.lib().glasma_batch_last_error(self._h)
            raise GlasmaError(f"{what} failed with status {rc}: {msg.decode() if msg else ''}")

    def update(self, nsteps: int):
        self._chk(_lib.lib().glasma_batch_update(self._h, int(nsteps)), "glasma_batch_update")

    def update_serial(self, nsteps: int):
        self._chk(_lib.lib().glasma_batch_update_serial(self._h, int(nsteps)), "glasma_batch_update_serial")

    def last_ms(self) -> float:
        ms = ctypes.c_double()
        self._chk(_lib.lib().glasma_batch_last_ms(self._h, ctypes.byref(ms)), "glasma_batch_last_ms")
        return ms.value

    def tune(self, reps: int = 3):
        """Tune the step's tile order once per distinct lattice of the ensemble (Glasma.tune_step) and
        rebuild the work list; the events' fields are unchanged."""
        self._chk(_lib.lib().glasma_batch_tune(self._h, int(reps)), "glasma_batch_tune")

    def close(self):
        if getattr(self, "_h", None):
            for ev in self.events:
                ev._h = None
            _lib.lib().glasma_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def fft_host(data: np.ndarray, dims, batch: int, sign: int) -> np.ndarray:
    """Device FFT of a host complex array (FFTW conventions, unnormalised)."""
    d = np.ascontiguousarray(data, np.complex128).copy()
    dv = (ctypes.c_int * len(dims))(*[int(x) for x in dims])
    check(_lib.lib().glasma_fft_host(_p(d), len(dims), dv, int(batch), int(sign)), None, "glasma_fft_host")
    return d


def su3_batch(op: int, c: float, inp: np.ndarray, in2: np.ndarray | None = None) -> np.ndarray:
    """Batched per-site SU(3) algebra on the device (see glasma_su3_batch in include/glasma.h)."""
    inp = np.ascontiguousarray(inp)
    n = inp.shape[0]
    out_w = {0: 18, 1: 8, 2: 8, 3: 8, 4: 8}[op]
    out = np.zeros((n, out_w))
    i2 = None if in2 is None else np.ascontiguousarray(in2, np.float64)
    a = inp.view(np.float64) if inp.dtype == np.complex128 else inp
    check(_lib.lib().glasma_su3_batch(int(op), n, float(c), _p(np.ascontiguousarray(a)), _p(i2), _p(out)), None,
          "glasma_su3_batch")
    return out
