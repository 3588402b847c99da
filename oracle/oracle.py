"""ctypes front end of the oracle (``oracle/cpu_ref.cpp``) — TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import this module,
and only as the checker (or the timed CPU baseline).  The product package never imports it.

Every array here is a numpy array in the reference's array-of-structs layout *including* the two z
ghost layers (SC/LATTICE/GaugeLinks.cpp:27-36, ElectricFields.cpp:25-35, ChargeDensity.cpp:27-37):

* U, UOld: complex128 of shape (Nz+2, Ny, Nx, 3, 3, 3) -> [z+1, y, x, mu, j, i] so that the flat order
  equals ``9*(mu + 3*Index3D) + i + 3*j`` (column-major matrices, SC/SUNc/SU3/GroupOperations.cpp:13-15).
* E: float64 (Nz+2, Ny, Nx, 3, Ng); J*: float64 (Nz+2, Ny, Nx, Ng).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libcpu_ref.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++ only)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-B" if force else "all"], cwd=_HERE)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        L = _lib
        L.ref_su3_retridminusu.restype = ctypes.c_double
        L.ref_su3_unitarity_norm.restype = ctypes.c_double
        L.ref_su3_expi.argtypes = [ctypes.c_double, _dp, _dp]
        L.ref_su3_ilog.argtypes = [ctypes.c_double, _dp, _dp]
        L.ref_su2_expi.argtypes = [ctypes.c_double, _dp, _dp]
        L.ref_su2_ilog.argtypes = [ctypes.c_double, _dp, _dp]
        L.ref_update.argtypes = [ctypes.c_int, _ip, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_int] + [_dp] * 7
        L.ref_bulk.argtypes = [ctypes.c_int, _ip, _dp, ctypes.c_int] + [_dp] * 5
        L.ref_gauss.argtypes = [ctypes.c_int, _ip, _dp, ctypes.c_int] + [_dp] * 7
        L.ref_unitarity.argtypes = [ctypes.c_int, _ip, _dp, ctypes.c_int] + [_dp] * 3
        L.ref_emt.argtypes = [ctypes.c_int, _ip, _dp, ctypes.c_int] + [_dp] * 4
        L.ref_field_profile.argtypes = [ctypes.c_int, _ip, _dp, ctypes.c_int] + [_dp] * 3
        L.ref_current_profile.argtypes = [ctypes.c_int, _ip] + [_dp] * 5
        L.ref_transverse_map.argtypes = [ctypes.c_int, _ip, _dp, ctypes.c_int] + [_dp] * 3
        L.ref_fft_nd.argtypes = [_dp, ctypes.c_int, _ip, ctypes.c_long, ctypes.c_int]
        L.ref_mt_raw.argtypes = [ctypes.c_ulong, ctypes.c_int, ctypes.c_long, ctypes.POINTER(ctypes.c_uint32)]
        L.ref_gauss_draws.argtypes = [ctypes.c_ulong, ctypes.c_long, _dp]
        L.ref_draw_charges.argtypes = [ctypes.c_ulong, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.ref_set_wilson_lines.argtypes = [_ip, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_double, _dp, ctypes.c_char, ctypes.c_int, _dp, _dp, _dp]
        L.ref_single_sheet.argtypes = [_ip, _dp, ctypes.c_int, ctypes.c_double, _dp, ctypes.c_int, ctypes.c_int, _dp,
                                       _dp, ctypes.c_char, _dp, _dp, _dp, _dp, _dp]
        L.ref_evolve_rho.argtypes = [_ip, _dp, ctypes.c_double, ctypes.c_char, _dp, _dp]
        L.ref_solve_laplace.argtypes = [_ip, _dp, _dp]
        L.ref_mv_initial_state.argtypes = ([ctypes.c_int, _ip, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp, _dp] + [_dp] * 7)
        L.ref_left_sheet_from_charge.argtypes = [_ip, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                 ctypes.c_double, _dp] + [_dp] * 5
        L.ref_coulomb_gauge.argtypes = [_ip, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int,
                                        ctypes.c_double, _dp]
        L.ref_gauge_transform.argtypes = [_ip, _dp, ctypes.c_int, _dp] + [_dp] * 6
        L.ref_coulomb_gauge_algo.argtypes = [_ip, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_double,
                                             ctypes.c_int, ctypes.c_double, _dp]
        L.ref_coulomb_gauge_algo_bc.argtypes = [_ip, _dp, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, ctypes.c_int,
                                                ctypes.c_double, ctypes.c_int, ctypes.c_double, _dp]
        L.ref_max_trace_projection.argtypes = [_dp, _dp]
        L.ref_tmd_draw.argtypes = [ctypes.c_ulong, _ip, ctypes.c_int] + [_dp] * 4
        L.ref_tmd_draw.restype = ctypes.c_int
        L.ref_charge_fluctuation.argtypes = [_ip, _dp, _dp, _ip, _dp, _dp]
        L.ref_nucleus_profile.argtypes = [_dp, _ip, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.ref_nucleus_profile.restype = ctypes.c_double
        L.ref_tmd_initial_state.argtypes = [ctypes.c_int, _ip, _dp, ctypes.c_int, ctypes.c_double, _dp, _ip] + [_dp] * 13
        L.ref_gauss_restore.argtypes = [_ip, _dp, ctypes.c_int] + [_dp] * 4 + [ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                                                            _dp]
        L.ref_sync_ghosts.argtypes = [ctypes.c_int, _ip, ctypes.c_int, _dp, ctypes.c_long]
        L.ref_num_threads.restype = ctypes.c_int
        L.ref_set_threads.argtypes = [ctypes.c_int]
    return _lib


def P(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    if a.dtype == np.complex128:
        return a.view(np.float64).ctypes.data_as(_dp)
    assert a.dtype == np.float64
    return a.ctypes.data_as(_dp)


def ivec(v):
    return (ctypes.c_int * len(v))(*[int(x) for x in v])


def dvec(v):
    return (ctypes.c_double * len(v))(*[float(x) for x in v])


# ------------------------------------------------------------------------------------------------
# field containers
# ------------------------------------------------------------------------------------------------
class Fields:
    """Reference-layout SU(Nc) fields with ghost layers: U, UOld, E, Jp, Jm, JpS, JmS."""

    def __init__(self, N, nc: int = 3):
        self.N = tuple(int(n) for n in N)
        self.nc = nc
        Nx, Ny, Nz = self.N
        ng = 8 if nc == 3 else 3
        self.ng = ng
        if nc == 3:
            self.U = np.zeros((Nz + 2, Ny, Nx, 3, 3, 3), np.complex128)
            for i in range(3):
                self.U[..., i, i] = 1.0
        else:
            self.U = np.zeros((Nz + 2, Ny, Nx, 3, 4), np.float64)
            self.U[..., 3] = 1.0
        self.UOld = self.U.copy()
        self.E = np.zeros((Nz + 2, Ny, Nx, 3, ng))
        self.Jp = np.zeros((Nz + 2, Ny, Nx, ng))
        self.Jm = np.zeros_like(self.Jp)
        self.JpS = np.zeros_like(self.Jp)
        self.JmS = np.zeros_like(self.Jp)

    def copy(self) -> "Fields":
        f = Fields.__new__(Fields)
        f.N, f.nc, f.ng = self.N, self.nc, self.ng
        for k in ("U", "UOld", "E", "Jp", "Jm", "JpS", "JmS"):
            setattr(f, k, getattr(self, k).copy())
        return f

    def arrays(self):
        return [self.U, self.UOld, self.E, self.Jp, self.Jm, self.JpS, self.JmS]

    def sync(self, bc: int = 0):
        """Open (bc=0) or periodic (bc=1) ghost synchronisation of every array."""
        for A in self.arrays():
            if bc == 0:
                A[0] = A[1]
                A[-1] = A[-2]
            else:
                A[0] = A[-2]
                A[-1] = A[1]


def update(f: Fields, a, dt: float, nsteps: int = 1, bc: int = 0):
    """Dynamics::SUNc::Update repeated nsteps times (SC/DYNAMICS/SUNcDynamics.cpp:353-378)."""
    lib().ref_update(f.nc, ivec(f.N), dvec(a), bc, dt, nsteps, P(f.U), P(f.UOld), P(f.E), P(f.Jp), P(f.Jm),
                     P(f.JpS), P(f.JmS))


def bulk(f: Fields, a, bc: int = 0) -> np.ndarray:
    out = np.zeros(9)
    lib().ref_bulk(f.nc, ivec(f.N), dvec(a), bc, P(f.U), P(f.E), P(f.Jp), P(f.Jm), P(out))
    return out


def gauss(f: Fields, a, bc: int = 0):
    pz = np.zeros(f.N[2])
    mx = np.zeros(1)
    rms = np.zeros(1)
    lib().ref_gauss(f.nc, ivec(f.N), dvec(a), bc, P(f.U), P(f.E), P(f.JpS), P(f.JmS), P(pz), P(mx), P(rms))
    return pz, float(mx[0]), float(rms[0])


def unitarity(f: Fields, a, bc: int = 0):
    mx = np.zeros(1)
    rms = np.zeros(1)
    lib().ref_unitarity(f.nc, ivec(f.N), dvec(a), bc, P(f.U), P(mx), P(rms))
    return float(mx[0]), float(rms[0])


def emt(f: Fields, a, bc: int = 0) -> np.ndarray:
    out = np.zeros((f.N[2], 10))
    lib().ref_emt(f.nc, ivec(f.N), dvec(a), bc, P(f.U), P(f.UOld), P(f.E), P(out))
    return out


def field_profile(f: Fields, a, bc: int = 0) -> np.ndarray:
    """IO::SaveFields columns per z (SC/IO/SaveConfiguration.cpp:112-158), shape (Nz, 4)."""
    out = np.zeros((f.N[2], 4))
    lib().ref_field_profile(f.nc, ivec(f.N), dvec(a), bc, P(f.U), P(f.E), P(out))
    return out


def current_profile(f: Fields) -> np.ndarray:
    """IO::SaveTransverselyIntegratedCurrents rows z = -1..Nz (SC/IO/SaveConfiguration.cpp:160-204), (Nz + 2, 4)."""
    out = np.zeros((f.N[2] + 2, 4))
    lib().ref_current_profile(f.nc, ivec(f.N), P(f.Jp), P(f.Jm), P(f.JpS), P(f.JmS), P(out))
    return out


def transverse_map(f: Fields, a, bc: int = 0) -> np.ndarray:
    """EnergyMomentumTensor::CreateMapBfCollision (SC/OBSERVABLES/EnergyMomentumTensor.cpp:10-77), (Ny, Nx, 2)."""
    out = np.zeros((f.N[1], f.N[0], 2))
    lib().ref_transverse_map(f.nc, ivec(f.N), dvec(a), bc, P(f.U), P(f.E), P(out))
    return out


def bulk_T(raw: np.ndarray, a) -> dict:
    """T00, TXX, TYY, TZZ, ELECTRIC, MAGNETIC, SOURCE from the raw sums (SC/OBSERVABLES/BulkObservables.cpp:22-62)."""
    a = np.asarray(a, float)
    acube = a.prod()
    c = (a / acube) ** 2
    cS = a / acube ** 2
    ES, BS, EJ = raw[0:3], raw[3:6], raw[6:9]
    t = c * ES + c * BS
    return {
        "T00": 0.5 * t.sum(),
        "TXX": 0.5 * (t[1] + t[2] - t[0]),
        "TYY": 0.5 * (t[0] + t[2] - t[1]),
        "TZZ": 0.5 * (t[0] + t[1] - t[2]),
        "ELECTRIC": 0.5 * (c * ES).sum(),
        "MAGNETIC": 0.5 * (c * BS).sum(),
        "SOURCE": (cS * EJ).sum(),
    }


# ------------------------------------------------------------------------------------------------
# SU(3) per-site helpers (flat complex[9] column-major)
# ------------------------------------------------------------------------------------------------
def su3_expi(c: float, alpha) -> np.ndarray:
    al = np.ascontiguousarray(alpha, np.float64)
    X = np.zeros(9, np.complex128)
    lib().ref_su3_expi(c, P(al), P(X))
    return X


def su3_ilog(c: float, U) -> np.ndarray:
    Uf = np.ascontiguousarray(U, np.complex128).reshape(9)
    out = np.zeros(8)
    lib().ref_su3_ilog(c, P(Uf), P(out))
    return out


def su3_call(name: str, *arrays, out_shape=8, out_dtype=np.float64):
    ins = [np.ascontiguousarray(x) for x in arrays]
    out = np.zeros(out_shape, out_dtype)
    getattr(lib(), name)(*[P(x) for x in ins], P(out))
    return out


def fft_nd(data: np.ndarray, dims, batch: int, sign: int) -> np.ndarray:
    d = np.ascontiguousarray(data, np.complex128).copy()
    lib().ref_fft_nd(P(d), len(dims), ivec(dims), batch, sign)
    return d


def mt_raw(seed: int, n: int, std_seeding: bool = False) -> np.ndarray:
    out = np.zeros(n, np.uint32)
    lib().ref_mt_raw(seed, int(std_seeding), n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    return out


def gauss_draws(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n)
    lib().ref_gauss_draws(seed, n, P(out))
    return out


def draw_charges(seed: int, Nx: int, Ny: int, qs_l: float, qs_r: float):
    xl = np.zeros((Ny, Nx, 8))
    xr = np.zeros((Ny, Nx, 8))
    lib().ref_draw_charges(seed, Nx, Ny, qs_l, qs_r, P(xl), P(xr))
    return xl, xr


def mv_initial_state(driver: int, N, a, dt, xi_l, xi_r, *, lam, meff, sigma_l, sigma_r, center_l, center_r,
                     mult_l=1.0, mult_r=1.0, bc=0) -> Fields:
    """Initial fields of Simulation::Run (SC/Simulation.cpp:327-378); driver 0 collision, 1 left, 2 right."""
    f = Fields(N, 3)
    xl = np.ascontiguousarray(xi_l, np.float64)
    xr = np.ascontiguousarray(xi_r, np.float64)
    lib().ref_mv_initial_state(driver, ivec(N), dvec(a), bc, dt, lam, meff, sigma_l, sigma_r, mult_l, mult_r,
                               int(center_l), int(center_r), P(xl), P(xr), P(f.U), P(f.UOld), P(f.E), P(f.Jp),
                               P(f.Jm), P(f.JpS), P(f.JmS))
    return f


def left_sheet_from_charge(N, a, dt, rho, *, lam, meff, bc=0) -> Fields:
    """The left-moving sheet over all z (SingleSheet 'L', the driver of SC/LeftNucleusSimulation.cpp:357)
    built from a given charge density rho shaped (Nz, Ny, Nx, 8): the density SetChargeDensities
    (COLLIDINGOBJECTS/CollidingObjects.cpp:34-102) would produce, before smoothing."""
    f = Fields(N, 3)
    Nx, Ny, Nz = N
    r = np.zeros((Nz + 2, Ny, Nx, 8))
    r[1:-1] = np.asarray(rho, np.float64).reshape(Nz, Ny, Nx, 8)
    lib().ref_left_sheet_from_charge(ivec(N), dvec(a), bc, dt, lam, meff, P(r), P(f.U), P(f.UOld), P(f.E), P(f.Jp),
                                     P(f.JpS))
    return f


def coulomb_gauge(f: Fields, a, G: np.ndarray, tol: float, maxit: int, alpha: float = 2.0):
    """CoulombGaugeFixing::SetCoulombGauge; G (Nz+2,Ny,Nx,3,3) complex is updated in place."""
    Nx, Ny, Nz = f.N
    UG = Fields(f.N).U
    EG = np.zeros_like(f.E)
    stats = np.zeros(7)
    lib().ref_coulomb_gauge(ivec(f.N), dvec(a), P(f.U), P(f.E), P(G), P(UG), P(EG), tol, maxit, alpha, P(stats))
    return UG, EG, stats


ALGORITHMS = {"fourier": 0, "jacobi": 1, "losalamos": 2, "consistent": 3}


def coulomb_gauge_algo(f: Fields, a, G: np.ndarray, algo: str, tol: float, maxit: int, param: float, bc: int = 0):
    """SetCoulombGauge with the Fourier-accelerated (param = alpha), Jacobi (param = step size) or Los Alamos
    (param = over-relaxation probability, 0) update; G is updated in place.  bc: z boundary (0 open, 1 periodic),
    which decides how G's z ghost layers are synchronised."""
    UG = Fields(f.N).U
    EG = np.zeros_like(f.E)
    stats = np.zeros(7)
    lib().ref_coulomb_gauge_algo_bc(ivec(f.N), dvec(a), int(bc), P(f.U), P(f.E), P(G), P(UG), P(EG), ALGORITHMS[algo],
                                    tol, maxit, param, P(stats))
    return UG, EG, stats


def gauss_restore(f: Fields, a, variant: int = 0, tol: float = 1e-14, maxit: int = 30000, bc: int = 0):
    """GaussLawRestoration::Restore (variant 0) / GaussLawRestoration2D (variant 1): E relaxed in place toward the
    Gauss law with the static charges; returns (iterations, max |G|, RMS G)."""
    stats = np.zeros(3)
    lib().ref_gauss_restore(ivec(f.N), dvec(a), bc, P(f.U), P(f.E), P(f.JpS), P(f.JmS), variant, tol, maxit, P(stats))
    return stats


# ------------------------------------------------------------------------------------------------
# physical-charge (TMD / GBW) initial conditions
# ------------------------------------------------------------------------------------------------
class TmdParams:
    """Parameters of the physical-charge nucleus (SC/LATTICE/Parameters.cpp): qs0a, meff (lattice units),
    com_energy (fm^-1), afm = (aXInFm, aYInFm, aZInFm), A (atomic number), centers = (left, right) z index."""

    def __init__(self, qs0a, meff, com_energy, afm, A, centers):
        self.qs0a, self.meff, self.com_energy = float(qs0a), float(meff), float(com_energy)
        self.afm = tuple(float(v) for v in afm)
        self.A = int(A)
        self.centers = tuple(int(c) for c in centers)

    def arrays(self):
        return (np.array([self.qs0a, self.meff, self.com_energy, *self.afm]),
                np.array([self.A, *self.centers], np.int32))


def tmd_draw(seed: int, N, A: int):
    """The draws of SimulationmitPC::Run: (noise_l, noise_r) shaped (Nz, Ny, Nx, 8) and nucleon positions
    (pos_l, pos_r) shaped (A, 3) in fm."""
    nl = np.zeros((N[2], N[1], N[0], 8))
    nr = np.zeros_like(nl)
    pl = np.zeros((A, 3))
    pr = np.zeros((A, 3))
    rc = lib().ref_tmd_draw(int(seed), ivec(N), A, P(nl), P(nr), P(pl), P(pr))
    if rc != 0:
        raise ValueError(f"no Woods-Saxon parameters for A = {A}")
    return nl, nr, pl, pr


def charge_fluctuation(N, a, prm: TmdParams, noise: np.ndarray) -> np.ndarray:
    d, i = prm.arrays()
    out = np.zeros((N[2], N[1], N[0], 8))
    lib().ref_charge_fluctuation(ivec(N), dvec(a), P(d), i.ctypes.data_as(_ip), P(np.ascontiguousarray(noise)), P(out))
    return out


def nucleus_profile(prm: TmdParams, pos: np.ndarray, x: float, y: float, z: float) -> float:
    d, i = prm.arrays()
    return lib().ref_nucleus_profile(P(d), i.ctypes.data_as(_ip), P(np.ascontiguousarray(pos, np.float64)), x, y, z)


def tmd_initial_state(driver: int, N, a, dt: float, prm: TmdParams, noise_l, noise_r, pos_l, pos_r, bc: int = 0):
    """SimulationmitPC::Run's initial state; returns (Fields, rho_l, rho_r) with the charges before the
    transverse filter (ghost layout (Nz+2, Ny, Nx, 8))."""
    f = Fields(N)
    d, i = prm.arrays()
    rl = np.zeros((N[2] + 2, N[1], N[0], 8))
    rr = np.zeros_like(rl)
    lib().ref_tmd_initial_state(driver, ivec(N), dvec(a), bc, dt, P(d), i.ctypes.data_as(_ip),
                                P(np.ascontiguousarray(noise_l)), P(np.ascontiguousarray(noise_r)),
                                P(np.ascontiguousarray(pos_l, np.float64)), P(np.ascontiguousarray(pos_r, np.float64)),
                                P(f.U), P(f.UOld), P(f.E), P(f.Jp), P(f.Jm), P(f.JpS), P(f.JmS), P(rl), P(rr))
    return f, rl, rr


def max_trace_projection(V: np.ndarray) -> np.ndarray:
    """SUNcGroup::Extended::MaxTraceProjection of a 3x3 complex matrix (reference column-major storage)."""
    V = np.ascontiguousarray(V, np.complex128)
    out = np.zeros(9, np.complex128)
    lib().ref_max_trace_projection(P(V.reshape(9)), P(out))
    return out


def gauge_transform(f: Fields, a, G: np.ndarray, bc: int = 0):
    lib().ref_gauge_transform(ivec(f.N), dvec(a), bc, P(G), P(f.U), P(f.E), P(f.Jp), P(f.Jm), P(f.JpS), P(f.JmS))
