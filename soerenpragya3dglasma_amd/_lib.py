"""Loader of libglasma.so (the HIP engine).  Fails loudly when the library is missing: there is no CPU path."""
from __future__ import annotations

import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "_lib")
# GLASMA_LIB selects another build of the same library (used for A/B measurements of build variants)
LIB_PATH = os.environ.get("GLASMA_LIB") or os.path.join(LIB_DIR, "libglasma.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "glasma.h")

_lib = None

dp = ctypes.POINTER(ctypes.c_double)
ip = ctypes.POINTER(ctypes.c_int)


class GlasmaError(RuntimeError):
    pass


class Params(ctypes.Structure):
    _fields_ = [("nc", ctypes.c_int), ("n", ctypes.c_int * 3), ("a", ctypes.c_double * 3), ("dt", ctypes.c_double),
                ("bc_z", ctypes.c_int)]


class MVParams(ctypes.Structure):
    _fields_ = [("lambda_", ctypes.c_double), ("m_eff", ctypes.c_double), ("sigma_l", ctypes.c_double),
                ("sigma_r", ctypes.c_double), ("nucleus_l", ctypes.c_double), ("nucleus_r", ctypes.c_double),
                ("center_l", ctypes.c_int), ("center_r", ctypes.c_int)]


class TmdParams(ctypes.Structure):
    _fields_ = [("qs0a", ctypes.c_double), ("m_eff", ctypes.c_double), ("com_energy", ctypes.c_double),
                ("a_fm", ctypes.c_double * 3), ("atomic_number", ctypes.c_int), ("center_l", ctypes.c_int),
                ("center_r", ctypes.c_int)]


def header_functions() -> list[str]:
    """Names of every function declared in include/glasma.h."""
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*|double|void)\s+(glasma_\w+)\s*\(", txt, re.M)))


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GlasmaError(f"libglasma.so not built ({LIB_PATH}); run __graft_entry__.build() or "
                          "python -m soerenpragya3dglasma_amd.build")
    L = ctypes.CDLL(LIB_PATH)
    c_int, c_long, c_double, c_char, c_ulong = ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_char, ctypes.c_ulong
    vp = ctypes.c_void_p
    sig = {
        "glasma_create": ([ctypes.POINTER(Params), ctypes.POINTER(vp)], c_int),
        "glasma_destroy": ([vp], c_int),
        "glasma_last_error": ([vp], ctypes.c_char_p),
        "glasma_version": ([], ctypes.c_char_p),
        "glasma_memory_bytes": ([vp, ctypes.POINTER(ctypes.c_size_t)], c_int),
        "glasma_set_fields": ([vp] + [dp] * 6, c_int),
        "glasma_get_fields": ([vp] + [dp] * 6, c_int),
        "glasma_set_uold": ([vp, dp], c_int),
        "glasma_get_uold": ([vp, dp], c_int),
        "glasma_get_time": ([vp, dp, ctypes.POINTER(c_long)], c_int),
        "glasma_set_time": ([vp, c_double, c_long], c_int),
        "glasma_update": ([vp, c_int], c_int),
        "glasma_update_async": ([vp, c_int], c_int),
        "glasma_synchronize": ([vp], c_int),
        "glasma_last_step_kernel_ms": ([vp, dp], c_int),
        "glasma_bulk": ([vp, dp], c_int),
        "glasma_bulk_noalloc": ([vp, dp], c_int),
        "glasma_gauss": ([vp, dp, dp, dp], c_int),
        "glasma_unitarity": ([vp, dp, dp], c_int),
        "glasma_emt": ([vp, dp], c_int),
        "glasma_checksum": ([vp, ctypes.POINTER(ctypes.c_uint64)], c_int),
        "glasma_field_profile": ([vp, dp], c_int),
        "glasma_current_profile": ([vp, dp], c_int),
        "glasma_transverse_energy_map": ([vp, dp], c_int),
        "glasma_draw_charges": ([c_ulong, c_int, c_int, c_double, c_double, dp, dp], c_int),
        "glasma_tmd_draw": ([c_ulong, c_int, c_int, c_int, c_int, dp, dp, dp, dp], c_int),
        "glasma_tmd_initial_state": ([vp, c_int, ctypes.POINTER(TmdParams), dp, dp, dp, dp], c_int),
        "glasma_mv_initial_state": ([vp, c_int, ctypes.POINTER(MVParams), dp, dp], c_int),
        "glasma_set_wilson_lines": ([vp, c_char, ctypes.POINTER(MVParams), dp, dp, dp], c_int),
        "glasma_coulomb_gauge": ([vp, c_double, c_int, c_double, dp], c_int),
        "glasma_coulomb_gauge_algorithm": ([vp, c_int, c_double, c_int, c_double, dp], c_int),
        "glasma_gauss_restore": ([vp, c_int, c_double, c_int, dp], c_int),
        "glasma_gauge_reset": ([vp], c_int),
        "glasma_get_gauge_fixed": ([vp, dp, dp, dp], c_int),
        "glasma_set_gauge_transformation": ([vp, dp], c_int),
        "glasma_fft_host": ([dp, c_int, ip, c_long, c_int], c_int),
        "glasma_fill_random": ([vp, c_ulong, c_double, c_double, c_double], c_int),
        "glasma_su3_batch": ([c_int, c_long, c_double, dp, dp, dp], c_int),
        "glasma_batch_create": ([ctypes.POINTER(Params), c_int, ctypes.POINTER(vp)], c_int),
        "glasma_batch_destroy": ([vp], c_int),
        "glasma_batch_size": ([vp, ctypes.POINTER(c_int)], c_int),
        "glasma_batch_context": ([vp, c_int, ctypes.POINTER(vp)], c_int),
        "glasma_batch_last_error": ([vp], ctypes.c_char_p),
        "glasma_batch_update": ([vp, c_int], c_int),
        "glasma_batch_update_serial": ([vp, c_int], c_int),
        "glasma_batch_last_ms": ([vp, dp], c_int),
        "glasma_save_configuration": ([vp, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p], c_int),
        "glasma_load_configuration": ([vp, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, c_double], c_int),
        "glasma_save_state": ([vp, ctypes.c_char_p], c_int),
        "glasma_load_state": ([vp, ctypes.c_char_p], c_int),
        "glasma_tune_step": ([vp, c_int, ip, ip, dp, ip], c_int),
        "glasma_tile_order": ([vp, ip, ip], c_int),
        "glasma_copy_ceiling": ([vp, c_int, dp], c_int),
        "glasma_batch_tune": ([vp, c_int], c_int),
        "glasma_slab_energy": ([vp, c_int, c_int, dp], c_int),
        "glasma_step_vjp": ([vp] + [dp] * 6, c_int),
        "glasma_slab_energy_from_charge": ([vp, ctypes.POINTER(MVParams), dp, c_int, c_int, c_int, dp], c_int),
        "glasma_slab_energy_gradient": ([vp, ctypes.POINTER(MVParams), dp, c_int, c_int, c_int, c_int, dp, dp], c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def check(rc: int, ctx=None, what: str = "") -> None:
    if rc != 0:
        msg = lib().glasma_last_error(ctx)
        raise GlasmaError(f"{what} failed with status {rc}: {msg.decode() if msg else ''}")
