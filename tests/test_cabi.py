"""The C-ABI library (libglasma.so): it loads without a GPU, exports every function include/glasma.h
declares, and rejects invalid arguments before touching the device.  CPU only: no compute call."""
import ctypes

import numpy as np

from soerenpragya3dglasma_amd import _lib


def test_header_lists_the_interface():
    names = _lib.header_functions()
    for n in ("glasma_create", "glasma_update", "glasma_bulk", "glasma_gauss", "glasma_emt", "glasma_unitarity",
              "glasma_mv_initial_state", "glasma_coulomb_gauge", "glasma_fft_host", "glasma_draw_charges"):
        assert n in names
    assert len(names) >= 29


def test_library_exports_every_declared_function():
    L = _lib.lib()
    missing = [n for n in _lib.header_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_argument_errors_are_reported_without_a_device():
    L = _lib.lib()
    p = _lib.Params()
    p.nc = 2  # only SU(3) runs on the device
    p.n[0] = p.n[1] = p.n[2] = 4
    p.a[0] = p.a[1] = p.a[2] = 1.0
    p.dt = 0.1
    h = ctypes.c_void_p()
    assert L.glasma_create(ctypes.byref(p), ctypes.byref(h)) == -1 and not h.value
    p.nc, p.bc_z = 3, 7
    assert L.glasma_create(ctypes.byref(p), ctypes.byref(h)) == -1
    assert L.glasma_create(None, ctypes.byref(h)) == -1
    x = np.zeros(16)
    dims = (ctypes.c_int * 1)(8)
    assert L.glasma_fft_host(x.ctypes.data_as(_lib.dp), 1, dims, 1, 3) == -1  # sign must be +-1
    assert L.glasma_su3_batch(9, 1, 0.0, x.ctypes.data_as(_lib.dp), None, x.ctypes.data_as(_lib.dp)) == -1
    assert L.glasma_draw_charges(1, 0, 4, 0.5, 0.5, x.ctypes.data_as(_lib.dp), x.ctypes.data_as(_lib.dp)) == -1
    assert L.glasma_version().decode().startswith("glasma")
    assert L.glasma_destroy(None) == 0


def test_creation_errors_are_reported_without_a_context():
    """glasma_create checks its parameters before touching a device, and glasma_last_error(NULL) explains
    a failed creation."""
    import ctypes
    from soerenpragya3dglasma_amd import _lib
    L = _lib.lib()
    h = ctypes.c_void_p()
    p = _lib.Params()
    p.nc = 2
    for k in range(3):
        p.n[k], p.a[k] = 4, 1.0
    p.dt = 0.1
    assert L.glasma_create(ctypes.byref(p), ctypes.byref(h)) == -1  # GLASMA_ERR_ARG
    assert b"SU(3)" in L.glasma_last_error(None)
    p.nc, p.n[2] = 3, 0
    assert L.glasma_create(ctypes.byref(p), ctypes.byref(h)) == -1
    assert b"positive" in L.glasma_last_error(None)
    assert not h.value
