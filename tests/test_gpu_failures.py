"""Failure handling forced from the host: a context larger than the device's memory is refused with
GLASMA_ERR_NOMEM (every partial allocation released), and the device stays usable afterwards."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_oversized_context_is_refused_and_the_device_recovers():
    from soerenpragya3dglasma_amd import Glasma, GlasmaError
    # 2048 x 2048 x 1024 sites need terabytes of fields, far beyond one device's memory
    with pytest.raises(GlasmaError, match=r"\(-3\).*out of memory"):
        Glasma((2048, 2048, 1024), (1.0, 1.0, 1.0), 0.25)
    with Glasma((8, 8, 8), (1.0, 1.0, 1.0), 0.25) as g:
        g.fill_random(seed=3, link_scale=0.3, e_scale=0.1, j_scale=0.05)
        g.Update(2)
        assert np.all(np.isfinite(g.bulk_raw()))



def _hip_runtime():
    """The HIP runtime libglasma.so itself uses (the ROCm installation's, among the process's mapped libraries), for
    filling device memory: torch's bundled runtime, mapped too once a test has imported torch, cannot share the
    device with it inside one process."""
    import ctypes
    from soerenpragya3dglasma_amd import _lib
    _lib.lib()
    with open("/proc/self/maps") as m:
        paths = {ln.split()[-1] for ln in m if "libamdhip64" in ln and "/torch/" not in ln}
    assert len(paths) == 1, paths
    hip = ctypes.CDLL(paths.pop())
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    return hip


def test_refused_gauge_allocation_leaves_the_context_retryable():
    """Device memory is filled (plain hipMalloc blocks) until the gauge-fixing arrays of a 128^3 context no longer
    fit: glasma_coulomb_gauge is refused with GLASMA_ERR_NOMEM, and after the memory is released the same context
    fixes the gauge exactly as a fresh context does."""
    import ctypes
    from soerenpragya3dglasma_amd import Glasma, GlasmaError
    hip = _hip_runtime()

    def free_bytes():
        f, t = ctypes.c_size_t(), ctypes.c_size_t()
        assert hip.hipMemGetInfo(ctypes.byref(f), ctypes.byref(t)) == 0
        return f.value

    # the filler takes blocks until hipMalloc itself refuses the smallest one (the runtime can reuse memory that
    # earlier contexts freed, which hipMemGetInfo does not count as free), so the 302 MB of G cannot fit
    N = (128, 128, 128)
    with Glasma(N, (1.0, 1.0, 1.0), 0.25) as g, Glasma(N, (1.0, 1.0, 1.0), 0.25) as ref:
        for c in (g, ref):
            c.fill_random(seed=5, link_scale=0.3, e_scale=0.1, j_scale=0.05)
        blocks = []
        try:
            for chunk in (1 << 30, 1 << 26, 1 << 22):
                while True:
                    p = ctypes.c_void_p()
                    if hip.hipMalloc(ctypes.byref(p), chunk) != 0:
                        break
                    blocks.append(p)
            hip.hipGetLastError()  # a refused block of the filler leaves no stale error for the next call
            print(f"filled {len(blocks)} blocks; {free_bytes()} bytes free", flush=True)
            with pytest.raises(GlasmaError, match=r"status -3: ensure_gauge: out of memory"):
                g.coulomb_gauge(tol=0.0, maxit=2, alpha=2.0)
        finally:
            for p in blocks:
                hip.hipFree(p)
        st = g.coulomb_gauge(tol=0.0, maxit=2, alpha=2.0)
        st_ref = ref.coulomb_gauge(tol=0.0, maxit=2, alpha=2.0)
        assert np.array_equal(st, st_ref)
        assert np.array_equal(g.get_gauge_fixed()[2], ref.get_gauge_fixed()[2])
