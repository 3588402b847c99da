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
