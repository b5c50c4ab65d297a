"""The largest model this suite runs on one GPU: a 512x256x256 Kerr torus
(33.5 M cells, ~34 GB of device buffers), built from the grid with the
metric tables on the device.  Without noise the torus is axisymmetric, and
the step keeps it so bit for bit: every phi row of the state after two steps
equals the first (a check that holds at any size and catches any index that
overflows or wraps).  The same torus built on the device
(raishin_model_set_torus) matches the host one within 1e-12 of each field's
maximum at this size too."""
from dataclasses import replace

import numpy as np
import pytest

from raishinsource_amd import configs, native

pytestmark = pytest.mark.gpu


def test_largest_model_stays_axisymmetric():
    p = replace(configs.C2_PPM, name="c2max", imax=512, jmax=256, kmax=256, ymax=None).resolved()
    g = native.setup_grid(p)
    mod = native.Model.from_grid(p, g)
    assert mod.device_bytes() > 30 * 2**30
    uri = native.setup_torus(p, g, mod.metric_row(), seed=0, noise=0.0)
    mod.set_torus(seed=0, noise=0.0)
    w_d, _, t_d = mod.get_state()
    assert t_d == 0.0
    for v in range(9):
        scale = max(float(np.abs(uri[v]).max()), 1e-300)
        err = float(np.abs(w_d[v] - uri[v]).max()) / scale
        assert err <= 1e-12, (v, err)
    del w_d
    mod.set_state(uri)
    assert mod.stored_rows() == p.jmax  # 246 interior phi cells: every row is kept
    t, dt = mod.step(2)
    assert np.isfinite(t) and dt > 0
    w, u, _ = mod.get_state()
    mod.close()
    for a in (w, u):
        v = a.view(np.uint64)
        # every phi row equal to row 0 (ghost rows included: periodic phi)
        assert np.array_equal(v, np.broadcast_to(v[:, :, 0:1, :], v.shape))
    n1 = p.nm1
    assert np.isfinite(w[:, n1:-n1, n1:-n1, n1:-n1]).all()
