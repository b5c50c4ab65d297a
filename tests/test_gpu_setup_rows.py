"""Models built from the one phi row of an axisymmetric metric
(raishin_model_create_ik) step bit for bit like models built from the
reference-layout arrays (raishin_model_create)."""
import numpy as np
import pytest

from raishinsource_amd import configs, native

pytestmark = pytest.mark.gpu


def bitdiff(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    both_nan = np.isnan(a) & np.isnan(b)
    return int(np.count_nonzero((a.view(np.uint64) != b.view(np.uint64)) & ~both_nan))


@pytest.mark.parametrize("name,nsteps", [("t2d", 6), ("uppm", 4), ("ump5", 3), ("umc", 4)])
def test_row_model_steps_like_full_model(name, nsteps):
    p = configs.ALL[name].resolved()
    g = native.setup_grid(p)
    out = []
    for row in (True, False):
        m = native.setup_metric(p, g, row=row)
        uri = native.setup_torus(p, g, m, seed=2, noise=0.02)
        mod = native.Model(p, g, m)
        mod.set_state(uri)
        t, dt = mod.step(nsteps)
        w, u, _ = mod.get_state()
        out.append((t, dt, w, u, mod.counters(), mod.device_bytes()))
        mod.close()
    (ta, da, wa, ua, ca, ba), (tb, db, wb, ub, cb, bb) = out
    assert (ta, da) == (tb, db)
    assert bitdiff(wa, wb) == 0 and bitdiff(ua, ub) == 0
    assert ca == cb and ba == bb


def test_row_tube_model():
    p = configs.U_TUBE.resolved()
    g = native.setup_grid(p)
    left, right = native.random_tube_states(np.random.default_rng(3))
    uri = native.setup_tube(p, left, right)
    res = []
    for row in (True, False):
        m = native.setup_metric(p, g, row=row)
        mod = native.Model(p, g, m)
        mod.set_state(uri)
        mod.step(10)
        res.append(mod.get_state())
        mod.close()
    assert res[0][2] == res[1][2]
    assert bitdiff(res[0][0], res[1][0]) == 0 and bitdiff(res[0][1], res[1][1]) == 0
