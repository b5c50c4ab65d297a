"""Set-up from the one phi row of an axisymmetric metric (setup_metric with
row=True, raishin_mdtorus1_ik): the same metric values as row j=1 of the full
reference-layout arrays and the same torus, bit for bit.  Host code only;
test_gpu_setup_rows.py checks the models built from the row."""
import numpy as np
import pytest
from dataclasses import replace

from raishinsource_amd import configs, native


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


@pytest.mark.parametrize("name", ["t2d", "uppm", "utube", "c1"])
def test_row_metric_and_torus_equal_full(name):
    p = configs.ALL[name].resolved()
    g = native.setup_grid(p)
    mr = native.setup_metric(p, g, row=True)
    mf = native.setup_metric(p, g)
    assert native.one_row(p, mr) and not native.one_row(p, mf)
    for k, a in mf.items():
        r = mr[k]
        full_row = a[..., 0:1, :]
        assert r.shape == full_row.shape, k
        assert np.array_equal(bits(r), bits(full_row)), k
        # and every row of the full arrays equals that row (the metric has no phi dependence)
        assert np.array_equal(bits(np.broadcast_to(full_row, a.shape)), bits(a)), k
    if p.metric == 303:
        for seed in (0, 5):
            ur = native.setup_torus(p, g, mr, seed=seed, noise=0.02, beta=10.0)
            uf = native.setup_torus(p, g, mf, seed=seed, noise=0.02, beta=10.0)
            assert np.array_equal(bits(ur), bits(uf)), seed


def test_torus_set_up_is_deterministic_with_threads():
    """The torus loops run on all host cores; the result does not depend on it."""
    p = replace(configs.T2D, name="t2d_big", imax=96, kmax=80).resolved()
    g = native.setup_grid(p)
    m = native.setup_metric(p, g, row=True)
    a = native.setup_torus(p, g, m, seed=7)
    b = native.setup_torus(p, g, m, seed=7)
    assert np.array_equal(bits(a), bits(b))
