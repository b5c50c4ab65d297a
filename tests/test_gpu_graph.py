"""Graph-captured stepping (two steps per HIP graph) and the concurrent
multi-model step against eager stepping: bit-identical after every replay,
with eager work and synchronisation between the replays."""
import numpy as np
import pytest

from raishinsource_amd import configs, native

pytestmark = pytest.mark.gpu


def bitdiff(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    both_nan = np.isnan(a) & np.isnan(b)
    return int(np.count_nonzero((a.view(np.uint64) != b.view(np.uint64)) & ~both_nan))


def build(name, mixed=False):
    p = configs.ALL[name].resolved()
    g = native.setup_grid(p)
    m = native.setup_metric(p, g)
    uri = native.setup_torus(p, g, m, seed=0, noise=0.02)
    mod = native.Model(p, g, m)
    mod.set_state(uri)
    mod.set_precision(mixed)
    return mod


@pytest.mark.parametrize("name,mixed", [("t2d", False), ("ump5", False), ("uppm", True), ("t2dmc", True), ("urc2", False)])
def test_graph_replays_match_eager(name, mixed):
    a, b = build(name, mixed), build(name, mixed)
    for rep in range(3):
        ta, dta = a.step(2)
        tb, dtb = b.step_graph(1)
        wa, ua, _ = a.get_state()
        wb, ub, _ = b.get_state()
        assert (ta, dta) == (tb, dtb), rep
        assert bitdiff(wa, wb) == 0 and bitdiff(ua, ub) == 0, rep
    # an odd number of eager steps moves the buffers: the graph is captured again
    a.step(1)
    b.step(1)
    a.step(2)
    b.step_graph(1)
    wa, ua, _ = a.get_state()
    wb, ub, _ = b.get_state()
    assert bitdiff(wa, wb) == 0 and bitdiff(ua, ub) == 0
    assert a.counters() == b.counters()
    a.close()
    b.close()


def test_models_step_concurrently_match_eager():
    names = ["t2d", "uppm", "ump5", "c1"]
    eager = [build(n) for n in names]
    conc = [build(n) for n in names]
    for mm in eager:
        mm.step(4)
    native.step_models_graph(conc, 2)
    for n, x, y in zip(names, eager, conc):
        wx, ux, tx = x.get_state()
        wy, uy, ty = y.get_state()
        assert tx == ty, n
        assert bitdiff(wx, wy) == 0 and bitdiff(ux, uy) == 0, n
    for mm in eager + conc:
        mm.close()
