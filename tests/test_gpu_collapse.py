"""One stored phi row for axisymmetric models with one interior phi cell:
bit-identical to the same model keeping every row (state, time, dt,
counters, the stage intermediates), eagerly, through graphs and in batches;
and the cases that must keep every row."""
from dataclasses import replace

import numpy as np
import pytest

from raishinsource_amd import configs, native

pytestmark = pytest.mark.gpu


def bitdiff(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    both_nan = np.isnan(a) & np.isnan(b)
    return int(np.count_nonzero((a.view(np.uint64) != b.view(np.uint64)) & ~both_nan))


def pair(p, mixed=False, seed=0, beta=100.0):
    p = p.resolved()
    g = native.setup_grid(p)
    m = native.setup_metric(p, g)
    uri = native.setup_torus(p, g, m, seed=seed, noise=0.02, beta=beta)
    out = []
    for allow in (True, False):
        mod = native.Model(p, g, m)
        mod.allow_collapse(allow)
        mod.set_state(uri)
        mod.set_precision(mixed)
        out.append(mod)
    return out, uri


def same(a, b, tag):
    wa, ua, ta = a.get_state()
    wb, ub, tb = b.get_state()
    assert ta == tb, tag
    assert bitdiff(wa, wb) == 0 and bitdiff(ua, ub) == 0, tag
    assert a.counters() == b.counters(), tag


@pytest.mark.parametrize("name,mixed", [("t2d", False), ("c1", False), ("t2d", True), ("t2dmc", False), ("t2drc2", False)])
def test_collapsed_rows_match_full(name, mixed):
    (col, full), _ = pair(configs.ALL[name], mixed)
    assert col.stored_rows() == 1 and full.stored_rows() == col.p.jmax
    for s in range(3):
        tc, dc = col.step(1)
        tf, df = full.step(1)
        assert (tc, dc) == (tf, df), s
    same(col, full, "eager")
    col.step_graph(2)
    full.step_graph(2)
    same(col, full, "graph")
    assert col.cdtcfl4() == full.cdtcfl4()
    wwc, wpc = col.stage(1, want_ww=True, want_wpre=True)
    wwf, wpf = full.stage(1, want_ww=True, want_wpre=True)
    # the rows the update reads: fluxes on flux-CT's phi range (outside it
    # the full model's transverse-B fluxes are the sweeps' unaveraged ones,
    # which nothing reads), recovered primitives on the interior row
    p = col.p
    rows = slice(p.nm1 - 1, p.jmax - p.nm1)
    assert bitdiff(wwc[:, :, :, rows, :], wwf[:, :, :, rows, :]) == 0
    assert bitdiff(wpc[:, :, p.nm1, :], wpf[:, :, p.nm1, :]) == 0
    col.close()
    full.close()


def test_collapsed_members_in_a_batch():
    (c1, f1), _ = pair(configs.T2D, False, seed=1)
    (c2, f2), _ = pair(replace(configs.T2D, name="t2d_b", akm=0.5, xmin=None, kmax=40), True, seed=2, beta=10.0)
    (c3, f3), _ = pair(configs.U_PPM, False, seed=3)  # 3D: keeps every row
    assert c3.stored_rows() == configs.U_PPM.jmax
    b = native.Batch([c1, c2, c3, f1])
    b.step(3)
    b.step_graph(1)
    for mm in (f2, f3):
        mm.step(5)
    same(c1, f1, "batch t2d")
    same(c2, f2, "batch t2d mixed")
    b.set([])
    b.close()
    for mm in (c1, f1, c2, f2, c3, f3):
        mm.close()


def test_rows_kept_when_they_differ():
    p = configs.T2D.resolved()
    g = native.setup_grid(p)
    m = native.setup_metric(p, g)
    uri = native.setup_torus(p, g, m, seed=0, noise=0.02)
    uri[4, :, 3, :] *= 1.0 + 1e-12  # one ghost row differs
    mod = native.Model(p, g, m)
    mod.set_state(uri)
    assert mod.stored_rows() == p.jmax
    w, _, _ = mod.get_state()
    assert bitdiff(w, uri) == 0
    mod.close()
