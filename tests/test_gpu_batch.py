"""Batched stepping (raishin_batch_*): many models of different sizes,
spins and precisions advanced by one launch per kernel of the step, against
each model stepped alone with raishin_model_step.  Bit-identical state,
time and counters after every call, including odd step counts, the two-step
graph, and members joining and leaving."""
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


def make(p, mixed=False, seed=0, beta=100.0):
    p = p.resolved()
    g = native.setup_grid(p)
    m = native.setup_metric(p, g)
    uri = native.setup_torus(p, g, m, seed=seed, noise=0.02, beta=beta)
    mod = native.Model(p, g, m)
    mod.set_state(uri)
    mod.set_precision(mixed)
    return mod


def ppm_members():
    # (pram, mixed, seed, beta): RK2 PPM models of four shapes, two spins, both precisions
    return [
        (configs.T2D, False, 0, 100.0),
        (configs.C1, False, 1, 10.0),
        (configs.U_PPM, True, 2, 1000.0),
        (replace(configs.T2D, name="t2d_a5", akm=0.5, xmin=None), False, 3, 100.0),
        (replace(configs.T2D, name="t2d_mx", kmax=40), True, 4, 100.0),
    ]


def pairs(members):
    alone = [make(p, mx, sd, bt) for p, mx, sd, bt in members]
    batched = [make(p, mx, sd, bt) for p, mx, sd, bt in members]
    return alone, batched


def assert_same(alone, batched, what):
    for q, (x, y) in enumerate(zip(alone, batched)):
        wx, ux, tx = x.get_state()
        wy, uy, ty = y.get_state()
        assert tx == ty, (what, q)
        assert bitdiff(wx, wy) == 0 and bitdiff(ux, uy) == 0, (what, q)
        assert x.counters() == y.counters(), (what, q)


def close_all(*groups):
    for grp in groups:
        for mm in grp:
            mm.close()


def test_batch_ppm_matches_models_alone():
    alone, batched = pairs(ppm_members())
    b = native.Batch(batched)
    for mm in alone:
        mm.step(3)
    b.step(3)  # odd: the plan's second step comes next
    assert_same(alone, batched, "after 3")
    assert b.times() == [mm.step(0)[0] for mm in alone]
    for mm in alone:
        mm.step(4)
    b.step_graph(2)  # starts on the plan's second step
    assert_same(alone, batched, "after graph from odd")
    for mm in alone:
        mm.step(5)
    b.step(1)
    b.step_graph(2)
    assert_same(alone, batched, "after graph from even")
    b.close()
    close_all(alone, batched)


def test_batch_members_change():
    members = ppm_members()
    alone, batched = pairs(members)
    b = native.Batch(batched[:3])
    for mm in alone[:3]:
        mm.step(2)
    b.step(2)
    # one leaves, two join; a member stepped alone in between is picked up where it is
    batched[3].step(1)
    alone[3].step(1)
    b.set(batched[1:])
    for mm in alone[1:]:
        mm.step(3)
    b.step(3)
    assert_same(alone, batched, "after change")
    b.set([])
    b.close()
    close_all(alone, batched)


def test_batch_mp5_rk3():
    members = [(configs.U_MP5, False, 0, 100.0), (configs.T3D, False, 1, 100.0), (configs.U_MP5, True, 2, 10.0)]
    alone, batched = pairs(members)
    b = native.Batch(batched)
    for mm in alone:
        mm.step(3)
    b.step(1)
    b.step_graph(1)
    assert_same(alone, batched, "mp5")
    b.close()
    close_all(alone, batched)


def test_batch_mc():
    members = [(configs.U_MC, False, 0, 100.0), (configs.T2D_MC, True, 1, 10.0), (configs.U_MC, True, 2, 100.0),
               (replace(configs.T2D_MC, name="t2dmc_leis", icha=1), False, 3, 10.0),
               (replace(configs.U_MC, name="umc_gam", icha=0), True, 4, 100.0),
               (replace(configs.T2D_MC, name="t2dmc_rc2", iwvec=2), False, 5, 10.0)]
    alone, batched = pairs(members)
    b = native.Batch(batched)
    for mm in alone:
        mm.step(3)
    b.step(1)
    b.step_graph(1)  # one replay: two steps
    assert_same(alone, batched, "mc")
    b.close()
    close_all(alone, batched)


def test_batch_rejects_mixed_schemes():
    a = make(configs.U_PPM)
    c = make(configs.U_MP5)
    b = native.Batch()
    with pytest.raises(native.RaishinError, match="different step sequences"):
        b.set([a, c])
    with pytest.raises(native.RaishinError, match="twice"):
        b.set([a, a])
    b.close()
    close_all([a, c])


def test_workgroup_sizes_do_not_change_results():
    """The autotuner's choice, explicit odd choices, and members of one batch
    with different sizes: bit-identical to the default launch shapes."""
    members = [(configs.T2D, False, 0, 100.0), (configs.U_PPM, True, 1, 100.0),
               (replace(configs.U_PPM, name="uppm16", jmax=16), False, 2, 100.0)]
    ref = [make(p, mx, sd, bt) for p, mx, sd, bt in members]
    tuned = [make(p, mx, sd, bt) for p, mx, sd, bt in members]
    for (p, mx, sd, bt), mm in zip(members, tuned):
        pr = p.resolved()
        g = native.setup_grid(pr)
        m = native.setup_metric(pr, g)
        uri = native.setup_torus(pr, g, m, seed=sd, noise=0.02, beta=bt)
        sizes = mm.autotune(uri, steps=1)
        assert {sizes[k] for k in native.TUNE_SLOTS} <= {64, 128, 256}, sizes
    tuned[0].set_block_sizes([64, 192, 128, 64, 192, 128, 64, 192])
    for mm in ref:
        mm.step(3)
    for mm in tuned:
        mm.step(1)
    b = native.Batch(tuned)
    b.step(2)
    assert_same(ref, tuned, "tuned")
    b.close()
    close_all(ref, tuned)


def test_batch_member_switches_precision_between_steps():
    """A member switches form between batch steps (its stored fluxes are
    converted on its own stream, which the batch waits for): bit for bit the
    same as the same switches on the model stepped alone."""
    members = ppm_members()[:3]
    alone, batched = pairs(members)
    b = native.Batch(batched)
    for mm in alone:
        mm.step(2)
    b.step(2)
    for grp in (alone, batched):
        grp[0].set_precision(True)
        grp[2].set_precision(False)
    for mm in alone:
        mm.step(3)
    b.step(3)
    assert_same(alone, batched, "after the switch")
    for grp in (alone, batched):
        grp[0].set_precision(False)
    for mm in alone:
        mm.step(2)
    b.step_graph(1)
    assert_same(alone, batched, "after switching back, graph")
    b.set([])
    b.close()
    close_all(alone, batched)
