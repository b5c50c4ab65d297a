"""PPM sweep tiles: the fused sweep (face values through LDS; along j and k
a wave holds `seg` cells along the axis by 64 / seg consecutive i) against
the split form (face values through
HBM from their own kernel), bit for bit: state, time, counters and the
stored fluxes, for 2D and 3D tori, one stored phi row, FP64 and mixed,
segment lengths that do and do not divide the extents, alone and in a
batch."""
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


def make(p, mixed=False, seed=0):
    p = p.resolved()
    g = native.setup_grid(p)
    m = native.setup_metric(p, g)
    uri = native.setup_torus(p, g, m, seed=seed, noise=0.02)
    mod = native.Model(p, g, m)
    mod.set_state(uri)
    mod.set_precision(mixed)
    return mod


def same(a, b, tag):
    wa, ua, ta = a.get_state()
    wb, ub, tb = b.get_state()
    assert ta == tb, tag
    assert bitdiff(wa, wb) == 0 and bitdiff(ua, ub) == 0, tag
    assert a.counters() == b.counters(), tag
    fa, ca = a.fluxes()
    fb, cb = b.fluxes()
    assert bitdiff(fa, fb) == 0 and bitdiff(ca, cb) == 0, tag


CASES = [
    ("t3dppm", False, (1, 0, 0), (0, 8, 16)),
    ("t3dppm", True, (1, 0, 0), (0, 64, 4)),
    ("uppm", False, (1, 0, 0), (0, 2, 32)),
    ("uppm", True, (1, 0, 0), (0, 0, 0)),
    ("c1", False, (1, 0, 0), (0, 16, 8)),      # one stored phi row
    ("t2d", True, (1, 0, 0), (0, 4, 64)),      # one stored phi row, mixed
]


@pytest.mark.parametrize("name,mixed,ta,tb", CASES)
def test_fused_sweep_equals_split(name, mixed, ta, tb):
    a, b = make(configs.ALL[name], mixed), make(configs.ALL[name], mixed)
    a.set_sweep_tiles(*ta)
    b.set_sweep_tiles(*tb)
    assert a.sweep_tiles()["ppm_split"] == ta[0] and b.sweep_tiles()["ppm_split"] == tb[0]
    for s in range(3):
        assert a.step(1) == b.step(1), s
    same(a, b, "eager")
    a.step_graph(1)
    b.step_graph(1)
    same(a, b, "graph")
    a.close()
    b.close()


def test_fused_sweep_in_a_batch():
    # RK2 members (a batch's members make the same launch sequence)
    ps = [configs.T2D, configs.U_PPM, replace(configs.T3D_PPM, name="t3dppm_b", jmax=20, irkt=2)]
    alone = [make(p, q == 1, seed=q) for q, p in enumerate(ps)]
    batched = [make(p, q == 1, seed=q) for q, p in enumerate(ps)]
    for q, mm in enumerate(alone):
        mm.set_sweep_tiles(1, 0, 0)
    for q, mm in enumerate(batched):
        mm.set_sweep_tiles(0, 4 << q, 32 >> q)
    for mm in alone:
        mm.step(4)
    b = native.Batch(batched)
    b.step(1)
    b.step_graph(1)
    b.step(1)
    for x, y in zip(alone, batched):
        same(x, y, "batch")
    b.set([])
    b.close()
    for mm in alone + batched:
        mm.close()


def test_autotuner_picks_tiles():
    p = configs.T3D_PPM.resolved()
    g = native.setup_grid(p)
    m = native.setup_metric(p, g)
    uri = native.setup_torus(p, g, m, seed=0, noise=0.02)
    ref = native.Model(p, g, m)
    ref.set_state(uri)
    mod = native.Model(p, g, m)
    tuned = mod.autotune(uri, steps=1)
    print("tuned", tuned)
    assert tuned["ppm_split"] == 0 and tuned["seg_j"] in (8, 16, 32) and tuned["seg_k"] in (8, 16, 32)
    with pytest.raises(native.RaishinError, match="power of two"):
        mod.set_sweep_tiles(0, 6, 8)
    ref.step(3)
    mod.step(3)
    same(ref, mod, "tuned")
    ref.close()
    mod.close()
