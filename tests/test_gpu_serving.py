"""The serving loop (continuous batching): each request's result is the same
served in the mix as run alone, bit for bit (state, step count, time), with
admission held back by a small HBM budget so requests queue and join the
running batch as others finish.  Also the device footprint estimate against
the bytes a model actually allocates."""
from dataclasses import replace

import numpy as np
import pytest

from raishinsource_amd import configs, native, serving

pytestmark = pytest.mark.gpu


def bitdiff(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    both_nan = np.isnan(a) & np.isnan(b)
    return int(np.count_nonzero((a.view(np.uint64) != b.view(np.uint64)) & ~both_nan))


def small_requests(n, rate, seed):
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(n):
        t += float(rng.exponential(1.0 / rate))
        u = rng.random()
        mixed = bool(rng.random() < 0.5)
        if u < 0.5:
            left, right = native.random_tube_states(rng)
            out.append(serving.Request(rid, "tube", t, mixed, configs.U_TUBE, {"left": left, "right": right},
                                       t_end=float(rng.uniform(0.02, 0.06))))
        elif u < 0.85:
            ni, nk = [(24, 24), (40, 32), (54, 60)][rng.integers(3)]
            p = replace(configs.T2D, name=f"t{rid}", imax=ni + 10, kmax=nk + 10, akm=float(rng.uniform(0, 0.99)),
                        xmin=None)
            rin = float(rng.uniform(6, 10))
            out.append(serving.Request(rid, "torus2d", t, mixed, p, {"rin": rin, "rmax": rin + 6, "beta": 100.0,
                                                                   "seed": rid},
                                       nsteps=int(rng.integers(5, 30))))
        else:
            p = replace(configs.U_PPM, name=f"t3_{rid}")
            out.append(serving.Request(rid, "torus3d", t, mixed, p, {"rin": 6.0, "rmax": 12.0, "beta": 100.0,
                                                                   "seed": rid},
                                       nsteps=int(rng.integers(3, 12))))
    return out


def test_footprint_matches_allocation():
    for p in (configs.U_TUBE, configs.T2D, configs.U_PPM, configs.U_MP5, configs.U_MC,
              replace(configs.C1, imax=74, jmax=74, kmax=74, ymax=None)):
        p = p.resolved()
        g = native.setup_grid(p)
        m = native.setup_metric(p, g)
        mod = native.Model(p, g, m)
        assert mod.device_bytes() == serving.footprint(p), p.name
        mod.close()


def test_served_results_equal_alone():
    reqs = small_requests(24, 400.0, seed=7)
    budget = 3 * max(serving.footprint(r.p) for r in reqs)
    srv = serving.Server(hbm_budget=budget, max_tick_steps=4, workers=4, keep_state=True)
    try:
        served = srv.run(reqs)
    finally:
        srv.close()
    assert srv.ticks > 0
    print("ticks", srv.ticks, "steps", srv.tick_steps, "deferred", len(srv.deferred),
          serving.summarize(served, 1.0))
    assert srv.deferred  # the budget held some requests back until others finished
    for req, r in zip(reqs, served):
        alone, _, _ = serving.run_alone(req)
        assert r.steps == alone.steps and r.time == alone.time, (req.rid, req.kind, r.steps, alone.steps)
        assert bitdiff(r.uri, alone.uri) == 0, (req.rid, req.kind)
        assert r.finished >= r.started >= r.admitted >= r.arrival
