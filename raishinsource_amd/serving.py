"""Serving loop: independent solves arriving over time, advanced together.

Requests (1D special-relativistic MHD Riemann tubes, 2D and 3D magnetised
tori, each in FP64 or mixed precision) arrive as a Poisson process.  A
request is admitted when its device footprint fits the HBM budget (and a
member slot is free); its host set-up (grid, metric, initial data) and the
upload run on worker threads while the GPU keeps stepping.  A ready model
joins the running batch at the next tick; every tick advances all members
together (one launch per kernel of the step, raishin_batch_step) and
retires the members that are done, freeing their memory for waiting
requests.  That is continuous batching: membership changes between ticks,
not between batches.

A tube runs until its time exceeds t_end, checked before every step as the
reference's main loop checks time against tmax (main.f90:624); a torus runs
a fixed number of steps.  The tick length is the largest step count no
member can overrun, so each member takes exactly the steps it would take
alone; ``run_alone`` is that reference (one model, one step at a time).

The scheme is the same for every request (PPM, HLLE, flux-CT, RK2), so all
members share one batch; the metric (Kerr-Schild or Minkowski), size, spin,
field strength and precision differ per member.
"""
import math
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field, replace

import numpy as np

from . import configs, native

TORUS_SIZES_2D = [(64, 64), (128, 128), (256, 128), (512, 256)]


@dataclass
class Request:
    rid: int
    kind: str            # "tube", "torus2d", "torus3d"
    arrival: float       # seconds after the start of the run
    mixed: bool
    p: configs.Pram
    init: dict           # initial-data parameters
    t_end: float = None  # tubes: run while time <= t_end
    nsteps: int = None   # tori: fixed step count

    @property
    def zones(self):
        a, b, c = self.p.interior
        return a * b * c


@dataclass
class Result:
    rid: int
    kind: str
    mixed: bool
    arrival: float
    admitted: float = None
    started: float = None
    finished: float = None
    steps: int = 0
    time: float = 0.0
    zones: int = 0
    uri: np.ndarray = field(default=None, repr=False)

    @property
    def latency(self):
        return self.finished - self.arrival


def footprint(p):
    """Device bytes a model of pram p allocates (the allocations of
    build_model in csrc/rh_model.cpp, with raishin_model_device_bytes as the
    measured value)."""
    p = p.resolved()
    nm1 = p.nm1
    ni, nj, nk = p.imax, p.jmax, p.kmax
    n = ni * nj * nk
    p2 = ni * nk
    cap = (ni - 2 * nm1) * (nj - 2 * nm1) * (nk - 2 * nm1)
    need_wo = any(2 in getattr(p, a) for a in ("iboux1in", "iboux1ot", "iboux2in", "iboux2ot", "iboux3in",
                                               "iboux3ot"))
    ncfl = max((max(ni - 2, 0) * max(nj - 2, 0) * max(nk - 2, 0) + 255) // 256, 1)
    planes = 7 * 9 + 27 + 6 + 1 + 9 + (9 if need_wo else 0) + (18 if p.irec in (1, 5) else 0)
    b = 8 * n * planes
    b += n + 8 * (2 + cap) + cap                     # cflag, carry count + list + counter, pend
    b += 8 * 64 * p2 + 2 * 8 * ni + 4 * 54 + 8 * 4 + 8 * 11 + 8 * 2
    b += 2 * 8 * 4 * ncfl                            # CFL per-block results
    b += 8 * 11 * ((n + 255) // 256) * 4             # per-wavefront counter rows
    b += 4 * 8 * 33 * p2                             # metric tables
    b += sum(8 * (4 * d + 2) for d in (ni, nj, nk))  # xs, dxb, dxs, xa
    if p.irec == 5:
        b += sum(8 * 7 * (d + 2) for d in (ni, nj, nk))
    return b


def draw_torus2d(rng, name, seed):
    """One 2D torus of the ensemble and of the serving mix: r x theta drawn
    from TORUS_SIZES_2D, spin a ~ U[0, 0.99], beta_min from {10, 100, 1000},
    r_in ~ U[6, 10], r_max = r_in + 6; C1's scheme (PPM, HLLE, flux-CT, RK2,
    CFL 0.4).  Returns (pram, initial-data parameters with the noise seed)."""
    ni, nk = TORUS_SIZES_2D[rng.integers(len(TORUS_SIZES_2D))]
    a = float(rng.uniform(0.0, 0.99))
    beta = float(rng.choice([10.0, 100.0, 1000.0]))
    rin = float(rng.uniform(6.0, 10.0))
    p = replace(configs.C1, name=name, imax=ni + 2 * configs.C1.nm1, kmax=nk + 2 * configs.C1.nm1, akm=a,
                xmin=None)
    return p, {"rin": rin, "rmax": rin + 6.0, "beta": beta, "seed": seed}


def make_ensemble(n, seed=2024):
    """n tori drawn by draw_torus2d from one generator (the noise seed of
    member q is q).  Returns [(pram, initial-data parameters)]."""
    rng = np.random.default_rng(seed)
    return [draw_torus2d(rng, f"ens{q}", q) for q in range(n)]


def build_torus(p, init):
    """Model (metric tables built on the device) and initial primitives of a
    torus member: seeded 2% pressure noise."""
    p = p.resolved()
    g = native.setup_grid(p)
    mod = native.Model.from_grid(p, g)
    uri = native.setup_torus(p, g, mod.metric_row(), rin=init["rin"], rmax=init["rmax"], beta=init["beta"],
                             noise=0.02, seed=init["seed"])
    return mod, uri


def make_requests(n, rate, seed=0, mix=(0.6, 0.3, 0.1), mixed_fraction=0.5, torus2d_steps=500,
                  torus3d_steps=100, tube_t_end=0.4):
    """n requests with exponential inter-arrival times of mean 1/rate."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(n):
        t += float(rng.exponential(1.0 / rate))
        u = rng.random()
        mixed = bool(rng.random() < mixed_fraction)
        if u < mix[0]:
            left, right = native.random_tube_states(rng)
            out.append(Request(rid, "tube", t, mixed, configs.TUBE, {"left": left, "right": right},
                               t_end=tube_t_end))
        elif u < mix[0] + mix[1]:
            p, init = draw_torus2d(rng, f"torus2d_{rid}", rid)
            out.append(Request(rid, "torus2d", t, mixed, p, init, nsteps=torus2d_steps))
        else:
            g = 2 * configs.C1.nm1
            p = replace(configs.C1, name=f"torus3d_{rid}", imax=64 + g, jmax=64 + g, kmax=64 + g, ymax=None)
            out.append(Request(rid, "torus3d", t, mixed, p, {"rin": 6.0, "rmax": 12.0, "beta": 100.0,
                                                           "seed": rid}, nsteps=torus3d_steps))
    return out


def build(req):
    """Device model and initial state of one request: metric tables built on
    the device, a torus too (raishin_model_set_torus: no host arrays), a tube
    on the host."""
    p = req.p.resolved()
    g = native.setup_grid(p)
    mod = native.Model.from_grid(p, g)  # metric tables built on the device
    if req.kind == "tube":
        mod.set_state(native.setup_tube(p, req.init["left"], req.init["right"]))
    else:
        mod.set_torus(rin=req.init["rin"], rmax=req.init["rmax"], beta=req.init["beta"], noise=0.02,
                      seed=req.init["seed"])
    mod.set_precision(req.mixed)
    mod.sync()
    return mod


def safe_steps(req, steps, t, dt):
    """Steps a request can take with no chance of overrunning its end: a
    torus's remaining count; for a tube, half the time left over the last
    dt (at least one), so time stays <= t_end before each of them unless dt
    more than doubles within them (dt follows the CFL limit, which moves
    slowly)."""
    if req.kind != "tube":
        return req.nsteps - steps
    if t > req.t_end:
        return 0
    if not dt or dt <= 0.0 or not math.isfinite(dt):
        return 1
    return max(1, int(0.5 * (req.t_end - t) / dt))


def run_alone(req, keep_state=True, strict=True):
    """The request by itself on one model, stopping as the reference's main
    loop does (time checked before every step).  strict: one step at a time
    (the reference for the served results); else in safe_steps chunks, as a
    one-request server would.  Returns (Result, seconds of the build,
    seconds of the steps)."""
    t0 = time.perf_counter()
    mod = build(req)
    t1 = time.perf_counter()
    steps = 0
    t = dt = 0.0
    if req.kind == "tube":
        while t <= req.t_end:
            k = 1 if strict else safe_steps(req, steps, t, dt)
            t_prev = t
            t, _ = mod.step(k)
            dt = (t - t_prev) / k
            steps += k
    else:
        t, _ = mod.step(req.nsteps)
        steps = req.nsteps
    t2 = time.perf_counter()
    r = Result(req.rid, req.kind, req.mixed, req.arrival, steps=steps, time=t, zones=req.zones)
    if keep_state:
        r.uri = mod.get_state()[0]
    mod.close()
    return r, t1 - t0, t2 - t1


class Server:
    """Continuous batching of arriving requests on one GPU."""

    def __init__(self, hbm_budget=64 << 30, max_members=512, max_tick_steps=32, workers=4, keep_state=False):
        self.budget = int(hbm_budget)
        self.max_members = max_members
        self.max_tick = max_tick_steps
        self.keep_state = keep_state
        self.pool = ThreadPoolExecutor(max_workers=workers)
        self.batch = native.Batch()
        self.ticks = 0
        self.tick_steps = 0
        self.deferred = set()  # requests that waited for memory or a slot after arriving

    def close(self):
        self.pool.shutdown(wait=True)
        self.batch.set([])
        self.batch.close()

    def run(self, requests):
        """Serve the requests; returns their Results in request order."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        results = {r.rid: Result(r.rid, r.kind, r.mixed, r.arrival, zones=r.zones) for r in reqs}
        waiting = []      # arrived, not admitted
        building = []     # (req, future, bytes)
        members = []      # dicts: req, model, steps, t, dt, bytes
        used = 0
        nxt = 0
        start = time.perf_counter()

        def clock():
            return time.perf_counter() - start

        while nxt < len(reqs) or waiting or building or members:
            now = clock()
            while nxt < len(reqs) and reqs[nxt].arrival <= now:
                waiting.append(reqs[nxt])
                nxt += 1
            # admission, in arrival order, by device footprint
            while waiting:
                req = waiting[0]
                need = footprint(req.p)
                if need > self.budget:
                    raise ValueError(f"request {req.rid} needs {need} bytes, over the whole budget")
                if used + need > self.budget or len(members) + len(building) >= self.max_members:
                    self.deferred.add(req.rid)
                    break
                waiting.pop(0)
                used += need
                results[req.rid].admitted = now
                building.append((req, self.pool.submit(build, req), need))
            # models whose set-up has finished join the batch
            joined = False
            still = []
            for req, fut, need in building:
                if fut.done():
                    members.append({"req": req, "model": fut.result(), "steps": 0, "t": 0.0, "dt": 0.0,
                                    "bytes": need})
                    results[req.rid].started = now
                    joined = True
                else:
                    still.append((req, fut, need))
            building = still
            if not members:
                if building:
                    time.sleep(0.0005)
                elif nxt < len(reqs) and not waiting:
                    time.sleep(max(0.0, min(reqs[nxt].arrival - clock(), 0.005)))
                continue
            if joined or len(members) != len(self.batch.models):
                self.batch.set([mm["model"] for mm in members])
            k = min(self.max_tick, min(safe_steps(mm["req"], mm["steps"], mm["t"], mm["dt"]) for mm in members))
            if k >= 2:
                self.batch.step_graph(k // 2)
            if k % 2:
                self.batch.step(1)
            self.ticks += 1
            self.tick_steps += k
            times = self.batch.times()
            now = clock()
            done, keep = [], []
            for mm, t in zip(members, times):
                mm["dt"] = (t - mm["t"]) / k if k else mm["dt"]
                mm["t"] = t
                mm["steps"] += k
                req = mm["req"]
                fin = (t > req.t_end) if req.kind == "tube" else (mm["steps"] >= req.nsteps)
                (done if fin else keep).append(mm)
            if done:
                members = keep
                self.batch.set([mm["model"] for mm in members])
                for mm in done:
                    r = results[mm["req"].rid]
                    r.finished = now
                    r.steps = mm["steps"]
                    r.time = mm["t"]
                    if self.keep_state:
                        r.uri = mm["model"].get_state()[0]
                    mm["model"].close()
                    used -= mm["bytes"]
        return [results[r.rid] for r in requests]


def summarize(results, wall):
    lat = np.array([r.latency for r in results])
    zu = sum(r.zones * r.steps for r in results)
    return {
        "requests": len(results),
        "p50_s": float(np.percentile(lat, 50)),
        "p99_s": float(np.percentile(lat, 99)),
        "mean_s": float(lat.mean()),
        "zone_updates_per_s": zu / wall if wall > 0 else None,
        "wall_s": wall,
        "by_kind": {k: int(sum(1 for r in results if r.kind == k)) for k in ("tube", "torus2d", "torus3d")},
    }


def one_at_a_time(requests, keep_state=False):
    """The same requests served by one worker, one request at a time in
    arrival order: a request starts when it has arrived and the previous one
    has finished (build and steps measured, queueing computed)."""
    free_at = 0.0
    out = []
    for req in sorted(requests, key=lambda r: r.arrival):
        r, t_build, t_steps = run_alone(req, keep_state=keep_state)
        start = max(req.arrival, free_at)
        r.started = start
        r.finished = start + t_build + t_steps
        free_at = r.finished
        out.append(r)
    order = {r.rid: q for q, r in enumerate(requests)}
    out.sort(key=lambda r: order[r.rid])
    return out
