#!/usr/bin/env python3
"""Benchmark of the GRMHD time step on one MI355X.

Workload: one GPU's 3D subdomain at its real shape -- a magnetised
Fishbone-Moncrief torus in Kerr-Schild coordinates, 256 (r) x 128 (phi) x
128 (theta) cells with ghosts, PPM reconstruction, HLLE, flux-CT, RK3, FP64
throughout.  A step is one full time step (CFL reduction + three RK stages)
over the whole subdomain with the state resident in HBM.  (With MP5 or WENO5
the reference itself turns this torus to NaN within a few steps; --config c2
runs MP5 anyway.)

Prints ONE JSON line: zone-updates/s (interior zones x steps / wall time of
the timed steps); the roofline of the whole step (compulsory state bytes per
zone-update x zone-updates/s against 8 TB/s), with the bytes the step really
moves and its VALU issue time from the committed rocprofv3 counter runs, and
the dominant kernel's own figure from HIP-event timings taken on the
model's stream during the timed steps (events around that kernel's launches
only; the per-kernel breakdown comes from the warm-up steps, which time
every kernel); and the CPU baseline (the reference's own Fortran routines,
one process per host core on phi slabs of the same workload, with the C
restatement in oracle/ on one core beside it).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E peak, GB/s


def algorithmic_bytes(p, kind):
    """Compulsory HBM bytes of one launch of a kernel kind (FP64 state).

    sweep_*: per interface, read the two face states (PPM: 18 edge values
             from ppm_edges; MP5/WENO5: the 9 primitive planes once) and
             write 9 flux planes, plus the face metric table (33 doubles per
             (i,k)).
    ppm_edges: per cell and direction, read the 9 primitive planes, write
             18 edge planes (all three directions per launch kind).
    update:  read W, U (one RK input at stage 1, two after), 3x9 fluxes;
             write the 9 RK-result planes; centre metric + Christoffel tables
             (97 doubles per (i,k)).
    recov:   read the RK result and W, write 9 recovered planes and the
             final W of the Newton iteration; centre metric table.
    post:    read 9 recovered planes, write 9 primitive + 9 conserved
             planes per cell, centre metric table.
    """
    ni, nj, nk, nm0, nm1 = p.imax, p.jmax, p.kmax, p.nm0, p.nm1
    P2 = ni * nk
    if kind.startswith("sweep"):
        units = (ni - 2 * nm0 + 1) * (nj - 2 * nm0 + 1) * (nk - 2 * nm0 + 1)
        per = 216.0 if p.irec == 5 else 144.0
        return units, per * units + 264.0 * P2, "interfaces"
    if kind == "ppm_edges":
        units = (ni - 2 * nm0 + 2) * (nj - 2 * nm0 + 1) * (nk - 2 * nm0 + 1)  # one direction's cells
        return units, 216.0 * units, "cells"
    if kind == "update":
        units = (ni - 2 * nm1) * (nj - 2 * nm1) * (nk - 2 * nm1)
        per = 72 + (72 if p.irkt == 2 else (72 + 2 * 144) / 3.0) + 216 + 72
        return units, per * units + 776.0 * P2, "interior cells"
    if kind == "recov":
        units = (ni - 2 * nm1) * (nj - 2 * nm1) * (nk - 2 * nm1)
        return units, (72 + 72 + 72 + 8) * units + 264.0 * P2, "interior cells"
    if kind == "post":
        units = ni * nj * nk
        return units, 216.0 * units + 264.0 * P2, "cells"
    if kind == "fluxct":
        units = (ni - 2 * nm1 + 1) * (nj - 2 * nm1 + 1) * (nk - 2 * nm1 + 1)
        return units, 96.0 * units, "cells"
    if kind == "cfl":
        units = (ni - 2) * (nj - 2) * (nk - 2)
        return units, 72.0 * units + 264.0 * P2, "cells"
    return 0, 0.0, ""


def step_bytes_per_zone_update(p):
    """Compulsory HBM bytes of one zone-update (one interior zone advanced one
    full step): every RK stage reads the stage's primitives W and conserved
    state U (plus U^n from the second stage on) and writes U and W, 72 B each
    (288 B at stage 1, 360 B after: RK2 648 B, RK3 1008 B), plus the
    axisymmetric metric tables, 103 doubles per (r, theta) column per stage
    (3 locations x (10 + 10 + 1) + 40 Christoffel symbols = 824 B), shared by
    the column's interior phi cells."""
    per_stage = [288.0] + [360.0] * (p.irkt - 1)
    nphi = max(p.jmax - 2 * p.nm1, 1)
    return sum(per_stage) + p.irkt * 824.0 / nphi


PROFILE_NOTE = ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes over this bench's command "
                "(profiles/hbm_traffic_<config>.md), not the timed run")


def profile_figures(config):
    """Per-kernel HBM traffic (bytes per launch) and VALU issue figures taken
    from the committed rocprofv3 counter runs of this workload
    (profiles/traffic.json), or {} when there are none for `config`."""
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(tfile):
        return {}
    with open(tfile) as f:
        tr = json.load(f)
    return tr.get(config, {})


def step_roofline(p, zups, s_per_step, prof, dominant):
    """The whole step against HBM: compulsory bytes per zone-update times
    zone-updates/s over the 8 TB/s peak.  Beside it: the HBM bytes the step
    really moves (sum of the per-kernel counter figures of the profile run),
    the step's VALU issue time at one wave instruction per clock per CU (from
    the same profile), and the dominant kernel's own figure."""
    A = step_bytes_per_zone_update(p)
    achieved = zups * A / 1e9
    nint = (p.imax - 2 * p.nm1) * (p.jmax - 2 * p.nm1) * (p.kmax - 2 * p.nm1)
    out = {
        "bound": "hbm",
        "scope": "whole time step (CFL reduction + RK stages): compulsory state bytes per zone-update",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "algorithmic_bytes_per_zone_update": A,
        "zone_updates_per_step": nint,
        "avg_step_ms": s_per_step * 1e3,
        "traffic": prof.get("step_bytes"),
        "traffic_over_algorithmic": (prof["step_bytes"] / (A * nint)) if prof.get("step_bytes") else None,
        "traffic_source": PROFILE_NOTE if prof.get("step_bytes") else None,
        "valu": None,
        "dominant_kernel": dominant,
    }
    v = prof.get("valu")
    if v:
        # the fraction of the step its VALU instructions alone need
        out["valu"] = dict(v, issue_frac=v["issue_floor_ms_per_step"] / (s_per_step * 1e3),
                           fp64_issue_frac=v["fp64_issue_floor_ms_per_step"] / (s_per_step * 1e3))
    return out


def cpu_baseline(p, phi_cells):
    """The oracle (C restatement, one core) on a bounded sample of the same
    workload: the same r-theta grid and torus with `phi_cells` interior phi
    cells; one warm-up step, then one timed step.  Returns (zone-updates/s,
    seconds timed, sample grid)."""
    from dataclasses import replace

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from harness import Oracle, fzeros  # noqa: E402
    from raishinsource_amd import native

    q = replace(p, jmax=phi_cells + 2 * p.nm1, ymax=None).resolved()
    g = native.setup_grid(q)
    m = native.setup_metric(q, g)
    uri = native.setup_torus(q, g, m, seed=0, noise=0.02)
    o = Oracle(q)
    uu = fzeros(9, q.imax, q.jmax, q.kmax)
    o.call("caluu1", uri, uu, m["gcov"], m["gcon"], m["detg"])
    _, t = o.step(uu, uri, g, m, 0.0)
    t0 = time.perf_counter()
    o.step(uu, uri, g, m, t)
    el = time.perf_counter() - t0
    o.lib.orc_free_work()
    nint = (q.imax - 2 * q.nm1) * (q.jmax - 2 * q.nm1) * (q.kmax - 2 * q.nm1)
    return nint / el, el, q


def _ref_sample(name):
    """Set-up of one reference-build sample: the reference's own grid, metric
    and Christoffel arrays, the seeded torus, its conserved variables and a
    stepper.  Returns (stepper, uu, uri, pram)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from harness import Oracle, RefStepper, fzeros, grid_and_metric  # noqa: E402
    from raishinsource_amd import native

    r, g, m = grid_and_metric(name)
    q = r.p
    uri = native.setup_torus(q, g, m, seed=0, noise=0.02)
    uu = fzeros(9, q.imax, q.jmax, q.kmax)
    Oracle(q).call("caluu1", uri, uu, m["gcov"], m["gcon"], m["detg"])
    return RefStepper(name, g, m), uu, uri, q


def _ref_worker(name, nsteps, barrier, out):
    """One rank of the all-core CPU baseline: set up, one warm-up step, wait
    for every rank, then time `nsteps` steps of its own slab."""
    rs, uu, uri, q = _ref_sample(name)
    _, t = rs.step(uu, uri, 0.0)
    barrier.wait()
    t0 = time.perf_counter()
    for _ in range(nsteps):
        _, t = rs.step(uu, uri, t)
    out.put(time.perf_counter() - t0)


def cpu_baseline_reference(name, ranks):
    """The reference's own routines (oracle/_ref/<name>: the Fortran sources
    built with flang -O2) stepped in main.f90:743-1149's order, with the
    grid, metric and Christoffel symbols from its own grid.f90 / coord1 /
    calchrist; bnd4 and infchk come from the oracle (bnd.f90 needs mpif.h).
    `ranks` processes, one per host core, each stepping its own phi slab of
    the same torus as one MPI rank of the reference's decomposition would
    (mpisub.f90:6-59; the halo exchange between slabs is left out): one
    warm-up step, then two timed steps, started together.  Aggregate
    zone-updates/s = ranks x slab zones x 2 / slowest rank's time.  Returns
    (zone-updates/s, seconds, pram), or None where the build is absent."""
    import multiprocessing as mp

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from harness import have_ref  # noqa: E402
    from raishinsource_amd import configs

    if not have_ref(name):
        return None
    q = configs.ALL[name].resolved()
    ctx = mp.get_context("spawn")  # fresh interpreters: this process holds the GPU
    bar, out = ctx.Barrier(ranks), ctx.Queue()
    procs = [ctx.Process(target=_ref_worker, args=(name, 2, bar, out)) for _ in range(ranks)]
    for pr in procs:
        pr.start()
    els = [out.get(timeout=600) for _ in procs]
    for pr in procs:
        pr.join()
    el = max(els)
    nint = (q.imax - 2 * q.nm1) * (q.jmax - 2 * q.nm1) * (q.kmax - 2 * q.nm1)
    return ranks * 2 * nint / el, el, q


def host_cores():
    """Host cores this process may use, at most 16 (one GPU's share of the
    box's CPUs)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c2ppm")
    ap.add_argument("--cpu-phi", type=int, default=64,
                    help="interior phi cells of the CPU-baseline sample (0: skip the CPU baseline)")
    ap.add_argument("--mixed-steps", type=int, default=100,
                    help="steps of the FP64-vs-mixed error comparison (0: skip the mixed-precision run)")
    ap.add_argument("--ensemble", type=int, default=64,
                    help="models of the 2D-torus ensemble run (0: skip)")
    ap.add_argument("--ensemble-steps", type=int, default=500, help="steps per ensemble model (even)")
    ap.add_argument("--no-autotune", action="store_true", help="keep 256-thread workgroups everywhere")
    ap.add_argument("--serve-rates", default="10,100,500",
                    help="Poisson arrival rates (requests/s) of the serving runs, comma-separated ('' skips)")
    ap.add_argument("--serve-requests", type=int, default=40, help="requests per serving run")
    ap.add_argument("--serve-budget-gib", type=float, default=64.0, help="HBM budget of the serving loop")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): every box here has one MI355X and "
              "this path has no multi-GPU form; sharding the subdomain over several GPUs is designed, not built.", file=sys.stderr)
        return 2

    from raishinsource_amd import configs, native

    # stdout carries the one result line.  Native code that prints (the
    # reference's Fortran routines of the CPU baseline write to unit 6, and
    # their runtime may flush at exit) goes to stderr: fd 1 points there for
    # the whole run, and the result line is written to the saved stdout.
    sys.stdout.flush()
    result_fd = os.dup(1)
    os.dup2(2, 1)

    p = configs.ALL[args.config].resolved()
    t_setup = time.perf_counter()
    g = native.setup_grid(p)
    m = native.setup_metric(p, g, row=True)  # the axisymmetric metric's one phi row
    uri = native.setup_torus(p, g, m, seed=0, noise=0.02)
    model = native.Model(p, g, m)
    model.set_state(uri)
    t_setup = time.perf_counter() - t_setup
    print(f"[bench] {args.config} {p.imax}x{p.jmax}x{p.kmax} set up in {t_setup:.1f}s", file=sys.stderr, flush=True)
    block_sizes = model.block_sizes() if args.no_autotune else model.autotune(uri, steps=2)
    print(f"[bench] workgroup sizes {block_sizes}", file=sys.stderr, flush=True)

    # the warm-up steps time every kernel kind (the per-kernel breakdown and the
    # dominant kernel); the timed steps carry events around the dominant
    # kernel's launches only, since events between launches cost time
    kt_all = None
    model.set_timing(args.warmup > 0)
    for _ in range(args.warmup):
        model.step(1, sync=False)
    model.sync()
    if args.warmup > 0:
        kt_all = model.get_timing()
        model.set_timing(False)
    dom = max(kt_all, key=lambda k: kt_all[k][0]) if kt_all else None
    print(f"[bench] warmup {args.warmup} steps done", file=sys.stderr, flush=True)
    model.reset_counters()
    if dom is None:
        model.set_timing(True)
    else:
        model.set_timing(kinds=[dom])
    model.sync()
    t0 = time.perf_counter()
    model.step(args.steps, sync=False)
    model.sync()
    el = time.perf_counter() - t0
    kt = model.get_timing()
    print(f"[bench] timed {args.steps} steps: {el:.3f}s", file=sys.stderr, flush=True)
    model.set_timing(False)
    sim_time, dt = model.step(0)
    counters = model.counters()

    nint = (p.imax - 2 * p.nm1) * (p.jmax - 2 * p.nm1) * (p.kmax - 2 * p.nm1)
    zups = nint * args.steps / el
    # dominant kernel by total time (in the warm-up steps, else the timed ones)
    if dom is None:
        dom = max(kt, key=lambda k: kt[k][0])
    if kt_all is None:
        kt_all = kt
    ms_tot, launches = kt[dom]
    avg_ms = ms_tot / max(launches, 1)
    units, abytes, unit_name = algorithmic_bytes(p, dom)
    achieved = abytes / (avg_ms * 1e-3) / 1e9
    prof = profile_figures(args.config)
    traffic = prof.get("traffic", {}).get(dom)
    kernel_ms = {k: round(v[0] / max(v[1], 1), 4) for k, v in kt_all.items()}  # per launch, warm-up steps

    model.close()
    cpu = None
    if args.cpu_phi > 0:
        print(f"[bench] cpu baseline: oracle on {args.cpu_phi} interior phi cells", file=sys.stderr, flush=True)
        v, el_cpu, q = cpu_baseline(p, args.cpu_phi)
        cpu = {"value": v, "unit": "zone-updates/s", "cores": 1, "kind": "port",
               "sample": f"one step (after one warm-up step) of the same torus on {q.imax}x{q.jmax}x{q.kmax} "
                         f"({args.cpu_phi} of the {p.jmax - 2 * p.nm1} interior phi cells), {el_cpu:.1f}s, "
                         "oracle/raishin_oracle.c single-threaded"}
        ranks = host_cores()
        ref = cpu_baseline_reference("c2cpu", ranks) if p.name == "c2ppm" else None
        if ref is not None:
            rv, rel, rq = ref
            print(f"[bench] cpu baseline: reference build on {ranks} cores {rv:.3g} zone-updates/s ({rel:.1f}s)",
                  file=sys.stderr, flush=True)
            cpu = {"value": rv, "unit": "zone-updates/s", "cores": ranks, "kind": "reference",
                   "sample": f"{ranks} processes, one per host core, each stepping its own "
                             f"{rq.imax}x{rq.jmax}x{rq.kmax} phi slab of the same torus "
                             f"({rq.jmax - 2 * rq.nm1} of the {p.jmax - 2 * p.nm1} interior phi cells) two steps "
                             f"after one warm-up step, started together; slowest rank {rel:.1f}s: the reference's "
                             "Fortran routines (flang -O2, oracle/_ref/c2cpu) in main.f90's step order, bnd4/infchk "
                             "from the oracle (bnd.f90 needs mpif.h), no halo exchange between slabs",
                   "port": cpu}

    line = {
        "metric": "zone-updates/s on 1 GPU, mixed vs FP64 at L1 err<=1e-5; %HBM peak; p99 latency",
        "value": zups,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded Fishbone-Moncrief torus (2% pressure noise, seed 0)",
        "config": {
            "workload": f"3D magnetised Kerr-Schild torus {p.imax}x{p.jmax}x{p.kmax} (interior {nint} zones), "
                        f"{'PPM' if p.irec == 5 else 'MP5' if p.irec == 6 else 'WENO5'}, HLLE, flux-CT, "
                        f"RK{p.irkt}, FP64",
            "name": args.config,
            "interior_zones": nint,
            "parallelism": "single GPU",
        },
        "roofline": step_roofline(p, zups, el / args.steps, prof, {
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": PROFILE_NOTE if traffic is not None else None,
            "algorithmic_bytes_per_launch": abytes,
            "units_per_launch": units,
            "unit_name": unit_name,
            "avg_launch_ms": avg_ms,
        }),
        "kernel_avg_ms": kernel_ms,
        "cpu_baseline": cpu,
        "workgroup_sizes": block_sizes,
        "sim_time": sim_time,
        "counters": counters,
    }
    if args.mixed_steps > 0:
        line["mixed"] = mixed_precision(p, g, m, uri, args, el / args.steps, dom)
    if args.ensemble > 0:
        line["ensemble"] = ensemble(args)
    if args.serve_rates:
        line["serving"] = serve(args)
    sys.stdout.flush()
    os.write(result_fd, (json.dumps(line) + "\n").encode())
    return 0


def ensemble(args):
    """Independent 2D tori of uneven size (r x theta drawn from 64x64,
    128x128, 256x128, 512x256; spin U[0,0.99]; beta_min from 10, 100, 1000;
    r_in U[6,10], r_max = r_in + 6; PPM, RK2, CFL 0.4; seeded).  The same
    steps are run one model at a time (each model's steps launched eagerly,
    then the next model), with one stream and one two-step HIP graph per
    model queued round robin, and as one batch (one launch per kernel for all
    models, two steps per graph launch), in FP64 and mixed precision.  The
    FP64 results of the three must be bit-identical."""
    import numpy as np

    from raishinsource_amd import native, serving

    steps = max(2, args.ensemble_steps - args.ensemble_steps % 2)
    t_set = time.perf_counter()
    models, states, zones = [], [], 0
    for p, init in serving.make_ensemble(args.ensemble, seed=2024):
        mod, uri = serving.build_torus(p, init)
        models.append(mod)
        states.append(uri)
        zones += (p.imax - 2 * p.nm1) * (p.kmax - 2 * p.nm1)
    # the autotuner once per grid shape; the other models of that shape take its choice
    tuned = {}
    if not args.no_autotune:
        for mod, uri in zip(models, states):
            key = (mod.p.imax, mod.p.kmax)
            if key in tuned:
                mod.set_block_sizes(tuned[key])
            else:
                tuned[key] = mod.autotune(uri, steps=2)
    t_set = time.perf_counter() - t_set
    print(f"[bench] ensemble: {len(models)} models set up (and tuned) in {t_set:.1f}s", file=sys.stderr, flush=True)
    for mod, uri in zip(models, states):
        mod.set_state(uri)
        mod.sync()
    t0 = time.perf_counter()
    for mod in models:
        mod.step(steps, sync=True)
    t_seq = time.perf_counter() - t0
    ref = [mod.get_state()[0] for mod in models]
    for mod, uri in zip(models, states):
        mod.set_state(uri)
        mod.step_graph(1)  # capture (and run) two steps, then start again
        mod.set_state(uri)
        mod.sync()
    t0 = time.perf_counter()
    native.step_models_graph(models, steps // 2)
    t_conc = time.perf_counter() - t0
    same_streams = all(np.array_equal(np.asarray(r).view(np.uint64), np.asarray(mod.get_state()[0]).view(np.uint64))
                       for r, mod in zip(ref, models))

    def batched(mixed):
        for mod, uri in zip(models, states):
            mod.set_precision(mixed)
            mod.set_state(uri)
        b = native.Batch(models)
        b.step_graph(1)  # plan and capture (two steps run), then start again
        for mod, uri in zip(models, states):
            mod.set_state(uri)
        b.sync()
        for mod in models:
            mod.sync()
        t0 = time.perf_counter()
        b.step_graph(steps // 2)
        b.sync()
        el = time.perf_counter() - t0
        b.set([])
        b.close()
        return el

    t_bat = batched(False)
    same_batch = all(np.array_equal(np.asarray(r).view(np.uint64), np.asarray(mod.get_state()[0]).view(np.uint64))
                     for r, mod in zip(ref, models))
    t_mix = batched(True)
    l1 = []
    for r, mod in zip(ref, models):
        d = np.asarray(mod.get_state()[0])[0]
        r0 = np.asarray(r)[0]
        l1.append(float(np.abs(d - r0).sum() / np.abs(r0).sum()))
    for mod in models:
        mod.close()
    print(f"[bench] ensemble: {steps} steps, one at a time {t_seq:.2f}s, streams {t_conc:.2f}s, "
          f"batched {t_bat:.2f}s, batched mixed {t_mix:.2f}s", file=sys.stderr, flush=True)
    rate = lambda t: {"value": zones * steps / t, "unit": "zone-updates/s", "seconds": t}
    return {
        "models": len(models),
        "steps_per_model": steps,
        "interior_zones": zones,
        "one_at_a_time": rate(t_seq),
        "streams_graphs": dict(rate(t_conc), bitwise_same=bool(same_streams)),
        "batched_graph": dict(rate(t_bat), bitwise_same=bool(same_batch), speedup_vs_one_at_a_time=t_seq / t_bat),
        "batched_graph_mixed": dict(rate(t_mix), max_rel_l1_density_vs_fp64=max(l1),
                                    speedup_vs_one_at_a_time=t_seq / t_mix),
        "setup_seconds": t_set,
        "workgroup_sizes_by_shape": {f"{k[0] - 10}x{k[1] - 10}": v for k, v in tuned.items()},
    }


def serve(args):
    """Continuous batching of Poisson arrivals (60% 1D SRMHD tubes of 1024
    zones to t=0.4, 30% 2D tori as in the ensemble for 500 steps, 10% 3D 64^3
    tori for 100 steps; FP64 or mixed per request), admitted by device
    footprint, against the same requests served one at a time.  One request
    mix (seed 1) is used at every rate, its arrival times scaled by the rate.
    One at a time: each request built and run alone, in arrival order,
    starting when it has arrived and the previous one has finished (service
    times measured once, queueing computed per rate)."""
    from dataclasses import replace as dc_replace

    import numpy as np

    from raishinsource_amd import serving

    rates = [float(x) for x in args.serve_rates.split(",") if x]
    base = serving.make_requests(args.serve_requests, 1.0, seed=1)
    budget = int(args.serve_budget_gib * (1 << 30))
    # one at a time: service times (build + steps) of each request, alone
    service = {}
    zu_alone = 0
    for req in base:
        r, tb, ts = serving.run_alone(req, keep_state=False, strict=False)
        service[req.rid] = (tb, ts, r.steps)
        zu_alone += req.zones * r.steps
        print(f"[bench] serving: request {req.rid} ({req.kind}) alone: build {tb:.2f}s, {r.steps} steps {ts:.2f}s",
              file=sys.stderr, flush=True)
    print(f"[bench] serving: {len(base)} requests run one at a time: "
          f"{sum(a + b for a, b, _ in service.values()):.1f}s of service", file=sys.stderr, flush=True)
    out = {"requests": len(base), "hbm_budget_gib": args.serve_budget_gib,
           "mix": {k: sum(1 for r in base if r.kind == k) for k in ("tube", "torus2d", "torus3d")},
           "mixed_precision_requests": sum(1 for r in base if r.mixed), "rates": {}}
    for lam in rates:
        reqs = [dc_replace(r, arrival=r.arrival / lam) for r in base]
        srv = serving.Server(hbm_budget=budget)
        t0 = time.perf_counter()
        res = srv.run(reqs)
        wall = time.perf_counter() - t0
        srv.close()
        summ = serving.summarize(res, wall)
        summ["ticks"] = srv.ticks
        summ["deferred_by_budget"] = len(srv.deferred)
        same_steps = all(r.steps == service[r.rid][2] for r in res)
        # the single worker: FIFO queue over the measured service times
        free_at, lat = 0.0, []
        for r in sorted(reqs, key=lambda x: x.arrival):
            tb, ts, _ = service[r.rid]
            start = max(r.arrival, free_at)
            free_at = start + tb + ts
            lat.append(free_at - r.arrival)
        wall1 = free_at - min(r.arrival for r in reqs)
        one = {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
               "zone_updates_per_s": zu_alone / wall1, "wall_s": wall1}
        out["rates"][str(int(lam) if lam == int(lam) else lam)] = {
            "batched": summ, "one_at_a_time": one, "same_step_counts": bool(same_steps)}
        print(f"[bench] serving at {lam:g}/s: batched p50 {summ['p50_s']:.2f}s p99 {summ['p99_s']:.2f}s "
              f"{summ['zone_updates_per_s']:.3g} zu/s; one at a time p50 {one['p50_s']:.2f}s "
              f"p99 {one['p99_s']:.2f}s {one['zone_updates_per_s']:.3g} zu/s", file=sys.stderr, flush=True)
    return out


def divb_max(uu, g, p):
    """Largest |div B| over the corners of the updated cells: the divergence
    of the densitised field (uu(7..9) = sqrt(-g) B^i) at each corner from
    the four cell-centred values on each side, divided by the distance
    between the cell centres.  The reference does not keep it at round-off
    (its HLLE flux of the normal field component is nonzero), so it is
    reported for FP64 and mixed side by side."""
    import numpy as np

    n = p.nm1
    core = (slice(n, p.imax - n), slice(n, p.jmax - n), slice(n, p.kmax - n))
    Bs = [np.asarray(uu[6 + a])[core] for a in range(3)]
    xs = [np.asarray(g["x1"])[n:p.imax - n], np.asarray(g["x2"])[n:p.jmax - n], np.asarray(g["x3"])[n:p.kmax - n]]
    if min(b.shape[a] for a, b in enumerate(Bs)) < 2:
        return None
    div = 0.0
    for a in range(3):
        B = Bs[a]
        d = np.diff(B, axis=a) / np.diff(xs[a]).reshape([-1 if q == a else 1 for q in range(3)])
        # average over the 2x2 cells around the corner in the other two axes
        for q in range(3):
            if q != a:
                sl0 = [slice(None)] * 3
                sl1 = [slice(None)] * 3
                sl0[q] = slice(0, -1)
                sl1[q] = slice(1, None)
                d = 0.5 * (d[tuple(sl0)] + d[tuple(sl1)])
        div = div + d
    return float(np.abs(div).max())


def mixed_precision(p, g, m, uri, args, s_per_step_fp64, dom):
    """The mixed form (reconstruction and face fluxes in FP32): its rate on
    the same timed steps, and its per-field errors against the FP64 run after
    args.mixed_steps steps from the same initial state."""
    import numpy as np

    from raishinsource_amd import native

    model = native.Model(p, g, m)
    model.set_state(uri)
    model.set_precision(True)
    sizes = model.block_sizes() if args.no_autotune else model.autotune(uri, steps=2)
    # as the FP64 run: the warm-up steps time every kernel kind, the timed
    # steps carry events around the launches of the FP64 run's dominant
    # kernel kind only, so both rates are taken under the same events
    model.set_timing(True)
    for _ in range(args.warmup):
        model.step(1, sync=False)
    model.sync()
    kt = model.get_timing()
    model.set_timing(False)
    if dom is not None:
        model.set_timing(kinds=[dom])
    model.sync()
    t0 = time.perf_counter()
    model.step(args.steps, sync=False)
    model.sync()
    el = time.perf_counter() - t0
    model.set_timing(False)
    model.close()
    nint = (p.imax - 2 * p.nm1) * (p.jmax - 2 * p.nm1) * (p.kmax - 2 * p.nm1)
    print(f"[bench] mixed: timed {args.steps} steps: {el:.3f}s", file=sys.stderr, flush=True)
    runs = {}
    for mixed in (False, True):
        mod = native.Model(p, g, m)
        mod.set_state(uri)
        mod.set_precision(mixed)
        mod.step(args.mixed_steps, sync=False)
        w, u, t = mod.get_state()
        runs[mixed] = (w, mod.counters(), t, u)
        mod.close()
        print(f"[bench] {'mixed' if mixed else 'fp64'} {args.mixed_steps} steps done", file=sys.stderr, flush=True)
    w64, c64, t64, u64 = runs[False]
    w32, c32, t32, u32 = runs[True]
    uu0 = native.Model(p, g, m)
    uu0.set_state(uri)
    u0 = uu0.get_state()[1]
    uu0.close()
    divb = {"initial": divb_max(u0, g, p), "fp64": divb_max(u64, g, p), "mixed": divb_max(u32, g, p)}
    n1 = p.nm1
    inner = (slice(None), slice(n1, p.imax - n1), slice(n1, p.jmax - n1), slice(n1, p.kmax - n1))
    a, b = w32[inner], w64[inner]
    names = ["rho", "u1", "u2", "u3", "p", "s", "B1", "B2", "B3"]
    l1 = {nm: float(np.abs(a[v] - b[v]).sum() / max(np.abs(b[v]).sum(), 1e-300)) for v, nm in enumerate(names)}
    linf = {nm: float(np.abs(a[v] - b[v]).max() / max(np.abs(b[v]).max(), 1e-300)) for v, nm in enumerate(names)}
    return {
        "value": nint * args.steps / el,
        "unit": "zone-updates/s",
        "ms_per_step": el / args.steps * 1e3,
        "speedup_vs_fp64": (s_per_step_fp64 * args.steps) / el,
        "dtype": "f32 reconstruction + face fluxes, f64 elsewhere",
        "kernel_avg_ms": {k: round(v[0] / max(v[1], 1), 4) for k, v in kt.items()},
        "workgroup_sizes": sizes,
        "error_steps": args.mixed_steps,
        "rel_l1_vs_fp64": l1,
        "rel_linf_vs_fp64": linf,
        "sim_time": {"fp64": t64, "mixed": t32},
        "recov_fallback": {"fp64": c64["recov_fallback"], "mixed": c32["recov_fallback"]},
        "recov_fail": {"fp64": c64["recov_fail"], "mixed": c32["recov_fail"]},
        "max_abs_div_b": divb,
    }


if __name__ == "__main__":
    sys.exit(main())
