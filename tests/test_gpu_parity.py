"""The HIP library (through its C-ABI) against the oracle and against the
reference's own golden trajectories, on an MI355X.

Tolerance: per field, max |gpu - ref| / max |ref| <= 1e-11 over the finite
values, with identical NaN positions.  Everything except the entropy
variable (var 6) is bit-identical as long as no inversion falls back to the
entropy equation: var 6 goes through pow(), whose last bit differs between
the device math library and glibc in some calls, and the fallback branch
feeds that bit back into the other fields.
"""
import ctypes as C
import hashlib
import os
from dataclasses import replace

import numpy as np
import pytest

from harness import Oracle, fzeros, synthetic_state
from raishinsource_amd import configs, native

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-11


def bitdiff(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    both_nan = np.isnan(a) & np.isnan(b)
    return int(np.count_nonzero((a.view(np.uint64) != b.view(np.uint64)) & ~both_nan))


def field_errors(a, b):
    """Per-variable max relative error (max-norm scaled by the field's max)."""
    out = []
    for v in range(a.shape[0]):
        x, y = a[v], b[v]
        assert np.array_equal(np.isnan(x), np.isnan(y)), f"var {v + 1}: NaN positions differ"
        fin = np.isfinite(y)
        if not fin.any():
            out.append(0.0)
            continue
        scale = np.abs(y[fin]).max()
        d = np.abs(x[fin] - y[fin]).max()
        out.append(0.0 if d == 0 else d / (scale if scale > 0 else 1.0))
    return out


def finite_fraction(p, uri):
    """Fraction of interior cells (nm1 ghosts per side excluded) whose nine
    primitives are all finite."""
    n1 = p.nm1
    w = uri[:, n1:p.imax - n1, n1:p.jmax - n1, n1:p.kmax - n1]
    return float(np.isfinite(w).all(axis=0).mean())


def assert_finite(tag, p, uri, least):
    """A comparison only pins what is finite: fail when too little of the
    compared state is (a NaN compared with a NaN says nothing)."""
    f = finite_fraction(p, uri)
    print(f"{tag}: finite interior fraction {f:.4f} (required {least})")
    assert f >= least, (tag, f, least)


def assert_close(tag, a, b, tol=TOL):
    errs = field_errors(a, b)
    print(f"{tag}: bitdiff {bitdiff(a, b)} / {a.size}; per-var rel err " + " ".join(f"{e:.1e}" for e in errs))
    assert max(errs) <= tol, (tag, errs)


# the unit grids with enough theta cells under the mirror axis for the
# device ghost fill (nm1 interior cells; the unit grids themselves have fewer)
SMALL = {
    "gppm": replace(configs.G_PPM, kmax=16),
    "gmp5": replace(configs.G_MP5, kmax=13),
    "gweno": replace(configs.G_WENO, kmax=13),
    # recov2 on the MP5 torus: many non-converged cells and negative
    # pressures, so the failure path and the skipped Newton cycles run
    "t3drc2": replace(configs.T3D, name="t3drc2", iwvec=2),
    # MP5 and WENO5 over many steps: smooth 3D states with a floor-level band
    "smp5": replace(configs.G_MP5, name="smp5", imax=24, jmax=14, kmax=20),
    "sweno": replace(configs.G_WENO, name="sweno", imax=24, jmax=14, kmax=20),
}


def make(name_or_p, state="torus", seed=0, noise=0.02):
    if isinstance(name_or_p, str):
        name_or_p = SMALL.get(name_or_p, configs.ALL.get(name_or_p))
    p = name_or_p.resolved()
    g = native.setup_grid(p)
    m = native.setup_metric(p, g)
    if state == "torus":
        uri = native.setup_torus(p, g, m, seed=seed, noise=noise)
    else:
        uri = synthetic_state(p, seed=seed)
    return p, g, m, uri


def oracle_run(p, g, m, uri, nsteps):
    o = Oracle(p)
    o.reset_counters()
    uu = fzeros(9, p.imax, p.jmax, p.kmax)
    o.call("caluu1", uri, uu, m["gcov"], m["gcon"], m["detg"])
    w = uri.copy(order="F")
    t = 0.0
    dts = []
    for _ in range(nsteps):
        dt, t = o.step(uu, w, g, m, t)
        dts.append(dt)
    cnt = o.counters()
    o.lib.orc_free_work()
    return w, uu, t, np.array(dts), cnt


def gpu_run(p, g, m, uri, nsteps, uu=None):
    mod = native.Model(p, g, m)
    mod.set_state(uri, uu)
    dts = []
    t = 0.0
    for _ in range(nsteps):
        t, dt = mod.step(1)
        dts.append(dt)
    w, u, t2 = mod.get_state()
    assert t2 == t
    cnt = mod.counters()
    mod.close()
    return w, u, t, np.array(dts), cnt


def assert_counters(gc, oc):
    for k, v in oc.items():
        assert gc[k] == v, (k, gc[k], v)


# ---------------------------------------------------------------- single ops


@pytest.mark.parametrize("name,state", [("t2d", "torus"), ("t3d", "torus"), ("gmp5", "synthetic")])
def test_caluu1(name, state, oracle_lib):
    """caluu1 (caluu.f90:2) on the device: bitwise except the entropy variable."""
    p, g, m, uri = make(name, state)
    o = Oracle(p)
    uu = fzeros(9, p.imax, p.jmax, p.kmax)
    o.call("caluu1", uri, uu, m["gcov"], m["gcon"], m["detg"])
    mod = native.Model(p, g, m)
    mod.set_state(uri)          # no uu: computed on the device
    _, u2, _ = mod.get_state()
    mod.close()
    assert bitdiff(np.delete(u2, 5, 0), np.delete(uu, 5, 0)) == 0
    assert_close("caluu1", u2, uu, 1e-15)


@pytest.mark.parametrize("name,state", [("t2d", "torus"), ("t3d", "torus"), ("gppm", "synthetic"),
                                        ("gweno", "synthetic")])
def test_cdtcfl4_bitwise(name, state, oracle_lib):
    """cdtcfl4 (cdtcfl.f90:2): the four time steps are bit-identical."""
    p, g, m, uri = make(name, state)
    o = Oracle(p)
    d = [C.c_double(0) for _ in range(4)]
    o.call("cdtcfl4", uri, m["gcov"], m["gcon"], C.byref(d[0]), g["dx1b"], g["dx2b"], g["dx3b"], C.byref(d[1]),
           C.byref(d[2]), C.byref(d[3]), p.nm1)
    want = np.array([x.value for x in d])
    mod = native.Model(p, g, m)
    mod.set_state(uri)
    got = np.array(mod.cdtcfl4())
    mod.close()
    assert bitdiff(got, want) == 0, (got, want)
    # the reference-named entry point (cdtcfl.f90:2 argument list)
    L = native.lib()
    ps = native.pram_struct(p)
    vp = lambda a: C.c_void_p(a.ctypes.data)  # noqa: E731
    L.raishin_init(C.byref(ps))
    L.raishin_geometry(*[vp(a) for a in (g["x1"], g["x2"], g["x3"], g["dx1b"], g["dx2b"], g["dx3b"], m["gcov"],
                                         m["gcon"], m["gcovi"], m["gconi"], m["gcovj"], m["gconj"], m["gcovk"],
                                         m["gconk"], m["detg"], m["detgi"], m["detgj"], m["detgk"], m["christ"])])
    native.check()
    out = [C.c_double(0) for _ in range(4)]
    ints = [C.c_int(v) for v in (p.nm1, 1, p.imax, 1, p.jmax, 1, p.kmax)]
    try:
        L.raishin_cdtcfl4(vp(uri), vp(m["gcov"]), vp(m["gcon"]), C.byref(out[0]), vp(g["dx1b"]), vp(g["dx2b"]),
                          vp(g["dx3b"]), C.byref(out[1]), C.byref(out[2]), C.byref(out[3]), *[C.byref(x) for x in ints])
        native.check()
    finally:
        L.raishin_finalize()
    assert bitdiff(np.array([x.value for x in out]), want) == 0


@pytest.mark.parametrize("name,state", [("gppm", "synthetic"), ("gmp5", "synthetic"), ("gweno", "synthetic"),
                                        ("uppm", "torus"), ("ump5", "torus")])
def test_stage1_intermediates(name, state, oracle_lib):
    """First RK stage: fluxes after rec2/calcha/hll/ct (rec.f90:2, calcha.f90:2,
    hll.f90:2, ct.f90:2) and the primitives recov1a returns (recov.f90:578)."""
    p, g, m, uri = make(name, state)
    ni, nj, nk, nm0, nm1 = p.imax, p.jmax, p.kmax, p.nm0, p.nm1
    o = Oracle(p)
    uu = fzeros(9, ni, nj, nk)
    o.call("caluu1", uri, uu, m["gcov"], m["gcon"], m["detg"])
    names = ["uriir", "urijr", "urikr", "uriil", "urijl", "urikl", "uuir", "uujr", "uukr", "uuil", "uujl", "uukl",
             "wwir", "wwjr", "wwkr", "wwil", "wwjl", "wwkl"]
    F = {n: fzeros(9, ni, nj, nk) for n in names}
    d = [C.c_double(0) for _ in range(4)]
    o.call("cdtcfl4", uri, m["gcov"], m["gcon"], C.byref(d[0]), g["dx1b"], g["dx2b"], g["dx3b"], C.byref(d[1]),
           C.byref(d[2]), C.byref(d[3]), nm1)
    dt = p.cfl * d[0].value
    a1, a2, a3 = dt / g["dx1b"], dt / g["dx2b"], dt / g["dx3b"]
    o.call("rec2", uri, g["x1"], g["x2"], g["x3"], g["dx1b"], g["dx2b"], g["dx3b"], *[F[n] for n in names],
           m["gcovi"], m["gcovj"], m["gcovk"], m["gconi"], m["gconj"], m["gconk"], m["detgi"], m["detgj"],
           m["detgk"], nm0)
    cR = [fzeros(ni, nj, nk) for _ in range(6)]
    o.call("calcha", *[F[n] for n in names[:6]], m["gcovi"], m["gcovj"], m["gcovk"], m["gconi"], m["gconj"],
           m["gconk"], *cR, nm0)
    ww = fzeros(3, 9, ni, nj, nk)
    o.call("hll", ww, *[F[n] for n in names[6:]], *cR, nm0)
    o.call("ct", ww, nm0, nm1)
    tenr = fzeros(4, 4, ni, nj, nk)
    sf = fzeros(4, ni, nj, nk)
    o.call("caltenr", uri, tenr, m["gcov"], m["gcon"])
    o.call("calsf", sf, tenr, m["christ"], m["detg"])
    uh = fzeros(9, ni, nj, nk)
    o.call("rk2fst", uh, ww, uu, nm1, a1, a2, a3)   # the first RK3 stage is the same Euler step
    o.call("rkadsf", uh, sf, dt, nm1, 0)
    w2 = uri.copy(order="F")
    o.call("recov1a", uh, w2, m["gcov"], m["gcon"], m["detg"], g["x1"], nm1)

    mod = native.Model(p, g, m)
    mod.set_state(uri, uu)
    wwg, wpg = mod.stage(1, True, True)
    mod.close()
    used = (slice(None), slice(None), slice(nm1 - 1, ni - nm1), slice(nm1 - 1, nj - nm1), slice(nm1 - 1, nk - nm1))
    a, b = wwg[used], ww[used]
    assert bitdiff(np.delete(a, 5, 1), np.delete(b, 5, 1)) == 0
    assert_close("fluxes", a.reshape(27, -1), b.reshape(27, -1), 1e-14)
    inner = (slice(None), slice(nm1, ni - nm1), slice(nm1, nj - nm1), slice(nm1, nk - nm1))
    assert_close("recovered primitives", wpg[inner], w2[inner], 1e-12)


# ------------------------------------------------------- whole time steps


def test_golden_t2d_reference(oracle_lib):
    """10 steps of the 2D torus (64 x 64, PPM, RK2) against the reference
    Fortran's own trajectory (tests/golden/traj_t2d.npz)."""
    p = configs.T2D.resolved()
    gold = dict(np.load(os.path.join(GOLD, "traj_t2d.npz")))
    g = native.setup_grid(p)
    m = native.setup_metric(p, g)
    uri = np.asfortranarray(np.repeat(gold["uri0"][:, :, None, :], p.jmax, axis=2))
    w, u, t, dts, _ = gpu_run(p, g, m, uri, len(gold["dt"]))
    j = p.nm1
    assert bitdiff(dts, gold["dt"]) == 0
    assert_close("t2d uri", w[:, :, j, :], gold["uri"])
    assert_close("t2d uu", u[:, :, j, :], gold["uu"])


def test_golden_c1_100_steps_reference(oracle_lib):
    """The reference's 2D torus (128 x 128, PPM, HLLE, flux-CT, RK2, CFL 0.4,
    2% pressure noise, seed 0): 100 steps against the reference Fortran's
    trajectory, per field within 1e-11."""
    p = configs.C1.resolved()
    gold = dict(np.load(os.path.join(GOLD, "traj_c1.npz")))
    p_, g, m, uri = make("c1")
    w, u, t, dts, cnt = gpu_run(p, g, m, uri, 100)
    j = p.nm1
    print("dt bitdiff", bitdiff(dts, gold["dt"]), "time", t, float(gold["time"]), "counters", cnt)
    assert np.abs(dts / gold["dt"] - 1).max() <= TOL
    assert abs(t / float(gold["time"]) - 1) <= TOL
    assert_close("c1 uri after 100 steps", w[:, :, j, :], gold["uri100"])


# (name, initial state, steps, least finite fraction of the interior at the
# compared step).  MP5 and WENO5 turn the small tori to NaN within two steps
# in the reference too (MP5's absolute 1e-10 bound lets face densities go
# negative and de**gam is then NaN), so those tori are compared after one step
# (16%, 7.5% and 89% of the interior still finite) and after two for t3d
# (21%); the multi-step MP5 and WENO5 evidence is the synthetic 3D states,
# which stay finite.
ORACLE_CASES = [
    ("uppm", "torus", 12, 1.0),
    ("t3dppm", "torus", 20, 1.0),
    ("ump5", "torus", 1, 0.15),
    ("uweno", "torus", 1, 0.07),
    ("t3d", "torus", 1, 0.85),
    ("t3d", "torus", 2, 0.2),
    ("umc", "torus", 8, 1.0),
    ("t2dmc", "torus", 6, 1.0),
    ("uleis", "torus", 6, 1.0),
    ("t2dgam", "torus", 6, 1.0),
    ("t2drc2", "torus", 8, 1.0),
    ("urc2", "torus", 6, 1.0),
    ("t3drc2", "torus", 4, 1.0),
    ("gppm", "synthetic", 4, 1.0),
    ("gmp5", "synthetic", 4, 1.0),
    ("gweno", "synthetic", 4, 1.0),
    ("smp5", "synthetic", 12, 1.0),
    ("sweno", "synthetic", 12, 1.0),
]


@pytest.mark.parametrize("name,state,nsteps,least", ORACLE_CASES,
                         ids=[f"{c[0]}-{c[2]}" for c in ORACLE_CASES])
def test_steps_match_oracle(name, state, nsteps, least, oracle_lib):
    """Whole steps (main.f90:743-1149 order) against the oracle, with the
    inversion-failure and floor counts."""
    p, g, m, uri = make(name, state)
    wo, uo, to, dto, co = oracle_run(p, g, m, uri, nsteps)
    wg, ug, tg, dtg, cg = gpu_run(p, g, m, uri, nsteps)
    print(name, "oracle counters", co, "\n", name, "gpu counters", cg)
    assert_finite(name, p, wo, least)
    assert np.array_equal(np.isnan(dtg), np.isnan(dto))
    assert_close(name + " dt", dtg[None, np.isfinite(dto)], dto[None, np.isfinite(dto)])
    assert_close(name + " uri", wg, wo)
    assert_close(name + " uu", ug, uo)
    assert_counters(cg, co)


# Tubes with strong fields send many cells through recov1a's entropy branch
# (recov.f90:1022-1147), which is built on pow (rho**gam and its
# derivatives); the device pow (ROCm ocml) and the host's glibc pow differ in
# the last bit for some arguments, and those differences grow over the steps.
# Same counters, same step sequence; per-field agreement to this tolerance.
TUBE_TOL = 1e-8


@pytest.mark.parametrize("name,seed,nsteps", [("utube", 1, 40), ("utube", 2, 40), ("tube", 3, 12)])
def test_tube_matches_oracle(name, seed, nsteps, oracle_lib):
    """1D SRMHD Riemann tubes (flat metric, Gamma 5/3, random left/right states)."""
    p = configs.ALL[name].resolved()
    g = native.setup_grid(p)
    m = native.setup_metric(p, g)
    left, right = native.random_tube_states(np.random.default_rng(seed))
    uri = native.setup_tube(p, left, right)
    wo, uo, to, dto, co = oracle_run(p, g, m, uri, nsteps)
    wg, ug, tg, dtg, cg = gpu_run(p, g, m, uri, nsteps)
    print(name, seed, "oracle counters", co, "gpu counters", cg)
    assert_close(name + " dt", dtg[None], dto[None], TUBE_TOL)
    assert_close(name + " uri", wg, wo, TUBE_TOL)
    assert_close(name + " uu", ug, uo, TUBE_TOL)
    assert_counters(cg, co)


def test_tube_failing_inversions_match_oracle(oracle_lib):
    """A strongly magnetised tube whose recovery often fails the 2D Newton
    iteration (all iter trips in the reference) and sometimes the entropy
    iteration too: the device skips whole periods of cycling iterations and
    must end in the reference's states and counts."""
    from raishinsource_amd import serving
    req = serving.make_requests(40, 1.0, seed=1)[33]
    assert req.kind == "tube"
    p = configs.TUBE.resolved()
    g = native.setup_grid(p)
    m = native.setup_metric(p, g)
    uri = native.setup_tube(p, req.init["left"], req.init["right"])
    wo, uo, to, dto, co = oracle_run(p, g, m, uri, 20)
    wg, ug, tg, dtg, cg = gpu_run(p, g, m, uri, 20)
    print("oracle counters", co, "gpu counters", cg)
    assert co["recov_fallback"] > 0 and co["recov_fail"] > 0
    assert_close("tube dt", dtg[None], dto[None], TUBE_TOL)
    assert_close("tube uri", wg, wo, TUBE_TOL)
    assert_close("tube uu", ug, uo, TUBE_TOL)
    assert_counters(cg, co)


def test_carried_guesses_match_oracle(oracle_lib):
    """Cells whose recov1a quadratic has no admissible root keep the previous
    cell's guess (recov.f90:815-817), so each depends on its predecessor in
    (k,j,i) order.  A 64 x 14 x 64 interior PPM/RK3 torus whose conserved
    field is scaled by 1000 in a band of 16 theta planes gives thousands of
    such cells per step in hundreds of chains (consecutive cells), all
    finite, spread over the 56 workgroups that share the chain heads."""
    p = replace(configs.ALL["t3dppm"], name="carry", imax=74, jmax=24, kmax=74).resolved()
    g = native.setup_grid(p)
    m = native.setup_metric(p, g)
    uri = native.setup_torus(p, g, m, seed=0, noise=0.02)
    o = Oracle(p)
    o.reset_counters()
    uu = fzeros(9, p.imax, p.jmax, p.kmax)
    o.call("caluu1", uri, uu, m["gcov"], m["gcon"], m["detg"])
    band = (slice(p.nm1 + 5, p.imax - p.nm1 - 5), slice(None), slice(p.kmax // 2 - 8, p.kmax // 2 + 8))
    for b in (6, 7, 8):
        uu[b][band] *= 1000.0
    # the oracle's step, recording the carried cells in loop order
    rec = np.zeros(1 << 20, dtype=np.int64)
    o.lib.orc_debug_carry(rec.ctypes.data_as(C.c_void_p), C.c_longlong(len(rec)))
    wo, uo = uri.copy(order="F"), uu.copy(order="F")
    dto, _ = o.step(uo, wo, g, m, 0.0)
    n = o.lib.orc_debug_carry_count()
    o.lib.orc_debug_carry(None, C.c_longlong(0))
    co = o.counters()
    o.lib.orc_free_work()
    chains = np.split(rec[:n], np.flatnonzero(np.diff(rec[:n]) != 1) + 1)
    lens = np.array([len(c) for c in chains])
    print("carried cells", n, "chains", len(lens), "longest", lens.max(), "counters", co)
    assert n >= 3000 and len(lens) >= 300 and lens.max() >= 8 and co["guard"] == 0
    wg, ug, tg, dtg, cg = gpu_run(p, g, m, uri, 1, uu=uu)
    assert finite_fraction(p, wo) == 1.0
    assert_close("carry dt", dtg[None], np.array([dto])[None])
    assert_close("carry uri", wg, wo)
    assert_close("carry uu", ug, uo)
    assert_counters(cg, co)


EDGE = [
    # ragged extents that are multiples of nothing the kernels tile by
    ("ragged_mp5", replace(configs.G_MP5, imax=17, jmax=10, kmax=13), 3),
    ("ragged_ppm_rk3", replace(configs.G_PPM, imax=19, jmax=12, kmax=16, irkt=3, cfl=0.3), 3),
    # the smallest grids the boundary codes allow: nm1 interior cells under the mirror axis
    ("min_ppm", replace(configs.G_PPM, imax=11, jmax=11, kmax=15), 3),
    ("min_weno", replace(configs.G_WENO, imax=9, jmax=9, kmax=12), 3),
]


@pytest.mark.parametrize("tag,p,nsteps", EDGE, ids=[e[0] for e in EDGE])
def test_edge_grids_match_oracle(tag, p, nsteps, oracle_lib):
    p, g, m, uri = make(p, "synthetic", seed=11)
    wo, uo, _, dto, co = oracle_run(p, g, m, uri, nsteps)
    wg, ug, _, dtg, cg = gpu_run(p, g, m, uri, nsteps)
    assert_close(tag + " dt", dtg[None], dto[None])
    assert_close(tag + " uri", wg, wo)
    assert_close(tag + " uu", ug, uo)
    assert_counters(cg, co)


def test_rejects_unbuilt_configuration():
    p, g, m, _ = make("gmp5", "synthetic")
    with pytest.raises(native.RaishinError, match="irec=2"):
        native.Model(replace(p, irec=2), g, m)
    with pytest.raises(native.RaishinError, match="mirror"):
        q = replace(configs.G_PPM, kmax=14).resolved()
        gq = native.setup_grid(q)
        native.Model(q, gq, native.setup_metric(q, gq))


# ----------------------------------------- the full 256 x 128 x 128 subdomain


def slab_of(p_full, g_full):
    """The same r-theta grid with one interior phi cell and the full grid's
    (uniform) phi spacing: an axisymmetric state evolves there exactly as
    every phi slice of the full grid does."""
    q = replace(p_full, jmax=2 * p_full.nm1 + 1, ymax=None).resolved()
    gq = native.setup_grid(q)
    gq["dx2b"][:] = g_full["dx2b"][0]
    return q, gq


@pytest.mark.parametrize("name,nsteps", [("c2ppm", 3)])
def test_full_subdomain_axisymmetric(name, nsteps, oracle_lib):
    """256 x 128 x 128, PPM, RK3: an axisymmetric torus stays bitwise
    axisymmetric on the device, every phi slice equals the device run on the
    one-slice grid bit for bit, and that run matches the oracle.  (With MP5 or
    WENO5 the reference itself breaks the symmetry: recov1a hands a cell with
    no admissible root the previous cell's W in (k,j,i) order, and the first
    phi row inherits from the last row of the previous theta plane.)"""
    p, g, m, uri = make(name, noise=0.0)
    # grid.f90 spaces phi by accumulation, so the widths differ in the last
    # bit; PPM weights by width, so exact translation symmetry needs one width
    g["dx2b"][:] = g["dx2b"][0]
    wf, uf, tf, dtf, _ = gpu_run(p, g, m, uri, nsteps)
    for j in range(p.jmax):
        assert bitdiff(wf[:, :, j, :], wf[:, :, p.nm1, :]) == 0, j
    q, gq = slab_of(p, g)
    mq = native.setup_metric(q, gq)
    uq = np.asfortranarray(uri[:, :, :q.jmax, :])
    ws, us, ts, dts, cs = gpu_run(q, gq, mq, uq, nsteps)
    assert bitdiff(dts, dtf) == 0
    assert bitdiff(ws[:, :, q.nm1, :], wf[:, :, p.nm1, :]) == 0
    assert bitdiff(us[:, :, q.nm1, :], uf[:, :, p.nm1, :]) == 0
    wo, uo, to, dto, co = oracle_run(q, gq, mq, uq, nsteps)
    print(name, "NaN cells", int(np.isnan(wo).any(axis=0).sum()), "counters oracle", co, "gpu", cs)
    assert_close(name + " slab uri", ws, wo)
    assert_close(name + " slab uu", us, uo)
    assert_counters(cs, co)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("name,nsteps,least", [("c2", 2, 0.98), ("c2weno", 1, 0.98)])
def test_full_subdomain_matches_oracle(name, nsteps, least, oracle_lib):
    """The 256 x 128 x 128 torus with 2% pressure noise (MP5 / WENO5, RK3)
    against the oracle at full size.  The reference turns this torus to NaN
    within a few steps (MP5's absolute 1e-10 limiter threshold lets face
    densities go negative); at the compared step 98.8% (MP5, 2 steps) and
    98.6% (WENO5, 1 step) of the interior is still finite.  NaN positions and
    all counters must agree."""
    p, g, m, uri = make(name)
    wg, ug, tg, dtg, cg = gpu_run(p, g, m, uri, nsteps)
    wo, uo, to, dto, co = oracle_run(p, g, m, uri, nsteps)
    print(name, "NaN cells", int(np.isnan(wo).any(axis=0).sum()), "\n oracle", co, "\n gpu", cg)
    assert_finite(name, p, wo, least)
    assert np.array_equal(np.isnan(dtg), np.isnan(dto))
    assert_close(name + " dt", dtg[None], dto[None])
    assert_close(name + " uri", wg, wo)
    assert_close(name + " uu", ug, uo)
    assert_counters(cg, co)


def test_full_subdomain_deterministic():
    """Two runs of the noisy 256 x 128 x 128 PPM torus are bit-identical."""
    p, g, m, uri = make("c2ppm")
    a = gpu_run(p, g, m, uri, 2)
    b = gpu_run(p, g, m, uri, 2)
    assert bitdiff(a[0], b[0]) == 0 and bitdiff(a[1], b[1]) == 0
    assert a[4] == b[4]


# ------------------------------------- 100 steps of the benched 3D workload


def test_c2ppm_100_steps_match_reference():
    """The benched workload itself -- 256 x 128 x 128 torus, 2% pressure
    noise from seed 0, PPM, HLLE, flux-CT, RK3 -- stepped 100 times on the
    device against the reference's Fortran routines stepped 100 times in
    main.f90's order (tests/golden/traj_c2ppm.npz, written by
    make_golden.py c2ppm with the oracle run beside it, bit for bit): the
    initial state, the dt of every step, two whole phi rows of the
    primitives and one of the conserved variables within 1e-11 per field,
    per-field L1 / L-infinity / signed sums and per-phi-row L1 norms of the
    interior within 1e-11 relative, the finite fraction, and the inversion
    and floor counts."""
    path = os.path.join(GOLD, "traj_c2ppm.npz")
    if not os.path.exists(path):
        pytest.skip("traj_c2ppm.npz not written yet (make_golden.py c2ppm)")
    gold = np.load(path)
    p = configs.C2_PPM.resolved()
    g = native.setup_grid(p)
    m = native.setup_metric(p, g, row=True)
    uri = native.setup_torus(p, g, m, seed=0, noise=0.02)
    assert hashlib.sha256(np.asfortranarray(uri).tobytes(order="F")).hexdigest() == str(gold["uri0_digest"])
    mod = native.Model(p, g, m)
    mod.set_state(uri)
    dts = []
    for _ in range(100):
        _, dt = mod.step(1)
        dts.append(dt)
    w, u, t = mod.get_state()
    cnt = mod.counters()
    mod.close()
    dts = np.array(dts)
    assert_close("c2ppm dt", dts[None], gold["dt"][None])
    assert abs(t - float(gold["time"])) <= TOL * abs(float(gold["time"]))
    n1 = p.nm1
    inner = w[:, n1:p.imax - n1, n1:p.jmax - n1, n1:p.kmax - n1]
    fin = float(np.isfinite(inner).all(axis=0).mean())
    assert fin == float(gold["finite_frac"]) and fin == 1.0
    for nm, got in (("l1", np.abs(inner).sum(axis=(1, 2, 3))), ("linf", np.abs(inner).max(axis=(1, 2, 3))),
                    ("sum", inner.sum(axis=(1, 2, 3)))):
        ref = gold[nm]
        rel = np.abs(got - ref) / np.maximum(np.abs(gold["l1"]) if nm == "sum" else np.abs(ref), 1e-300)
        print(f"c2ppm {nm}: max rel {rel.max():.2e}")
        assert rel.max() <= TOL, (nm, rel)
    rows = np.abs(inner).sum(axis=(1, 3))
    # per field, relative to its largest row (a field that is zero everywhere, such
    # as the entropy-like variable 6 here, must stay exactly zero)
    rel = np.abs(rows - gold["l1_rows"]) / np.maximum(np.abs(gold["l1_rows"]).max(axis=1, keepdims=True), 1e-300)
    print(f"c2ppm phi-row L1: max rel {rel.max():.2e}")
    assert rel.max() <= TOL, rel.max()
    for j in (5, 66):
        assert_close(f"c2ppm uri row {j}", w[:, :, j, :], gold[f"uri_row{j}"])
    assert_close("c2ppm uu row 5", u[:, :, 5, :], gold["uu_row5"])
    names = [str(x) for x in gold["counter_names"]]
    ref_cnt = dict(zip(names, (int(x) for x in gold["counters"])))
    print("counters gpu", cnt, "reference run", ref_cnt)
    for k in names:
        assert cnt[k] == ref_cnt[k], k
