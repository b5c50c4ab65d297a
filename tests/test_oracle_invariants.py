"""Checks that pin the oracle by mathematics rather than by the reference:
exactness of the reconstructions on linear data, HLLE consistency, flux-CT
preserving the corner divergence of B, the recovery round trip, the CFL
limit of static gas in flat space, the Kerr-Schild metric identities, the
ghost-fill codes by brute force and the single-precision RK constants."""
from dataclasses import replace

import numpy as np
import pytest

from harness import Oracle, fzeros
from raishinsource_amd import configs, native


def flat(base, **kw):
    return replace(base, metric=1, ix1=1, xmin=1.0, xmax=2.0, ymin=0.0, ymax=1.0, zmin=0.0, zmax=1.0, model=2,
                   **kw).resolved()


def rec2(o, p, g, m, uri):
    ni, nj, nk = p.imax, p.jmax, p.kmax
    names = ["uriir", "urijr", "urikr", "uriil", "urijl", "urikl", "uuir", "uujr", "uukr", "uuil", "uujl", "uukl",
             "wwir", "wwjr", "wwkr", "wwil", "wwjl", "wwkl"]
    F = {n: fzeros(9, ni, nj, nk) for n in names}
    o.call("rec2", uri, g["x1"], g["x2"], g["x3"], g["dx1b"], g["dx2b"], g["dx3b"], *[F[n] for n in names],
           m["gcovi"], m["gcovj"], m["gcovk"], m["gconi"], m["gconj"], m["gconk"], m["detgi"], m["detgj"], m["detgk"],
           p.nm0)
    return F


@pytest.mark.parametrize("irec", [5, 6, 7, 1])
def test_reconstruction_exact_on_linear_data(irec, oracle_lib):
    """PPM, MP5, WENO5 and the MC limiter give exact interface values for
    linear data on a uniform grid (the limiters do not act): right edge of
    cell i = u(i+1/2)."""
    p = flat(configs.G_PPM if irec == 5 else configs.G_MP5, irec=irec, imax=16, jmax=11, kmax=12)
    g = native.setup_grid(p)
    m = native.setup_metric(p, g)
    o = Oracle(p)
    o.set_faces(g)
    uri = fzeros(9, p.imax, p.jmax, p.kmax)
    x = g["x1"][:, None, None] + 0 * g["x2"][None, :, None] + 0 * g["x3"][None, None, :]
    uri[0] = 1.0 + 0.3 * x
    uri[4] = 0.7            # constant pressure: no flattening or steepening
    uri[1] = 0.05 * x
    uri[6] = -0.1 + 0.02 * x
    F = rec2(o, p, g, m, uri)
    nm0 = p.nm0
    i = np.arange(nm0 - 1, p.imax - nm0)            # interfaces i+1/2 consumed by hll (0-based i)
    xa = g["x1a"][i]                                 # x1a(i) = x1(i+1/2)
    # MC divides the differences by (dx + 1e-12) (rec.f90:263-264): a slope
    # error of 1e-12/dx, i.e. ~1e-13 at the edges here
    tol = 1e-12 if irec == 1 else 1e-13
    for n, (c0, c1) in {0: (1.0, 0.3), 1: (0.0, 0.05), 6: (-0.1, 0.02)}.items():
        exact = c0 + c1 * xa
        left = F["uriil"][n, i, 5, 5]                # right edge of cell i
        right = F["uriir"][n, i, 5, 5]               # left edge of cell i+1
        assert np.allclose(left, exact, rtol=0, atol=tol), (n, left - exact)
        assert np.allclose(right, exact, rtol=0, atol=tol), (n, right - exact)


def test_hlle_consistency(oracle_lib):
    """Identical left and right states give the physical flux: 0.5*(F+F) = F exactly."""
    p = configs.G_MP5.resolved()
    o = Oracle(p)
    rng = np.random.default_rng(0)
    sh = (9, p.imax, p.jmax, p.kmax)
    U = np.asfortranarray(rng.standard_normal(sh))
    Fx = [np.asfortranarray(rng.standard_normal(sh)) for _ in range(3)]
    c = [np.asfortranarray(np.full(sh[1:], v)) for v in (0.7, -0.4, 0.3, -0.9, 1.0, -1.0)]
    ww = fzeros(3, *sh)
    o.call("hll", ww, U, U, U, U, U, U, Fx[0], Fx[1], Fx[2], Fx[0], Fx[1], Fx[2], *c, p.nm0)
    nm0 = p.nm0
    sl = (slice(None), slice(nm0 - 1, p.imax - nm0), slice(nm0 - 1, p.jmax - nm0), slice(nm0 - 1, p.kmax - nm0))
    for d in range(3):
        assert np.array_equal(ww[d][sl], Fx[d][sl])


def test_fluxct_preserves_corner_divergence(oracle_lib):
    """With zero normal B fluxes, CT'd fluxes change the corner-centred
    divergence of B (Toth 2000) only by round-off."""
    p = replace(configs.G_MP5, imax=16, jmax=14, kmax=15).resolved()
    o = Oracle(p)
    rng = np.random.default_rng(1)
    ww = np.asfortranarray(rng.standard_normal((3, 9, p.imax, p.jmax, p.kmax)))
    ww[0, 6] = ww[1, 7] = ww[2, 8] = 0.0
    o.call("ct", ww, p.nm0, p.nm1)
    n1 = p.nm1  # corners whose surrounding CT'd fluxes all lie inside the CT range
    scale = np.abs(ww[:, 6:9]).max()
    assert np.abs(corner_divergence(ww)[n1:-n1, n1:-n1, n1:-n1]).max() < 1e-13 * scale
    # an unbalanced CT flux (one EMF term scaled by 1.001) is seen
    bad = ww.copy(order="F")
    bad[0, 7, 6, 6, 6] *= 1.001
    assert np.abs(corner_divergence(bad)[n1:-n1, n1:-n1, n1:-n1]).max() > 1e-6 * scale


def corner_divergence(ww):
    """Corner-centred divergence of dB/dt = -div F for B^x, B^y, B^z
    (vars 7..9) on a unit grid, averaged over the 8 cells around a corner."""
    dB = []
    for n in (6, 7, 8):
        f = ww[:, n]
        d = (f[0, 1:, 1:, 1:] - f[0, :-1, 1:, 1:]) + (f[1, 1:, 1:, 1:] - f[1, 1:, :-1, 1:]) + \
            (f[2, 1:, 1:, 1:] - f[2, 1:, 1:, :-1])
        dB.append(-d)
    bx, by, bz = dB
    return 0.25 * ((bx[1:, :-1, :-1] + bx[1:, 1:, :-1] + bx[1:, :-1, 1:] + bx[1:, 1:, 1:])
                   - (bx[:-1, :-1, :-1] + bx[:-1, 1:, :-1] + bx[:-1, :-1, 1:] + bx[:-1, 1:, 1:])
                   + (by[:-1, 1:, :-1] + by[1:, 1:, :-1] + by[:-1, 1:, 1:] + by[1:, 1:, 1:])
                   - (by[:-1, :-1, :-1] + by[1:, :-1, :-1] + by[:-1, :-1, 1:] + by[1:, :-1, 1:])
                   + (bz[:-1, :-1, 1:] + bz[1:, :-1, 1:] + bz[:-1, 1:, 1:] + bz[1:, 1:, 1:])
                   - (bz[:-1, :-1, :-1] + bz[1:, :-1, :-1] + bz[:-1, 1:, :-1] + bz[1:, 1:, :-1]))


@pytest.mark.parametrize("iwvec,tol", [(1, 1e-6), (2, 1e-5)])
def test_recovery_round_trip(oracle_lib, iwvec, tol):
    """uri -> caluu1 -> recov returns the primitives: recov1a's 2D
    Newton-Raphson (tolerance 1e-7, recov.f90:645) and recov2's 1D one on
    W - D (tolerance 1e-8, recov.f90:1986), all four velocity and pressure
    fields, from a poor initial guess."""
    from dataclasses import replace as _r
    p = _r(configs.T2D, iwvec=iwvec).resolved()
    g = native.setup_grid(p)
    m = native.setup_metric(p, g)
    uri = native.setup_torus(p, g, m)
    o = Oracle(p)
    uu = fzeros(9, p.imax, p.jmax, p.kmax)
    o.call("caluu1", uri, uu, m["gcov"], m["gcon"], m["detg"])
    w = uri.copy(order="F")
    w[[0, 4]] *= 1.3   # a poor initial guess
    o.reset_counters()
    o.call("recov", uu, w, m["gcov"], m["gcon"], m["detg"], g["x1"], p.nm1)
    assert o.counters()["recov_fail"] == 0
    n1 = p.nm1
    s = (slice(None), slice(n1, p.imax - n1), slice(n1, p.jmax - n1), slice(n1, p.kmax - n1))
    for v in (0, 4):
        rel = np.abs(w[v][s[1:]] - uri[v][s[1:]]) / np.abs(uri[v][s[1:]])
        assert rel.max() < tol, (v, rel.max())
    vmax = np.abs(uri[1:4][s]).max()
    for v in (1, 2, 3):
        err = np.abs(w[v][s[1:]] - uri[v][s[1:]]).max() / vmax
        assert err < tol, (v, err)


def test_cfl_static_gas_flat_metric(oracle_lib):
    """Flat metric, gas at rest, no field: the fast speed is the sound speed
    and dt = min dx / (c_s + 1e-12) (cdtcfl.f90:75-104)."""
    import ctypes as C
    p = flat(configs.G_MP5, imax=16, jmax=14, kmax=15)
    g = native.setup_grid(p)
    m = native.setup_metric(p, g)
    o = Oracle(p)
    rng = np.random.default_rng(2)
    uri = fzeros(9, p.imax, p.jmax, p.kmax)
    uri[0] = 0.5 + rng.random(uri[0].shape)
    uri[4] = 0.1 + rng.random(uri[0].shape)
    d = [C.c_double(0) for _ in range(4)]
    o.call("cdtcfl4", uri, m["gcov"], m["gcon"], C.byref(d[0]), g["dx1b"], g["dx2b"], g["dx3b"],
           C.byref(d[1]), C.byref(d[2]), C.byref(d[3]), p.nm1)
    roh = uri[0] + (p.gam / (p.gam - 1.0)) * uri[4]
    cs = np.sqrt(p.gam * uri[4] / roh)[1:-1, 1:-1, 1:-1]
    dts = [np.abs(g["dx%db" % a][idx] / (cs + 1e-12)).min() for a, idx in
           ((1, np.arange(1, p.imax - 1)[:, None, None]), (2, np.arange(1, p.jmax - 1)[None, :, None]),
            (3, np.arange(1, p.kmax - 1)[None, None, :]))]
    assert np.isclose(d[0].value, min(dts), rtol=1e-15, atol=0)
    assert np.isclose(d[1].value, dts[0], rtol=1e-15, atol=0)


def test_kerr_schild_metric_identities():
    p = configs.T3D.resolved()
    g = native.setup_grid(p)
    m = native.setup_metric(p, g)
    a = p.akm
    gc = np.moveaxis(m["gcov"], (0, 1), (-2, -1))
    gi = np.moveaxis(m["gcon"], (0, 1), (-2, -1))
    prod = gc @ gi
    assert np.abs(prod - np.eye(4)).max() < 1e-12
    r = g["x1"][:, None, None]
    th = g["x3"][None, None, :]
    sig = r * r + a * a * np.cos(th) ** 2
    sdg = sig * np.abs(np.sin(th))
    inner = np.abs(np.sin(th[0, 0])) > 1e-6
    rel = np.abs(m["detg"][:, :, inner] - np.broadcast_to(sdg, m["detg"].shape)[:, :, inner]) / sdg[:, :, inner]
    assert rel.max() < 1e-12


def test_flat_metric_no_source(oracle_lib):
    p = flat(configs.G_MP5)
    g = native.setup_grid(p)
    m = native.setup_metric(p, g)
    assert not m["christ"].any()
    from harness import synthetic_state
    o = Oracle(p)
    uri = synthetic_state(p, 5, atmosphere=False)
    tenr = fzeros(4, 4, p.imax, p.jmax, p.kmax)
    sf = fzeros(4, p.imax, p.jmax, p.kmax)
    o.call("caltenr", uri, tenr, m["gcov"], m["gcon"])
    o.call("calsf", sf, tenr, m["christ"], m["detg"])
    assert not sf.any()


def test_ghost_fill_codes(oracle_lib):
    """bnd4 by brute force: periodic phi, copy (5) and outflow-no-inflow (18)
    in r, symmetric (14) and antisymmetric (12) in theta (pram.f90:75-80)."""
    p = configs.G_MP5.resolved()
    p = replace(p, imax=14, jmax=13, kmax=15).resolved()
    o = Oracle(p)
    n1 = p.nm1
    rng = np.random.default_rng(3)
    u = np.asfortranarray(rng.standard_normal((9, p.imax, p.jmax, p.kmax)))
    u0 = u.copy(order="F")
    o.call("bnd4", u, u0, n1)
    I = slice(n1, p.imax - n1)
    W = p.jmax - 2 * n1
    for j in range(n1):                                   # periodic in j (interior i, k)
        jj = n1 + ((j - n1) % W)
        assert np.array_equal(u[:, I, j, n1:p.kmax - n1], u0[:, I, jj, n1:p.kmax - n1])
    for n in range(1, n1 + 1):                            # theta mirror: k=n <- k=2nm1-n+1 (1-based)
        src = u[:, :, :, 2 * n1 - n]
        assert np.array_equal(u[3, :, :, n - 1], -src[3])    # var 4 antisymmetric (12)
        assert np.array_equal(u[0, :, :, n - 1], src[0])     # var 1 symmetric (14)
    for n in range(n1):                                   # r inner copy (5): u(n) = u(nm1+1) before theta pass
        k = p.kmax // 2
        assert np.array_equal(u[0, n, n1:p.jmax - n1, k], u[0, n1, n1:p.jmax - n1, k])
    k = p.kmax // 2
    for j in range(n1, p.jmax - n1):                      # r outer: velocity zero where v^r < 0 at the edge
        edge = u[1, p.imax - n1 - 1, j, k]
        for n in range(n1):
            expect = 0.0 if edge < 0 else edge
            assert u[1, p.imax - 1 - n, j, k] == expect


def ghost_fill_by_index(u, uo, p, order=(0, 1, 2)):
    """bnd4 (bnd.f90:4-1059, one rank) as whole-face index maps rather than
    loops: the periodic copies over the interior ranges of the other two axes
    (bndpefb4, bnd.f90:25-381), then for each axis in turn (x1, x2, x3,
    bnd.f90:419, 624, 817) and each variable the inner and outer ghost layers
    over the full extent of the other axes (bndryfb4).  Each code, as the
    reference writes it: 2 keep uo; 3 (4u(n+1)-u(n+2))/3 and 4 2u(n+1)-u(n+2),
    outward one layer at a time; 5 copy the first interior layer; 7 uo plus
    the change of the next layer inward; 12/14 mirror (negated / not) about
    the boundary face; 13 zero; 18 (x1 outer) copy the last interior layer,
    or zero where its v^1 (var 2) is negative."""
    u = u.copy(order="F")
    n1 = p.nm1
    ext = (p.imax, p.jmax, p.kmax)
    bcs = ((p.iboux1in, p.iboux1ot), (p.iboux2in, p.iboux2ot), (p.iboux3in, p.iboux3ot))

    def sl(ax, idx, inner_only=False):
        s = [slice(None)] * 3
        if inner_only:
            s = [slice(n1, e - n1) for e in ext]
        s[ax] = idx
        return (slice(None),) + tuple(s)

    # periodic pass: ghost layer g (0-based) takes the interior layer one
    # interior length away; only interior cells of the other two axes
    for ax in range(3):
        e = ext[ax]
        nint = e - 2 * n1
        per = [m for m in range(9) if bcs[ax][0][m] == 1 or bcs[ax][1][m] == 1]
        if not per:
            continue
        lo = np.arange(n1)
        hi = np.arange(e - n1, e)
        for m in per:
            v = u[m:m + 1]
            v[sl(ax, lo, True)] = v[sl(ax, lo + nint, True)]
            v[sl(ax, hi, True)] = v[sl(ax, hi - nint, True)]
    for ax in order:
        e = ext[ax]
        for m in range(9):
            for outer in (0, 1):
                code = bcs[ax][outer][m]
                v, vo = u[m:m + 1], uo[m:m + 1]
                # ghost layers ordered outward from the boundary face, and
                # the interior layers mirrored onto them
                ghosts = list(range(n1 - 1, -1, -1)) if not outer else list(range(e - n1, e))
                inside = list(range(n1, 2 * n1)) if not outer else list(range(e - n1 - 1, e - 2 * n1 - 1, -1))
                step = -1 if not outer else 1
                if code == 1:
                    continue
                for q, g in enumerate(ghosts):
                    if code == 2:
                        v[sl(ax, g)] = vo[sl(ax, g)]
                    elif code == 3:
                        v[sl(ax, g)] = (4.0 * v[sl(ax, g - step)] - v[sl(ax, g - 2 * step)]) / 3.0
                    elif code == 4:
                        v[sl(ax, g)] = 2.0 * v[sl(ax, g - step)] - v[sl(ax, g - 2 * step)]
                    elif code == 5:
                        v[sl(ax, g)] = v[sl(ax, inside[0])]
                    elif code == 7:
                        v[sl(ax, g)] = vo[sl(ax, g)] + v[sl(ax, g - step)] - vo[sl(ax, g - step)]
                    elif code in (12, 14):
                        v[sl(ax, g)] = (-1.0 if code == 12 else 1.0) * v[sl(ax, inside[q])]
                    elif code == 13:
                        v[sl(ax, g)] = 0.0
                    elif code == 18:
                        assert ax == 0 and outer
                        edge = u[1:2][sl(ax, e - n1 - 1)]
                        v[sl(ax, g)] = np.where(edge < 0.0, 0.0, v[sl(ax, e - n1 - 1)])
                    else:
                        raise ValueError(code)
    return u


GHOST_SETS = {
    # the torus: copy / outflow-no-inflow in r, periodic phi, mirror theta
    "torus": {},
    # every other code on some variable of some face
    "codes": dict(iboux1in=(2, 3, 4, 5, 7, 12, 13, 14, 5), iboux1ot=(18, 18, 18, 7, 5, 4, 3, 2, 13),
                  iboux2in=(12, 14, 13, 2, 3, 4, 5, 7, 14), iboux2ot=(14, 12, 2, 3, 4, 5, 7, 13, 12),
                  iboux3in=(3, 4, 5, 7, 12, 13, 14, 2, 12), iboux3ot=(5, 7, 12, 13, 14, 2, 3, 4, 14)),
    # periodic in every axis (the Riemann tubes' x2, x3)
    "periodic": dict(iboux1in=(1,) * 9, iboux1ot=(1,) * 9, iboux3in=(1,) * 9, iboux3ot=(1,) * 9),
}


@pytest.mark.parametrize("which", sorted(GHOST_SETS))
def test_ghost_fill_every_cell(which, oracle_lib):
    """bnd4 against the index-map ghost fill on every cell, edges and corners
    included, for every variable and every boundary code the oracle builds.
    With every code present, swapping the axis order changes the corners, so
    the comparison pins the x1 -> x2 -> x3 order too (the torus's copy in r
    and mirror in theta commute at the corners)."""
    p = replace(configs.G_MP5, imax=13, jmax=12, kmax=14, **GHOST_SETS[which]).resolved()
    o = Oracle(p)
    rng = np.random.default_rng(21)
    u = np.asfortranarray(rng.standard_normal((9, p.imax, p.jmax, p.kmax)))
    uo = np.asfortranarray(rng.standard_normal((9, p.imax, p.jmax, p.kmax)))
    want = ghost_fill_by_index(u, uo, p)
    got = u.copy(order="F")
    o.call("bnd4", got, uo, p.nm1)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
    if which == "codes":
        swapped = ghost_fill_by_index(u, uo, p, order=(2, 1, 0))
        assert not np.array_equal(swapped, want)


def test_infchk_stops_inflow_at_the_outer_boundary(oracle_lib):
    """infchk (bnd.f90:1636-1824) on a Kerr-Schild grid with inflow (v^r < 0)
    and outflow in the outer r ghosts: where v^r < 0 the radial velocity is
    replaced by beta^r / alpha, i.e. the contravariant u^r = u~^r - gamma
    beta^r / alpha becomes zero; v^theta, v^phi are kept; the new u~ is
    gamma v with gamma = 1/sqrt(1 - v^2), v^2 clamped at 1 - 1/10^2 (so
    gamma = 10) where it reaches 1; every other cell, the inner r ghosts
    (whose branch writes nothing) and the other variables are unchanged."""
    p = replace(configs.T3D, imax=14, jmax=10, kmax=12).resolved()
    g = native.setup_grid(p)
    m = native.setup_metric(p, g)
    o = Oracle(p)
    rng = np.random.default_rng(8)
    uri = np.asfortranarray(rng.standard_normal((9, p.imax, p.jmax, p.kmax)) * 0.3)
    uri[0] = 1.0
    uri[4] = 0.1
    n1 = p.nm1
    # a quarter of the outer ghosts get a large tangential speed, so that
    # v^2 >= 1 once v^r is reset and the clamp acts
    big = rng.random((n1, p.jmax, p.kmax)) < 0.25
    uri[3, p.imax - n1:][big] = 40.0
    w = uri.copy(order="F")
    o.call("infchk", w, m["gcov"], m["gcon"], n1)

    def metric(name, i, j, k):
        return m[name][:, :, i, j, k]

    def quad(gc, v):
        # g_ij v^i v^j summed m-outer, n-inner as physfunc.f90:17-21 and
        # bnd.f90:1795-1800 do (v^2 near 1 amplifies the summation order)
        acc = 0.0
        for a in range(3):
            for b in range(3):
                acc = acc + gc[1 + a, 1 + b] * v[a] * v[b]
        return acc

    def gamma_of(ut, gc):
        # calgfl (physfunc.f90:2-41): g_ij u~^i u~^j clamped to [.., 500]
        # (small negatives taken absolute, larger ones set to 1e-10)
        utsq = quad(gc, ut)
        if utsq < 0.0:
            utsq = abs(utsq) if abs(utsq) <= 1e-10 else 1e-10
        return np.sqrt(1.0 + min(utsq, 500.0))

    changed = clamped = kept = 0
    for i in range(p.imax):
        for j in range(p.jmax):
            for k in range(p.kmax):
                before, after = uri[:, i, j, k], w[:, i, j, k]
                gc, gi = metric("gcov", i, j, k), metric("gcon", i, j, k)
                ut = before[1:4]
                gam0 = gamma_of(ut, gc)
                v0 = ut * (1.0 / gam0)          # bnd.f90 multiplies by reciprocals
                if i < p.imax - n1 or v0[0] >= 0.0:
                    assert np.array_equal(after, before), (i, j, k)
                    kept += 1
                    continue
                assert np.array_equal(np.delete(after, [1, 2, 3]), np.delete(before, [1, 2, 3]))
                alpha = 1.0 / np.sqrt(-gi[0, 0])
                beta_r = alpha * alpha * gi[0, 1]
                v = np.array([beta_r * (1.0 / alpha), v0[1], v0[2]])
                vsq = quad(gc, v)
                gam = 10.0 if vsq >= 1.0 else 1.0 / np.sqrt(1.0 - vsq)
                clamped += vsq >= 1.0
                np.testing.assert_allclose(after[1:4], gam * v, rtol=1e-13, atol=1e-14)
                if vsq < 1.0 and gam * gam - 1.0 < 500.0:
                    # the new u~ is self-consistent (below calgfl's clamp):
                    # gamma from it equals gamma, and the radial 4-velocity
                    # vanishes
                    np.testing.assert_allclose(gamma_of(after[1:4], gc), gam, rtol=1e-11)
                    ucon_r = after[1] - gam * beta_r / alpha
                    assert abs(ucon_r) <= 1e-13 * max(1.0, abs(after[1])), (i, j, k, ucon_r)
                changed += 1
    assert changed > 10 and clamped > 3 and kept > changed


def test_rk3_single_precision_constants(oracle_lib):
    """rk3trd uses the float32 literals 1./3. and 2./3. (rkt.f90:169-172)."""
    p = configs.G_MP5.resolved()
    o = Oracle(p)
    rng = np.random.default_rng(4)
    sh = (9, p.imax, p.jmax, p.kmax)
    uo = np.asfortranarray(rng.random(sh))
    us = np.asfortranarray(rng.random(sh))
    ww = fzeros(3, *sh)
    un = fzeros(*sh)
    a = np.ones(max(sh[1:]))
    o.call("rk3trd", un, ww, uo, us, p.nm1, a, a, a)
    n1 = p.nm1
    s = (slice(None), slice(n1, p.imax - n1), slice(n1, p.jmax - n1), slice(n1, p.kmax - n1))
    third32 = float(np.float32(1.0) / np.float32(3.0))
    assert third32 != 1.0 / 3.0
    expect = (uo + 2.0 * us) * third32 + 0.0 * float(np.float32(2.0) / np.float32(3.0))
    assert np.array_equal(un[s], expect[s])


def test_tube_uniform_state_stays_uniform(oracle_lib):
    """A Riemann tube with equal left and right states on the flat metric:
    every interface sees the same face states, so the flux divergence is
    exactly zero and every cell along the tube keeps one value (bit for bit),
    within the recovery's tolerance of the initial state.  (The x2-x3 corner
    ghosts keep their initial values: the periodic pass copies only over the
    interior ranges of the other two axes, bnd.f90:50-70.)"""
    p = configs.U_TUBE.resolved()
    g = native.setup_grid(p)
    m = native.setup_metric(p, g)
    st = {"de": 0.3, "pr": 0.2, "v": [0.1, -0.2, 0.05], "b": [0.3, 0.2, -0.1]}
    uri = native.setup_tube(p, st, st)
    o = Oracle(p)
    uu = fzeros(9, p.imax, p.jmax, p.kmax)
    o.call("caluu1", uri, uu, m["gcov"], m["gcon"], m["detg"])
    w = uri.copy(order="F")
    t = 0.0
    for _ in range(5):
        dt, t = o.step(uu, w, g, m, t)
    o.lib.orc_free_work()
    j = p.nm1
    line = w[:, :, j, j]
    for v in range(9):
        assert np.unique(line[v]).size == 1, v
    assert np.allclose(line[[0, 1, 2, 3, 4, 6, 7, 8], 0], uri[[0, 1, 2, 3, 4, 6, 7, 8], 0, 0, 0], rtol=1e-7)
    # recov1a rewrites the entropy tracer as rho * p / rho**(gam-1) (recov.f90:1186)
    assert np.isclose(line[5, 0], 0.3 * 0.2 / 0.3 ** (p.gam - 1.0), rtol=1e-7)
    # a tube of cube cells at |v| < 1: dt = cfl * dx / (fastest speed) with speed < 1
    assert dt > 0.4 * g["dx1b"][1]


def _speeds(p, m, uri):
    o = Oracle(p)
    sh = (p.imax, p.jmax, p.kmax)
    c = [np.asfortranarray(np.zeros(sh)) for _ in range(6)]
    o.call("calcha", uri, uri, uri, uri, uri, uri, m["gcovi"], m["gcovj"], m["gcovk"], m["gconi"], m["gconj"],
           m["gconk"], *c, p.nm0)
    n = p.nm0
    sl = (slice(n - 1, p.imax - n), slice(n - 1, p.jmax - n), slice(n - 1, p.kmax - n))
    return np.stack([x[sl] for x in c])


def _state(p, rng, field, velocity):
    uri = fzeros(9, p.imax, p.jmax, p.kmax)
    uri[0] = 0.5 + rng.random(uri[0].shape)
    uri[4] = 0.1 + rng.random(uri[0].shape)
    if velocity:
        uri[1:4] = 0.3 * rng.standard_normal(uri[1:4].shape)
    if field:
        uri[6:9] = 0.5 * rng.standard_normal(uri[6:9].shape)
    return uri


def test_wave_speed_formulas_agree_where_they_must(oracle_lib):
    """calcha.f90:489-527.  Without a field (v.B = 0, so r1 = 0) Leismann's
    f2 and f3 equal Del Zanna's algebraically: the speeds agree to round-off
    in the Kerr-Schild metric.  At rest in flat space all three formulas give
    +-sqrt(omsq) (Del Zanna's f3 = omsq * gcon3a(d,d) with gcon3a = 1).  With
    a field and motion the Leismann term r1 changes the speeds."""
    p = configs.U_PPM.resolved()
    g = native.setup_grid(p)
    m = native.setup_metric(p, g)
    rng = np.random.default_rng(4)
    uri = _state(p, rng, field=False, velocity=True)
    s1, s2 = _speeds(replace(p, icha=1), m, uri), _speeds(replace(p, icha=2), m, uri)
    # the few states that leave f3 < 0 give NaN in both, at the same places
    assert np.array_equal(np.isnan(s1), np.isnan(s2))
    assert np.allclose(s1, s2, rtol=1e-12, atol=1e-14, equal_nan=True)
    assert not np.array_equal(_speeds(replace(p, icha=0), m, uri), s2, equal_nan=True)
    uri = _state(p, rng, field=True, velocity=True)
    s1, s2 = _speeds(replace(p, icha=1), m, uri), _speeds(replace(p, icha=2), m, uri)
    assert np.nanmax(np.abs(s1 - s2)) > 1e-3
    q = flat(configs.G_MP5, imax=12, jmax=10, kmax=11)
    gq = native.setup_grid(q)
    mq = native.setup_metric(q, gq)
    uri = _state(q, rng, field=True, velocity=False)
    s = [_speeds(replace(q, icha=i), mq, uri) for i in (0, 1, 2)]
    roh = uri[0] + (q.gam / (q.gam - 1.0)) * uri[4]
    bsq = (uri[6] ** 2 + uri[7] ** 2 + uri[8] ** 2)
    va = bsq / (roh + bsq)
    om = np.sqrt(va + q.gam * uri[4] / roh * (1.0 - va))
    n = q.nm0
    om = om[n - 1:q.imax - n, n - 1:q.jmax - n, n - 1:q.kmax - n]
    for x in s:
        assert np.allclose(x[0::2], np.minimum(om, 1.0)[None], rtol=1e-14, atol=0)
        assert np.allclose(x[1::2], -np.minimum(om, 1.0)[None], rtol=1e-14, atol=0)
