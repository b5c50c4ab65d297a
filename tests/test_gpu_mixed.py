"""The mixed-precision form (reconstruction and face fluxes in FP32, state,
metric, CT, RK, sources and recovery in FP64) against the FP64 path: the
density's relative L1 error stays within 1e-5 and the form is really active."""
import numpy as np
import pytest

from raishinsource_amd import configs, native

pytestmark = pytest.mark.gpu


def run(p, g, m, uri, nsteps, mixed):
    mod = native.Model(p, g, m)
    mod.set_state(uri)
    mod.set_precision(mixed)
    t, _ = mod.step(nsteps)
    w, u, _ = mod.get_state()
    cnt = mod.counters()
    mod.close()
    return w, t, cnt


@pytest.mark.parametrize("name,nsteps", [("t2d", 20), ("uppm", 20), ("c1", 10), ("umc", 20)])
def test_mixed_density_l1_within_1e5(name, nsteps):
    p = configs.ALL[name].resolved()
    g = native.setup_grid(p)
    m = native.setup_metric(p, g)
    uri = native.setup_torus(p, g, m, seed=0, noise=0.02)
    w64, t64, c64 = run(p, g, m, uri, nsteps, False)
    w32, t32, c32 = run(p, g, m, uri, nsteps, True)
    n1 = p.nm1
    s = (slice(n1, p.imax - n1), slice(n1, p.jmax - n1), slice(n1, p.kmax - n1))
    a, b = w32[0][s], w64[0][s]
    fin = np.isfinite(a) & np.isfinite(b)
    assert np.array_equal(np.isfinite(a), np.isfinite(b))
    l1 = np.abs(a[fin] - b[fin]).sum() / np.abs(b[fin]).sum()
    print(f"{name}: density rel L1 {l1:.3e} after {nsteps} steps; time {t32!r} vs {t64!r}; "
          f"fallbacks {c32['recov_fallback']} vs {c64['recov_fallback']}")
    assert 0.0 < l1 <= 1e-5
    assert abs(t32 / t64 - 1) < 1e-4


@pytest.mark.parametrize("name", ["uppm", "ump5"])
def test_precision_round_trip_keeps_stored_fluxes(name):
    """The mixed form stores the face fluxes as float; switching to FP64 and
    back converts them there and back exactly, so a model that switched in
    the middle continues bit for bit like one that never did (ump5 has cells
    whose unordered wave speeds make a sweep keep the stored flux)."""
    p = configs.ALL[name].resolved()
    g = native.setup_grid(p)
    m = native.setup_metric(p, g)
    uri = native.setup_torus(p, g, m, seed=0, noise=0.02)
    a, b = native.Model(p, g, m), native.Model(p, g, m)
    for mod in (a, b):
        mod.set_state(uri)
        mod.set_precision(True)
        mod.step(3)
    a.set_precision(False)
    a.set_precision(True)
    ta, _ = a.step(4)
    tb, _ = b.step(4)
    wa, ua, _ = a.get_state()
    wb, ub, _ = b.get_state()
    a.close()
    b.close()
    assert ta == tb
    assert np.array_equal(wa, wb, equal_nan=True) and np.array_equal(ua, ub, equal_nan=True)


@pytest.mark.parametrize("name,nsteps", [("uppm", 2), ("t2d", 3), ("umc", 2)])
def test_stored_flux_conversion(name, nsteps):
    """A change of form converts the stored face fluxes (ww and flux-CT's
    six transverse-B copies) in place: FP64 -> mixed rounds each value to
    float, mixed -> FP64 widens it back exactly.  Read through the flux tap,
    on states whose fluxes are finite and nonzero."""
    p = configs.ALL[name].resolved()
    g = native.setup_grid(p)
    m = native.setup_metric(p, g)
    uri = native.setup_torus(p, g, m, seed=0, noise=0.02)
    mod = native.Model(p, g, m)
    mod.set_state(uri)
    mod.step(nsteps)
    ww64, ct64 = mod.fluxes()
    assert np.isfinite(ww64).all() and np.isfinite(ct64).all()
    assert np.count_nonzero(ww64) > 1000 and np.count_nonzero(ct64) > 1000
    exp_ww = ww64.astype(np.float32).astype(np.float64)
    exp_ct = ct64.astype(np.float32).astype(np.float64)
    assert not np.array_equal(exp_ww, ww64)  # the rounding is visible
    mod.set_precision(True)
    ww32, ct32 = mod.fluxes()
    assert np.array_equal(ww32, exp_ww) and np.array_equal(ct32, exp_ct)
    mod.set_precision(False)
    ww_back, ct_back = mod.fluxes()
    assert np.array_equal(ww_back, exp_ww) and np.array_equal(ct_back, exp_ct)
    # a model that switched still steps: the state stays finite
    mod.set_precision(True)
    mod.step(2)
    w, _, _ = mod.get_state()
    mod.close()
    assert np.isfinite(w).all()
