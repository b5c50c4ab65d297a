"""raishin_model_set_torus: the magnetised torus built on the device into a
model whose metric tables the device computed (raishin_model_create_grid).

- Against the host set-up (raishin_mdtorus1 on setup_metric's arrays, the
  same grid and parameters): the same torus cells, every primitive and
  conserved variable within 1e-12 of each field's maximum (the two share
  their arithmetic; the device's log, exp and pow round differently in the
  last bit, which the density and field normalisations carry to every cell:
  about 1e-13 measured), the same noise draws, time 0.
- The closed forms that pin the host torus (tests/test_torus_setup.py) hold
  for the device torus: one l = u^t u_phi equal to the circular geodesic's at
  r_max, one polytrope, min 2p/b^2 = beta.
- A model set this way steps like one given the host torus.
"""
from dataclasses import replace

import numpy as np
import pytest

from raishinsource_amd import configs, native
from test_torus_setup import l_fm, slice_fields

pytestmark = pytest.mark.gpu

CASES = [
    ("c1", configs.C1, dict(seed=0, noise=0.02)),                              # 2D, one stored phi row
    ("c1_spin05_beta10", replace(configs.C1, akm=0.5), dict(rin=8.0, rmax=14.0, beta=10.0, seed=4, noise=0.02)),
    ("t3dppm", configs.ALL["t3dppm"], dict(seed=7, noise=0.02)),              # 3D: a draw per cell
]


def host_and_device(p, kw):
    p = p.resolved()
    g = native.setup_grid(p)
    m = native.setup_metric(p, g)
    uri_h = native.setup_torus(p, g, m, **kw)
    ref = native.Model(p, g, m)
    ref.set_state(uri_h)
    w_h, u_h, _ = ref.get_state()
    ref.close()
    mod = native.Model.from_grid(p, g)
    mod.set_torus(**kw)
    w_d, u_d, t = mod.get_state()
    return p, g, m, mod, (w_h, u_h), (w_d, u_d, t)


def rel_err(a, b):
    return [float(np.abs(a[v] - b[v]).max() / max(np.abs(b[v]).max(), 1e-300)) for v in range(a.shape[0])]


@pytest.mark.parametrize("name,p,kw", CASES, ids=[c[0] for c in CASES])
def test_device_torus_matches_host(name, p, kw):
    p, g, m, mod, (w_h, u_h), (w_d, u_d, t) = host_and_device(p, kw)
    mod.close()
    assert t == 0.0
    torus_h = np.abs(w_h[1:4]).sum(axis=0) > 0
    torus_d = np.abs(w_d[1:4]).sum(axis=0) > 0
    assert torus_h.sum() > 100 and np.array_equal(torus_h, torus_d)
    ew, eu = rel_err(w_d, w_h), rel_err(u_d, u_h)
    print(name, "primitives", " ".join(f"{e:.1e}" for e in ew), "| conserved", " ".join(f"{e:.1e}" for e in eu),
          "| bitwise equal", float(np.mean(w_d == w_h)))
    assert max(ew) <= 1e-12 and max(eu) <= 1e-12


def test_device_torus_closed_forms():
    p, g, m, mod, _, (w_d, _, _) = host_and_device(configs.C1, dict(seed=0, noise=0.0))
    mod.close()
    w, u, ul, b2, _ = slice_fields(p, m, w_d)
    cells = (w[1] != 0) | (w[2] != 0) | (w[3] != 0)
    assert np.abs(u[0][cells] * ul[2][cells] / l_fm(p.akm, 12.0) - 1.0).max() < 1e-12
    K = w[4][cells] / w[0][cells] ** p.gam
    assert np.abs(K / K[0] - 1.0).max() < 1e-13
    assert abs(w[0].max() - 1.0) <= 2.3e-16
    has = b2 > 0
    assert abs((2.0 * w[4][has] / b2[has]).min() / 100.0 - 1.0) < 1e-12


def test_device_torus_steps_like_host_torus():
    p = configs.C1.resolved()
    g = native.setup_grid(p)
    m = native.setup_metric(p, g)
    a = native.Model(p, g, m)
    a.set_state(native.setup_torus(p, g, m, seed=0, noise=0.02))
    b = native.Model.from_grid(p, g)
    b.set_torus(seed=0, noise=0.02)
    ta, _ = a.step(10)
    tb, _ = b.step(10)
    wa, ua, _ = a.get_state()
    wb, ub, _ = b.get_state()
    a.close()
    b.close()
    assert np.isfinite(wb).all()
    assert abs(tb / ta - 1.0) < 1e-12
    e = rel_err(wb, wa)
    print("after 10 steps: primitives", " ".join(f"{x:.1e}" for x in e))
    assert max(e) <= 1e-10


def test_device_torus_rejects_flat_metric():
    p = configs.TUBE.resolved()
    g = native.setup_grid(p)
    mod = native.Model.from_grid(p, g)
    with pytest.raises(native.RaishinError):
        mod.set_torus()
    mod.close()
