"""MI355X-native RAISHIN GRMHD time step (HIP kernels for gfx950 behind a C-ABI).

    from raishinsource_amd import configs, native
    p = configs.C1
    g = native.setup_grid(p); m = native.setup_metric(p, g)
    uri = native.setup_torus(p, g, m)
    model = native.Model(p, g, m); model.set_state(uri)
    time, dt = model.step(100)
"""
from . import configs, native  # noqa: F401

__all__ = ["configs", "native"]
