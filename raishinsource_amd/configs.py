"""Run configurations expressed as the reference's compile-time ``module pram``.

Names the parameter sets that the HIP library, the benchmark and the tests
run with.  Every field mirrors one ``pram.f90`` parameter (reference
``pram.f90:11-80``); the derived ghost widths follow ``main.f90:179-211``.
"""
from dataclasses import dataclass, field, replace
import math

PI = 3.14159265358979323  # pram.f90:31


def _xmin_ks(akm):
    # pram.f90:52  xmin=0.98d0*(1.+sqrt(1.-akm*akm))
    return 0.98 * (1.0 + math.sqrt(1.0 - akm * akm))


@dataclass(frozen=True)
class Pram:
    name: str
    imax: int
    jmax: int
    kmax: int
    irec: int = 5          # 1 MC, 5 PPM, 6 MP5, 7 WENO5
    irkt: int = 2
    ict: int = 1
    ihll: int = 1
    icha: int = 2          # 0 Gammie, 1 Leismann, 2 Del Zanna wave speeds
    iwvec: int = 1          # 1 recov1a, 2 recov2 (recov.f90:2-47)
    ieos: int = 0
    model: int = 7
    metric: int = 303
    cfl: float = 0.4
    gam: float = 4.0 / 3.0
    akm: float = 0.9375
    dmin: float = 1.0e-4
    pmin: float = 1.0e-6
    pmax: float = 1.0e6
    iter: int = 5000
    ix1: int = 2
    ix2: int = 1
    ix3: int = 1
    xmin: float = None
    xmax: float = 50.0
    ymin: float = 0.0
    ymax: float = None      # default: periodic phi, dx2 = 2 pi / n_interior
    zmin: float = 0.0
    zmax: float = PI
    iboux1in: tuple = (5, 5, 5, 5, 5, 5, 5, 5, 5)
    iboux1ot: tuple = (5, 18, 18, 18, 5, 5, 5, 5, 5)
    iboux2in: tuple = (1,) * 9
    iboux2ot: tuple = (1,) * 9
    iboux3in: tuple = (14, 14, 14, 12, 14, 14, 14, 14, 12)
    iboux3ot: tuple = (14, 14, 14, 12, 14, 14, 14, 14, 12)

    # ghost widths: main.f90:179-211
    @property
    def nm0(self):
        if self.irec in (1, 2, 3, 11, 12, 13):
            return 2
        if self.irec in (4, 6, 7, 8, 9, 10):
            return 3
        if self.irec == 5:
            return 4
        return 1

    @property
    def nm1(self):
        return self.nm0 + {0: 0, 1: 1, 2: 2, 3: 3}[self.ict]

    def resolved(self):
        """Fill the grid extents that default from other parameters."""
        p = self
        if p.xmin is None:
            p = replace(p, xmin=_xmin_ks(p.akm))
        if p.ymax is None:
            nint = p.jmax - 2 * p.nm1
            dx2 = 2.0 * PI / max(nint, 1)
            p = replace(p, ymax=p.ymin + dx2 * (p.jmax - 1))
        return p

    @property
    def interior(self):
        n = self.nm1
        return ((self.imax - 2 * n), (self.jmax - 2 * n), (self.kmax - 2 * n))


# 2D axisymmetric FM torus, PPM + RK2 (one phi cell plus ghosts)
C1 = Pram("c1", 128, 11, 128, irec=5, irkt=2, cfl=0.4)
# one GPU's 3D subdomain, MP5 + RK3
C2 = Pram("c2", 256, 128, 128, irec=6, irkt=3, cfl=0.5)
C2_PPM = replace(C2, name="c2ppm", irec=5, jmax=128)
# the CPU baseline's sample of c2ppm: the same r-theta grid and torus with 8
# interior phi cells (a bounded run of the reference build on one core)
C2_CPU = replace(C2_PPM, name="c2cpu", jmax=8 + 2 * 5)
C2_WENO = replace(C2, name="c2weno", irec=7)

# small grids used for unit-level golden vectors and short trajectories
U_PPM = Pram("uppm", 24, 12, 22, irec=5, irkt=2, cfl=0.4)
U_MP5 = Pram("ump5", 20, 12, 18, irec=6, irkt=3, cfl=0.5)
U_WENO = Pram("uweno", 20, 12, 18, irec=7, irkt=3, cfl=0.5)
T2D = Pram("t2d", 64, 11, 64, irec=5, irkt=2, cfl=0.4)
# golden-vector grids (tests/golden)
G_PPM = Pram("gppm", 14, 11, 12, irec=5, irkt=2, cfl=0.4)
G_MP5 = Pram("gmp5", 12, 9, 10, irec=6, irkt=3, cfl=0.5)
G_WENO = Pram("gweno", 12, 9, 10, irec=7, irkt=3, cfl=0.5)
T3D = Pram("t3d", 32, 16, 32, irec=6, irkt=3, cfl=0.5)
# a small 3D torus with the scheme of the 256x128x128 PPM run (PPM, RK3, CFL
# 0.5, 2% pressure noise): 30 x 14 x 30 interior cells
T3D_PPM = Pram("t3dppm", 40, 24, 40, irec=5, irkt=3, cfl=0.5)



def tube(name, nzones, nm1=5):
    """1D special-relativistic MHD Riemann tube: Minkowski (metric 1),
    Gamma 5/3, model 2 (mdshoctub, mdgrmhd.f90:55), x1 in [1, 2] uniform
    over nzones interior cells, one interior cell in x2 and x3 (cube cells),
    PPM, RK2, CFL 0.4; outflow (copy) in x1, periodic in x2 and x3."""
    imax = nzones + 2 * nm1
    d = 1.0 / (imax - 1)
    w = d * (2 * nm1)
    return Pram(name, imax, 2 * nm1 + 1, 2 * nm1 + 1, irec=5, irkt=2, cfl=0.4, gam=5.0 / 3.0, model=2, metric=1,
                ix1=1, xmin=1.0, xmax=2.0, ymin=0.0, ymax=w, zmin=0.0, zmax=w,
                iboux1in=(5,) * 9, iboux1ot=(5,) * 9, iboux2in=(1,) * 9, iboux2ot=(1,) * 9,
                iboux3in=(1,) * 9, iboux3ot=(1,) * 9)


# MC-limited linear reconstruction (irec=1, rec.f90:208): nm0=2, nm1=3
U_MC = Pram("umc", 16, 10, 18, irec=1, irkt=2, cfl=0.4)
T2D_MC = Pram("t2dmc", 64, 7, 64, irec=1, irkt=2, cfl=0.4)
# the other wave-speed formulas (calcha.f90:489-509): Leismann et al. (icha=1,
# with MC and RK3 as pram_3Dtorus.f90:66-70 pairs them) and Gammie et al. (icha=0)
U_LEIS = Pram("uleis", 16, 10, 18, irec=1, irkt=3, cfl=0.3, icha=1)
T2D_GAM = Pram("t2dgam", 64, 11, 64, irec=5, irkt=2, cfl=0.4, icha=0)
# recov2 (iwvec=2, recov.f90:1919-2454): alone, and with the whole scheme set
# of pram_3Dtorus.f90:66-72 (MC, Leismann speeds, RK3)
T2D_RC2 = Pram("t2drc2", 64, 11, 64, irec=5, irkt=2, cfl=0.4, iwvec=2)
U_RC2 = Pram("urc2", 16, 10, 18, irec=1, irkt=3, cfl=0.3, icha=1, iwvec=2)

# the serving mix's tube (1024 zones) and a small one for the oracle
TUBE = tube("tube", 1024)
U_TUBE = tube("utube", 64)

ALL = {p.name: p for p in (C1, C2, C2_PPM, C2_CPU, C2_WENO, U_PPM, U_MP5, U_WENO, T2D, T3D, T3D_PPM, G_PPM, G_MP5, G_WENO, TUBE,
                           U_TUBE, U_MC, T2D_MC, U_LEIS, T2D_GAM, T2D_RC2, U_RC2)}
