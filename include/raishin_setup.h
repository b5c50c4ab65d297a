/* raishin_setup.h -- host-side set-up of a model (runs once per model).
 *
 * These build the inputs of the time step in the reference's layouts: the
 * grid (grid.f90:2), the metric and sqrt(-g) at centres and faces
 * (coord.f90:5 coord1, coord.f90:1016 caldetg), the Christoffel symbols
 * (coord.f90:1089 calchrist) and a magnetised Fishbone-Moncrief torus
 * (mdgrmhd.f90:879 mdtorus1).  Metrics built: metric=303 (Kerr-Schild) and
 * metric=1 (Minkowski, coord.f90:149-198).
 */
#ifndef RAISHIN_SETUP_H
#define RAISHIN_SETUP_H
#include "raishin_hip.h"
#ifdef __cplusplus
extern "C" {
#endif

/* Grid extents (pram.f90:52-61): ix1=1 uniform, 2 logarithmic radius. */
typedef struct {
  int ix1;
  double xmin, xmax, ymin, ymax, zmin, zmax;
} raishin_grid_ext;

/* Torus parameters (mdgrmhd.f90:938-947 fixes rin=6, rmax=12, beta=100). */
typedef struct {
  double rin, rmax;     /* inner edge, pressure maximum */
  double beta;          /* minimum plasma beta after normalisation */
  double acut;          /* A_phi = max(rho/rho_max - acut, 0) */
  double noise;         /* p <- p (1 + noise (U - 0.5)) */
  unsigned long long seed;
} raishin_torus;

/* grid.f90:2.  x*, dx*, x*a, x*b have n entries, dx*b has n-1 (n = imax, jmax, kmax). */
void raishin_grid(const raishin_pram *p, const raishin_grid_ext *ext, double *x1, double *x2, double *x3,
                  double *dx1, double *dx2, double *dx3, double *dx1b, double *dx2b, double *dx3b,
                  double *x1a, double *x2a, double *x3a, double *x1b, double *x2b, double *x3b);
/* coord.f90:5 with caldetg: metrics (4,4,imax,jmax,kmax), sqrt(-g) (imax,jmax,kmax). */
void raishin_coord1(const raishin_pram *p, double *gcov, double *gcovi, double *gcovj, double *gcovk,
                    double *gcon, double *gconi, double *gconj, double *gconk, double *detg,
                    double *detgi, double *detgj, double *detgk, const double *x1, const double *x2,
                    const double *x3, const double *x1a, const double *x2a, const double *x3a);
/* coord.f90:1089: christ(4,4,4,imax,jmax,kmax), christ(l,m,n) = Gamma^l_{mn}. */
void raishin_calchrist(const raishin_pram *p, double *christ, const double *x1, const double *x2,
                       const double *x3);
/* Primitive variables uri(9,imax,jmax,kmax) of the torus; returns 0 on success. */
int raishin_mdtorus1(const raishin_pram *p, const raishin_torus *t, double *uri, const double *gcov,
                     const double *gcon, const double *detg, const double *x1, const double *x2,
                     const double *x3);
/* The same torus (bit for bit) from the one phi row of the axisymmetric
 * metric: gcov/gcon(4,4,imax,kmax), detg(imax,kmax), as raishin_coord1 writes
 * them for a pram with jmax = 1.  The full metric of a 512x256 torus is
 * 2.3 GB of host arrays whose first touch costs seconds; the row is 16 MB. */
int raishin_mdtorus1_ik(const raishin_pram *p, const raishin_torus *t, double *uri, const double *gcov,
                        const double *gcon, const double *detg, const double *x1, const double *x2,
                        const double *x3);
/* The same torus built on the device, straight into the model's state
 * (primitives, conserved variables by caluu1, time 0, fluxes zeroed), from
 * the model's own device metric tables: a set-up with no host arrays, for a
 * model from raishin_model_create_grid (or any model of metric 303).  It
 * follows raishin_mdtorus1 step for step, including its noise stream; the
 * values differ from the host torus only where the device's log, exp and pow
 * round differently from glibc's (about 1e-15 relative).  Reference:
 * mdgrmhd.f90:879-1654 (mdtorus1), as raishin_mdtorus1; no counterpart of
 * its own in the reference, whose set-up is host Fortran.  Returns 0, or -1
 * with the error set (metric not 303, no torus cells on the grid). */
int raishin_model_set_torus(raishin_model *m, const raishin_torus *t);

/* Riemann-tube states (left [0], right [1]): density, pressure, 3-velocity,
 * magnetic field. */
typedef struct {
  double de[2], pr[2];
  double v[2][3];
  double b[2][3];
} raishin_tube;

/* Primitive variables uri(9,imax,jmax,kmax) of a Riemann tube on a flat
 * metric, jump at i = imax/2 (mdgrmhd.f90:55 mdshoctub with the states as
 * parameters); returns 0, or 1 for a non-physical state. */
int raishin_mdshoctub(const raishin_pram *p, const raishin_tube *t, double *uri);

#ifdef __cplusplus
}
#endif
#endif
