/* CPU restatement of the RAISHIN GRMHD time step -- TEST INFRASTRUCTURE ONLY.
 *
 * Each function restates one reference subroutine literally: same loop
 * ranges, same operation order, same single-precision literals, so that with
 * gcc -O2 -ffp-contract=off on x86-64 (no FMA) it reproduces the reference
 * Fortran (flang -O2) bit for bit.  Fortran semantics kept:
 *   MAX(a,b) = a>b ? a : b ; MIN(a,b) = a<b ? a : b  (flang lowering)
 *   SIGN(1d0,x) = copysign(1,x) ; default-real literals are float.
 * Citations are reference file:line.
 *
 * Used only by the tests, the smoke check and the benchmark's CPU-baseline
 * leg, as the checker; the library never links it.
 *
 * Pinning (what fixes each function's results):
 *   - against the reference's own routines, compiled from their sources with
 *     flang into oracle/_ref, bit for bit: rec2 with PPM, MC, MP5 and WENO5;
 *     calcha with icha 0, 1 and 2; hll, ct, caltenr/calsf, the rk* and
 *     rk*ads* routines, recov1a and recov2, atmosphere1a, pminmax2b, caluu1,
 *     cdtcfl4, and the composed step in main.f90's order
 *     (tests/test_reference_live.py), and the golden vectors and 10- and
 *     100-step trajectories that build wrote (tests/test_oracle_golden.py);
 *   - bnd4 and infchk (bnd.f90 needs mpif.h and is not built): by brute force
 *     over every boundary code on tiny grids (tests/test_oracle_invariants.py);
 *   - and by mathematics where no reference output reaches: exactness of the
 *     reconstructions on linear data, HLLE consistency, the CT corner
 *     divergence, the recovery round trips, the static-gas CFL, the
 *     wave-speed identities, the Kerr-Schild identities.
 * No function here is left unpinned.
 */
#include "raishin_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

static orc_pram P;
static orc_counters CNT;
/* optional record of the cells whose initial W is inherited (diagnostics) */
static long long *DBG_CARRY;
static long long DBG_CAP, DBG_N;
void orc_debug_carry(long long *buf, long long cap) { DBG_CARRY = buf; DBG_CAP = cap; DBG_N = 0; }
long long orc_debug_carry_count(void) { return DBG_N; }
static int NI, NJ, NK;

#define CELL(i, j, k) ((size_t)((i)-1) + (size_t)NI * ((size_t)((j)-1) + (size_t)NJ * (size_t)((k)-1)))
#define V(a, n, i, j, k) (a)[((n)-1) + 9 * CELL(i, j, k)]
#define G16(a, i, j, k) (&(a)[16 * CELL(i, j, k)])
#define S3(a, i, j, k) (a)[CELL(i, j, k)]
#define WW(a, d, n, i, j, k) (a)[((d)-1) + 3 * (((n)-1) + 9 * CELL(i, j, k))]
#define SF(a, n, i, j, k) (a)[((n)-2) + 4 * CELL(i, j, k)]
#define GM(g, m, n) (g)[(m) + 4 * (n)]

static inline double FMAX(double a, double b) { return a > b ? a : b; }
static inline double FMIN(double a, double b) { return a < b ? a : b; }
static inline double FSIGN1(double x) { return copysign(1.0, x); }

void orc_set_pram(const orc_pram *p) {
  P = *p;
  NI = p->imax; NJ = p->jmax; NK = p->kmax;
}
void orc_get_counters(orc_counters *c) { *c = CNT; }
void orc_reset_counters(void) { memset(&CNT, 0, sizeof CNT); }

/* ------------------------------------------------------------------ */
/* physfunc.f90 pointwise helpers                                      */
/* ------------------------------------------------------------------ */

/* calgfl: physfunc.f90:2-41 */
static double calgfl(const double util[4], const double *g) {
  double utsq = 0.0;
  for (int m = 1; m <= 3; m++)
    for (int n = 1; n <= 3; n++) utsq = utsq + GM(g, m, n) * util[m] * util[n];
  if (utsq < 0.0) {
    if (fabs(utsq) <= 1.0e-10) utsq = fabs(utsq);
    else utsq = 1.0e-10;
  }
  if (fabs(utsq) > 5.0e2) utsq = 5.0e2;
  return sqrt(1.0 + utsq);
}

/* calucon: physfunc.f90:44-74 */
static void calucon(const double util[4], double ucon[4], double gfl, const double *gi) {
  double alpha1 = 1.0 / sqrt(-GM(gi, 0, 0));
  double b1 = alpha1 * alpha1 * GM(gi, 0, 1);
  double b2 = alpha1 * alpha1 * GM(gi, 0, 2);
  double b3 = alpha1 * alpha1 * GM(gi, 0, 3);
  ucon[0] = gfl * (1.0 / alpha1);
  ucon[1] = util[1] - gfl * b1 * (1.0 / alpha1);
  ucon[2] = util[2] - gfl * b2 * (1.0 / alpha1);
  ucon[3] = util[3] - gfl * b3 * (1.0 / alpha1);
}

/* calbbcon: physfunc.f90:77-95 */
static void calbbcon(const double bcon[4], double bbcon[4], const double ucov[4],
                     const double ucon[4]) {
  bbcon[0] = bcon[1] * ucov[1] + bcon[2] * ucov[2] + bcon[3] * ucov[3];
  bbcon[1] = (bcon[1] + bbcon[0] * ucon[1]) * (1.0 / ucon[0]);
  bbcon[2] = (bcon[2] + bbcon[0] * ucon[2]) * (1.0 / ucon[0]);
  bbcon[3] = (bcon[3] + bbcon[0] * ucon[3]) * (1.0 / ucon[0]);
}

/* calbbsq: physfunc.f90:98-114 */
static double calbbsq(const double bbcon[4], const double bbcov[4]) {
  return bbcon[0] * bbcov[0] + bbcon[1] * bbcov[1] + bbcon[2] * bbcov[2] + bbcon[3] * bbcov[3];
}

/* lower / upper: physfunc.f90:158-230 */
static void lower(const double v[4], double w[4], const double *g) {
  for (int m = 0; m < 4; m++)
    w[m] = GM(g, m, 0) * v[0] + GM(g, m, 1) * v[1] + GM(g, m, 2) * v[2] + GM(g, m, 3) * v[3];
}
static void upper(const double v[4], double w[4], const double *g) { lower(v, w, g); }

/* lower1: physfunc.f90:233-258 (3-metric, gcov3a(1:3,1:3) stored as g3[m-1 + 3(n-1)]) */
static void lower1(const double v[4], double w[4], const double *g3) {
  for (int m = 1; m <= 3; m++)
    w[m] = g3[(m - 1) + 0] * v[1] + g3[(m - 1) + 3] * v[2] + g3[(m - 1) + 6] * v[3];
}

/* cal3met: coord.f90:970-1013 */
static void cal3met(const double *g, const double *gi, double *g3, double *gi3) {
  double alpha1 = 1.0 / sqrt(-GM(gi, 0, 0));
  double beta1[4];
  beta1[1] = alpha1 * alpha1 * GM(gi, 0, 1);
  beta1[2] = alpha1 * alpha1 * GM(gi, 0, 2);
  beta1[3] = alpha1 * alpha1 * GM(gi, 0, 3);
  double oncon[4], oncov[4];
  oncon[0] = 1.0 / alpha1;
  oncon[1] = beta1[1] / alpha1;
  oncon[2] = beta1[2] / alpha1;
  oncon[3] = beta1[3] / alpha1;
  oncov[0] = -alpha1; oncov[1] = 0.0; oncov[2] = 0.0; oncov[3] = 0.0;
  for (int m = 1; m <= 3; m++)
    for (int n = 1; n <= 3; n++) {
      g3[(m - 1) + 3 * (n - 1)] = GM(g, m, n) + (oncov[m] * oncov[n]);
      gi3[(m - 1) + 3 * (n - 1)] = GM(gi, m, n) + (oncon[m] * oncon[n]);
    }
}

static inline double enthalpy(double de, double pr) {
  /* ieos=0: roh=de+(gam/(gam-1.0))*pr (caluu.f90:52) */
  return de + (P.gam / (P.gam - 1.0)) * pr;
}

/* Shared prelude of caluu1/caluu3/calflx3*: primitive -> u^mu, u_mu, b^mu, b_mu, b^2 */
typedef struct { double de, pr, roh, gfl, bbsq, util[4], bcon[4], ucon[4], ucov[4], bbcon[4], bbcov[4]; } prel_t;

static void prelude(const double *w, const double *g, const double *gi, prel_t *s) {
  s->de = w[0];
  s->util[1] = w[1]; s->util[2] = w[2]; s->util[3] = w[3];
  s->pr = w[4];
  s->bcon[1] = w[6]; s->bcon[2] = w[7]; s->bcon[3] = w[8];
  s->roh = enthalpy(s->de, s->pr);
  s->gfl = calgfl(s->util, g);
  calucon(s->util, s->ucon, s->gfl, gi);
  lower(s->ucon, s->ucov, g);
  calbbcon(s->bcon, s->bbcon, s->ucov, s->ucon);
  lower(s->bbcon, s->bbcov, g);
  s->bbsq = calbbsq(s->bbcon, s->bbcov);
}

/* conserved variables: caluu.f90:85-103 (d=0) and fluxes calflx.f90:468-507 (d=1..3) */
static void tmunu_row(const prel_t *s, int d, double detg, double out[9]) {
  const double de = s->de, pr = s->pr, roh = s->roh, bbsq = s->bbsq;
  double delta[4] = {0.0, 0.0, 0.0, 0.0};
  delta[d] = 1.0; /* caldelta: physfunc.f90:117-155 */
  out[0] = de * s->ucon[d];
  out[1] = (roh + bbsq) * s->ucon[d] * s->ucov[1] + (pr + 0.5 * bbsq) * delta[1] - s->bbcon[d] * s->bbcov[1];
  out[2] = (roh + bbsq) * s->ucon[d] * s->ucov[2] + (pr + 0.5 * bbsq) * delta[2] - s->bbcon[d] * s->bbcov[2];
  out[3] = (roh + bbsq) * s->ucon[d] * s->ucov[3] + (pr + 0.5 * bbsq) * delta[3] - s->bbcon[d] * s->bbcov[3];
  out[4] = (roh + bbsq) * s->ucon[d] * s->ucov[0] + (pr + 0.5 * bbsq) * delta[0] - s->bbcon[d] * s->bbcov[0] + de * s->ucon[d];
  out[5] = de * (pr / pow(de, P.gam)) * s->ucon[d];
  if (d == 0) {
    out[6] = s->bcon[1]; out[7] = s->bcon[2]; out[8] = s->bcon[3];
  } else {
    for (int c = 1; c <= 3; c++) {
      if (c == d) out[5 + c] = 0.0;
      else out[5 + c] = (s->ucon[d] * (1.0 / s->ucon[0])) * s->bcon[c] - (s->ucon[c] * (1.0 / s->ucon[0])) * s->bcon[d];
    }
  }
  for (int n = 0; n < 9; n++) out[n] = detg * out[n];
}

/* ------------------------------------------------------------------ */
/* caluu1: caluu.f90:2-126 (all cells)                                  */
/* ------------------------------------------------------------------ */
void orc_caluu1(const double *uri, double *uu, const double *gcov, const double *gcon,
                const double *detg) {
  for (int k = 1; k <= NK; k++)
    for (int j = 1; j <= NJ; j++)
      for (int i = 1; i <= NI; i++) {
        prel_t s;
        prelude(&V(uri, 1, i, j, k), G16(gcov, i, j, k), G16(gcon, i, j, k), &s);
        tmunu_row(&s, 0, S3(detg, i, j, k), &V(uu, 1, i, j, k));
      }
}

/* caluu3 (caluu.f90:236-356) / calflx3i,j,k (calflx.f90:391-795) on one face array */
static void face_cons_flux(const double *ur, double *uu, double *ww, const double *gcov,
                           const double *gcon, const double *detg, int d, int nm0) {
  for (int k = 1 + nm0 - 1; k <= NK - nm0; k++)
    for (int j = 1 + nm0 - 1; j <= NJ - nm0; j++)
      for (int i = 1 + nm0 - 1; i <= NI - nm0; i++) {
        prel_t s;
        prelude(&V(ur, 1, i, j, k), G16(gcov, i, j, k), G16(gcon, i, j, k), &s);
        tmunu_row(&s, 0, S3(detg, i, j, k), &V(uu, 1, i, j, k));
        tmunu_row(&s, d, S3(detg, i, j, k), &V(ww, 1, i, j, k));
      }
}

/* ------------------------------------------------------------------ */
/* PPM: rec.f90:1261-2056                                               */
/* dx*b are passed as the dummy dx1b(2:imax-1) of rec2 (rec.f90:58): the */
/* dummy element d is the actual element d-1; DXB(a,d) restates that.   */
/* d-1 == 0 reads one element before the array in the reference; that   */
/* value only reaches interfaces hll never reads, so it is clamped here. */
/* ------------------------------------------------------------------ */
#define DXB(a, d) ((a)[((d) - 1) >= 1 ? ((d) - 2) : 0])

static void ppm_interp(const double *uri, const double *dx1b, const double *dx2b, const double *dx3b,
                       double *uriir, double *urijr, double *urikr, double *uriil, double *urijl,
                       double *urikl, double *dmai, double *dmaj, double *dmak) {
  /* rec.f90:1340-1418 */
  for (int k = 2; k <= NK - 1; k++)
    for (int j = 2; j <= NJ - 1; j++)
      for (int i = 2; i <= NI - 1; i++)
        for (int n = 1; n <= 9; n++) {
          double t1a, t1b, t1c, sp1, sm1, spm1, da, t1, t1d;
          t1a = DXB(dx1b, i) / (DXB(dx1b, i - 1) + DXB(dx1b, i) + DXB(dx1b, i + 1));
          t1b = (2.0 * DXB(dx1b, i - 1) + DXB(dx1b, i)) / (DXB(dx1b, i + 1) + DXB(dx1b, i));
          t1c = (DXB(dx1b, i) + 2.0 * DXB(dx1b, i + 1)) / (DXB(dx1b, i - 1) + DXB(dx1b, i));
          sp1 = (V(uri, n, i + 1, j, k) - V(uri, n, i, j, k));
          sm1 = (V(uri, n, i, j, k) - V(uri, n, i - 1, j, k));
          spm1 = sp1 * sm1;
          da = t1a * (t1b * sp1 + t1c * sm1);
          if (spm1 <= 0.0) V(dmai, n, i, j, k) = 0.0;
          else {
            t1 = (da < 0.0) ? -1.0 : 1.0;
            t1d = FMIN(fabs(sm1), fabs(sp1));
            V(dmai, n, i, j, k) = t1 * FMIN(fabs(da), 2.0 * t1d);
          }
          t1a = DXB(dx2b, j) / (DXB(dx2b, j - 1) + DXB(dx2b, j) + DXB(dx2b, j + 1));
          t1b = (2.0 * DXB(dx2b, j - 1) + DXB(dx2b, j)) / (DXB(dx2b, j + 1) + DXB(dx2b, j));
          t1c = (DXB(dx2b, j) + 2.0 * DXB(dx2b, j + 1)) / (DXB(dx2b, j - 1) + DXB(dx2b, j));
          sp1 = (V(uri, n, i, j + 1, k) - V(uri, n, i, j, k));
          sm1 = (V(uri, n, i, j, k) - V(uri, n, i, j - 1, k));
          spm1 = sp1 * sm1;
          da = t1a * (t1b * sp1 + t1c * sm1);
          if (spm1 <= 0.0) V(dmaj, n, i, j, k) = 0.0;
          else {
            t1 = (da < 0.0) ? -1.0 : 1.0;
            t1d = FMIN(fabs(sm1), fabs(sp1));
            V(dmaj, n, i, j, k) = t1 * FMIN(fabs(da), 2.0 * t1d);
          }
          t1a = DXB(dx3b, k) / (DXB(dx3b, k - 1) + DXB(dx3b, k) + DXB(dx3b, k + 1));
          t1b = (2.0 * DXB(dx3b, k - 1) + DXB(dx3b, k)) / (DXB(dx3b, k + 1) + DXB(dx3b, k));
          t1c = (DXB(dx3b, k) + 2.0 * DXB(dx3b, k + 1)) / (DXB(dx3b, k - 1) + DXB(dx3b, k));
          sp1 = (V(uri, n, i, j, k + 1) - V(uri, n, i, j, k));
          sm1 = (V(uri, n, i, j, k) - V(uri, n, i, j, k - 1));
          spm1 = sp1 * sm1;
          da = t1a * (t1b * sp1 + t1c * sm1);
          if (spm1 <= 0.0) V(dmak, n, i, j, k) = 0.0;
          else {
            t1 = (da < 0.0) ? -1.0 : 1.0;
            t1d = FMIN(fabs(sm1), fabs(sp1));
            /* rec.f90:1414: no factor 2 in the k direction */
            V(dmak, n, i, j, k) = t1 * FMIN(fabs(da), t1d);
          }
        }
  /* rec.f90:1421-1473, equation (51) */
  for (int k = 2; k <= NK - 2; k++)
    for (int j = 2; j <= NJ - 2; j++)
      for (int i = 2; i <= NI - 2; i++)
        for (int n = 1; n <= 9; n++) {
#define EQ51(dxb, c, UP, DM, OUTR, OUTL)                                                              \
  {                                                                                                  \
    double t2 = (DXB(dxb, c) / (DXB(dxb, c) + DXB(dxb, c + 1))) * (UP - V(uri, n, i, j, k));           \
    double t3a = 1.0 / (DXB(dxb, c - 1) + DXB(dxb, c) + DXB(dxb, c + 1) + DXB(dxb, c + 2));            \
    double t3b = (2.0 * DXB(dxb, c + 1) * DXB(dxb, c)) / (DXB(dxb, c) + DXB(dxb, c + 1));              \
    double t3c = ((DXB(dxb, c - 1) + DXB(dxb, c)) / (2.0 * DXB(dxb, c) + DXB(dxb, c + 1))) -           \
                 ((DXB(dxb, c + 2) + DXB(dxb, c + 1)) / (2.0 * DXB(dxb, c + 1) + DXB(dxb, c)));        \
    double t3 = t3a * t3b * t3c * (UP - V(uri, n, i, j, k));                                          \
    double t4 = t3a * DXB(dxb, c) * ((DXB(dxb, c - 1) + DXB(dxb, c)) / (2.0 * DXB(dxb, c) + DXB(dxb, c + 1))); \
    double t5 = t3a * DXB(dxb, c + 1) * ((DXB(dxb, c + 1) + DXB(dxb, c + 2)) / (DXB(dxb, c) + 2.0 * DXB(dxb, c + 1))); \
    OUTR = V(uri, n, i, j, k) + t2 + t3 - t4 * DM + t5 * V(dma_c, n, i, j, k);                        \
    OUTL = OUTR;                                                                                     \
  }
          {
            const double *dma_c = dmai;
            EQ51(dx1b, i, V(uri, n, i + 1, j, k), V(dmai, n, i + 1, j, k), V(uriir, n, i, j, k), V(uriil, n, i, j, k));
          }
          {
            const double *dma_c = dmaj;
            EQ51(dx2b, j, V(uri, n, i, j + 1, k), V(dmaj, n, i, j + 1, k), V(urijr, n, i, j, k), V(urijl, n, i, j, k));
          }
          {
            const double *dma_c = dmak;
            EQ51(dx3b, k, V(uri, n, i, j, k + 1), V(dmak, n, i, j, k + 1), V(urikr, n, i, j, k), V(urikl, n, i, j, k));
          }
#undef EQ51
        }
}

static void ppm_detect(const double *uri, const double *x1, const double *x2, const double *x3,
                       const double *dx1b, const double *dx2b, const double *dx3b,
                       double *uriir, double *urijr, double *urikr, double *uriil, double *urijl,
                       double *urikl, const double *dmai, const double *dmaj, const double *dmak,
                       double *dda3, double *eta3) {
  /* rec.f90:1479-1686 */
  const double eta1 = 5.0, eta2 = 0.05, eps1 = 0.1, ak0 = 1.0;
  const size_t N = (size_t)NI * NJ * NK;
  double *ddai = dda3, *ddaj = dda3 + N, *ddak = dda3 + 2 * N;
  double *etai = eta3, *etaj = eta3 + N, *etak = eta3 + 2 * N;
  for (int k = 2; k <= NK - 1; k++)
    for (int j = 2; j <= NJ - 1; j++)
      for (int i = 2; i <= NI - 1; i++) {
        double a, b, c;
        a = 1.0 / (DXB(dx1b, i - 1) + DXB(dx1b, i) + DXB(dx1b, i + 1));
        b = (V(uri, 1, i + 1, j, k) - V(uri, 1, i, j, k)) / (DXB(dx1b, i + 1) + DXB(dx1b, i));
        c = (V(uri, 1, i, j, k) - V(uri, 1, i - 1, j, k)) / (DXB(dx1b, i) + DXB(dx1b, i - 1));
        S3(ddai, i, j, k) = a * (b - c);
        a = 1.0 / (DXB(dx2b, j - 1) + DXB(dx2b, j) + DXB(dx2b, j + 1));
        b = (V(uri, 1, i, j + 1, k) - V(uri, 1, i, j, k)) / (DXB(dx2b, j + 1) + DXB(dx2b, j));
        c = (V(uri, 1, i, j, k) - V(uri, 1, i, j - 1, k)) / (DXB(dx2b, j) + DXB(dx2b, j - 1));
        S3(ddaj, i, j, k) = a * (b - c);
        a = 1.0 / (DXB(dx3b, k - 1) + DXB(dx3b, k) + DXB(dx3b, k + 1));
        b = (V(uri, 1, i, j, k + 1) - V(uri, 1, i, j, k)) / (DXB(dx3b, k + 1) + DXB(dx3b, k));
        c = (V(uri, 1, i, j, k) - V(uri, 1, i, j, k - 1)) / (DXB(dx3b, k) + DXB(dx3b, k - 1));
        S3(ddak, i, j, k) = a * (b - c);
      }
  for (int k = 3; k <= NK - 2; k++)
    for (int j = 3; j <= NJ - 2; j++)
      for (int i = 3; i <= NI - 2; i++) {
#define ETA(ddA, xs, cm, c0, cp, UP, UM, OUT)                                                      \
  {                                                                                               \
    double t1 = ddA##P * ddA##M;                                                                  \
    double t2 = fabs(UP - UM) - eps1 * FMIN(fabs(UP), fabs(UM));                                  \
    double etatil;                                                                                \
    if (t1 >= 0.0 && t2 <= 0.0) etatil = 0.0;                                                     \
    else {                                                                                        \
      double t2a = -(ddA##P - ddA##M) / (xs[cp] - xs[cm]);                                        \
      double dl = xs[c0] - xs[cm], dr = xs[cp] - xs[c0];                                          \
      double t2b = ((dl * dl * dl) + (dr * dr * dr)) / (UP - UM);                                 \
      etatil = t2a * t2b;                                                                         \
    }                                                                                             \
    OUT = FMAX(0.0, FMIN(eta1 * (etatil - eta2), 1.0));                                           \
  }
        double ddiP = S3(ddai, i + 1, j, k), ddiM = S3(ddai, i - 1, j, k);
        double ddjP = S3(ddaj, i, j + 1, k), ddjM = S3(ddaj, i, j - 1, k);
        double ddkP = S3(ddak, i, j, k + 1), ddkM = S3(ddak, i, j, k - 1);
        ETA(ddi, x1, i - 2, i - 1, i, V(uri, 1, i + 1, j, k), V(uri, 1, i - 1, j, k), S3(etai, i, j, k));
        ETA(ddj, x2, j - 2, j - 1, j, V(uri, 1, i, j + 1, k), V(uri, 1, i, j - 1, k), S3(etaj, i, j, k));
        ETA(ddk, x3, k - 2, k - 1, k, V(uri, 1, i, j, k + 1), V(uri, 1, i, j, k - 1), S3(etak, i, j, k));
#undef ETA
      }
  for (int k = 3; k <= NK - 2; k++)
    for (int j = 3; j <= NJ - 2; j++)
      for (int i = 3; i <= NI - 2; i++) {
        /* equation (63): rec.f90:1599-1629 */
        double t3, t4;
        t3 = fabs(V(uri, 1, i + 1, j, k) - V(uri, 1, i - 1, j, k)) / FMIN(V(uri, 1, i + 1, j, k), V(uri, 1, i - 1, j, k));
        t4 = fabs(V(uri, 5, i + 1, j, k) - V(uri, 5, i - 1, j, k)) / FMIN(V(uri, 5, i + 1, j, k), V(uri, 5, i - 1, j, k));
        if (P.gam * ak0 * t3 < t4) S3(etai, i, j, k) = 0.0;
        t3 = fabs(V(uri, 1, i, j + 1, k) - V(uri, 1, i, j - 1, k)) / FMIN(V(uri, 1, i, j + 1, k), V(uri, 1, i, j - 1, k));
        t4 = fabs(V(uri, 5, i, j + 1, k) - V(uri, 5, i, j - 1, k)) / FMIN(V(uri, 5, i, j + 1, k), V(uri, 5, i, j - 1, k));
        if (P.gam * ak0 * t3 < t4) S3(etaj, i, j, k) = 0.0;
        t3 = fabs(V(uri, 1, i, j, k + 1) - V(uri, 1, i, j, k - 1)) / FMIN(V(uri, 1, i, j, k + 1), V(uri, 1, i, j, k - 1));
        t4 = fabs(V(uri, 5, i, j, k + 1) - V(uri, 5, i, j, k - 1)) / FMIN(V(uri, 5, i, j, k + 1), V(uri, 5, i, j, k - 1));
        if (P.gam * ak0 * t3 < t4) S3(etak, i, j, k) = 0.0;
      }
  for (int k = 3; k <= NK - 2; k++)
    for (int j = 3; j <= NJ - 2; j++)
      for (int i = 3; i <= NI - 2; i++) {
        /* equations (64), (65): rec.f90:1632-1677 */
        double ei = S3(etai, i, j, k), ej = S3(etaj, i, j, k), ek = S3(etak, i, j, k);
        V(uriil, 1, i, j, k) = V(uriil, 1, i, j, k) * (1.0 - ei) + (V(uri, 1, i + 1, j, k) - 0.5 * V(dmai, 1, i + 1, j, k)) * ei;
        V(urijl, 1, i, j, k) = V(urijl, 1, i, j, k) * (1.0 - ej) + (V(uri, 1, i, j + 1, k) - 0.5 * V(dmaj, 1, i, j + 1, k)) * ej;
        V(urikl, 1, i, j, k) = V(urikl, 1, i, j, k) * (1.0 - ek) + (V(uri, 1, i, j, k + 1) - 0.5 * V(dmak, 1, i, j, k + 1)) * ek;
        V(uriir, 1, i - 1, j, k) = V(uriir, 1, i - 1, j, k) * (1.0 - ei) + (V(uri, 1, i - 1, j, k) + 0.5 * V(dmai, 1, i - 1, j, k)) * ei;
        V(urijr, 1, i, j - 1, k) = V(urijr, 1, i, j - 1, k) * (1.0 - ej) + (V(uri, 1, i, j - 1, k) + 0.5 * V(dmaj, 1, i, j - 1, k)) * ej;
        V(urikr, 1, i, j, k - 1) = V(urikr, 1, i, j, k - 1) * (1.0 - ek) + (V(uri, 1, i, j, k - 1) + 0.5 * V(dmak, 1, i, j, k - 1)) * ek;
      }
}

static void ppm_flaten(const double *uri, double *uriir, double *urijr, double *urikr,
                       double *uriil, double *urijl, double *urikl, double *f1a3, double *f2a3) {
  /* rec.f90:1689-1884 */
  const double eps2 = 1.0, omg2 = 0.52, omg3 = 10.0;
  const size_t N = (size_t)NI * NJ * NK;
  double *f1ai = f1a3, *f1aj = f1a3 + N, *f1ak = f1a3 + 2 * N;
  double *f2ai = f2a3, *f2aj = f2a3 + N, *f2ak = f2a3 + 2 * N;
  for (int k = 3; k <= NK - 2; k++)
    for (int j = 3; j <= NJ - 2; j++)
      for (int i = 3; i <= NI - 2; i++) {
#define F1(P1, M1, P2, M2, VP, VM, OUT)                                   \
  {                                                                       \
    double t3a = fabs(P1 - M1);                                           \
    double t1 = eps2 * FMIN(P1, M1) - t3a;                                \
    double t2 = VP - VM;                                                  \
    double omg1 = (t1 < 0.0 && t2 < 0.0) ? 1.0 : 0.0;                     \
    double t3b = (P2 - M2);                                               \
    double t4;                                                            \
    if (t3b == 0.0) t4 = (t3a == 0.0) ? -omg2 : 1.0 - omg2;               \
    else t4 = (t3a / t3b) - omg2;                                         \
    OUT = FMIN(1.0, omg1 * FMAX(0.0, t4 * omg3));                         \
  }
        F1(V(uri, 5, i + 1, j, k), V(uri, 5, i - 1, j, k), V(uri, 5, i + 2, j, k), V(uri, 5, i - 2, j, k),
           V(uri, 2, i + 1, j, k), V(uri, 2, i - 1, j, k), S3(f1ai, i, j, k));
        F1(V(uri, 5, i, j + 1, k), V(uri, 5, i, j - 1, k), V(uri, 5, i, j + 2, k), V(uri, 5, i, j - 2, k),
           V(uri, 3, i, j + 1, k), V(uri, 3, i, j - 1, k), S3(f1aj, i, j, k));
        F1(V(uri, 5, i, j, k + 1), V(uri, 5, i, j, k - 1), V(uri, 5, i, j, k + 2), V(uri, 5, i, j, k - 2),
           V(uri, 4, i, j, k + 1), V(uri, 4, i, j, k - 1), S3(f1ak, i, j, k));
#undef F1
      }
  for (int k = 4; k <= NK - 3; k++)
    for (int j = 4; j <= NJ - 3; j++)
      for (int i = 4; i <= NI - 3; i++) {
        double f1b;
        f1b = (V(uri, 5, i + 1, j, k) - V(uri, 5, i - 1, j, k) < 0.0) ? S3(f1ai, i + 1, j, k) : S3(f1ai, i - 1, j, k);
        S3(f2ai, i, j, k) = FMAX(S3(f1ai, i, j, k), f1b);
        f1b = (V(uri, 5, i, j + 1, k) - V(uri, 5, i, j - 1, k) < 0.0) ? S3(f1aj, i, j + 1, k) : S3(f1aj, i, j - 1, k);
        S3(f2aj, i, j, k) = FMAX(S3(f1aj, i, j, k), f1b);
        f1b = (V(uri, 5, i, j, k + 1) - V(uri, 5, i, j, k - 1) < 0.0) ? S3(f1ak, i, j, k + 1) : S3(f1ak, i, j, k - 1);
        S3(f2ak, i, j, k) = FMAX(S3(f1ak, i, j, k), f1b);
      }
  for (int k = 4; k <= NK - 3; k++)
    for (int j = 4; j <= NJ - 3; j++)
      for (int i = 4; i <= NI - 3; i++)
        for (int n = 1; n <= 9; n++) {
          double u = V(uri, n, i, j, k);
          double fi = S3(f2ai, i, j, k), fj = S3(f2aj, i, j, k), fk = S3(f2ak, i, j, k);
          V(uriil, n, i, j, k) = u * fi + V(uriil, n, i, j, k) * (1.0 - fi);
          V(uriir, n, i - 1, j, k) = u * fi + V(uriir, n, i - 1, j, k) * (1.0 - fi);
          V(urijl, n, i, j, k) = u * fj + V(urijl, n, i, j, k) * (1.0 - fj);
          V(urijr, n, i, j - 1, k) = u * fj + V(urijr, n, i, j - 1, k) * (1.0 - fj);
          V(urikl, n, i, j, k) = u * fk + V(urikl, n, i, j, k) * (1.0 - fk);
          V(urikr, n, i, j, k - 1) = u * fk + V(urikr, n, i, j, k - 1) * (1.0 - fk);
        }
}

/* monoto on one cell edge pair (L = left edge a_L = uri*r(c-1), R = right edge a_R = uri*l(c)) */
static inline void monoto_pair(double u, double *aR, double *aL) {
  /* equation (73): rec.f90:1920-1940 */
  if ((*aR - u) * (*aL - u) >= 0.0) { *aR = u; *aL = u; }
}
static inline void monoto_pair2(double u, double *aR, double *aL) {
  /* equations (74), (75): rec.f90:1991-2050; the uri*1 arrays of 1953-1988 are never read */
  double t3 = 6.0 * (*aR - *aL) * (u - 0.5 * (*aR + *aL));
  double t4 = (*aR - *aL) * (*aR - *aL);
  if (t3 > t4) *aL = 3.0 * u - 2.0 * *aR;
  if (-t3 > t4) *aR = 3.0 * u - 2.0 * *aL;
}

static void ppm_monoto(const double *uri, double *uriir, double *urijr, double *urikr,
                       double *uriil, double *urijl, double *urikl) {
  for (int k = 3; k <= NK - 2; k++)
    for (int j = 3; j <= NJ - 2; j++)
      for (int i = 3; i <= NI - 2; i++)
        for (int n = 1; n <= 9; n++) {
          double u = V(uri, n, i, j, k);
          monoto_pair(u, &V(uriil, n, i, j, k), &V(uriir, n, i - 1, j, k));
          monoto_pair(u, &V(urijl, n, i, j, k), &V(urijr, n, i, j - 1, k));
          monoto_pair(u, &V(urikl, n, i, j, k), &V(urikr, n, i, j, k - 1));
        }
  for (int k = 3; k <= NK - 2; k++)
    for (int j = 3; j <= NJ - 2; j++)
      for (int i = 3; i <= NI - 2; i++)
        for (int n = 1; n <= 9; n++) {
          double u = V(uri, n, i, j, k);
          monoto_pair2(u, &V(uriil, n, i, j, k), &V(uriir, n, i - 1, j, k));
          monoto_pair2(u, &V(urijl, n, i, j, k), &V(urijr, n, i, j - 1, k));
          monoto_pair2(u, &V(urikl, n, i, j, k), &V(urikr, n, i, j, k - 1));
        }
}

/* ------------------------------------------------------------------ */
/* MP5: rec.f90:2059-2370                                               */
/* ------------------------------------------------------------------ */
static inline double dmm(double x, double y) {
  return 0.5 * (FSIGN1(x) + FSIGN1(y)) * FMIN(fabs(x), fabs(y));
}
static inline double dm4(double w, double x, double y, double z) {
  return 0.125 * (FSIGN1(w) + FSIGN1(x)) * fabs(FSIGN1(w) + FSIGN1(y)) * (FSIGN1(w) + FSIGN1(z)) *
         FMIN(FMIN(FMIN(fabs(w), fabs(x)), fabs(y)), fabs(z));
}

/* s[0..4] = u(c-2..c+2).  Right edge (uri*l(c)): rec.f90:2113-2145 */
static double mp5_right(const double s[5]) {
  const double tmp1 = 1.0 / 60.0, tmp2 = 4.0 / 3.0, alpha = 4.0, eps1 = 1.0e-10;
  double u1 = tmp1 * (2.0 * s[0] - 13.0 * s[1] + 47.0 * s[2] + 27.0 * s[3] - 3.0 * s[4]);
  double ump = s[2] + dmm(s[3] - s[2], alpha * (s[2] - s[1]));
  if ((u1 - s[2]) * (u1 - ump) <= eps1) return u1;
  double dim1 = s[0] - 2.0 * s[1] + s[2];
  double di1 = s[1] - 2.0 * s[2] + s[3];
  double dip1 = s[2] - 2.0 * s[3] + s[4];
  double dm4iph = dm4(4.0 * di1 - dip1, 4.0 * dip1 - di1, di1, dip1);
  double dm4imh = dm4(4.0 * di1 - dim1, 4.0 * dim1 - di1, di1, dim1);
  double ul = s[2] + alpha * (s[2] - s[1]);
  double av = 0.5 * (s[2] + s[3]);
  double md = av - 0.5 * dm4iph;
  double lc = s[2] + 0.5 * (s[2] - s[1]) + tmp2 * dm4imh;
  double m1 = FMIN(FMIN(s[2], s[3]), md);
  double m2 = FMIN(FMIN(s[2], ul), lc);
  double m3 = FMAX(FMAX(s[2], s[3]), md);
  double m4 = FMAX(FMAX(s[2], ul), lc);
  double umax = FMAX(m1, m2);
  double umin = FMIN(m3, m4);
  return u1 + dmm(umin - u1, umax - u1);
}
/* Left edge (uri*r(c-1)): rec.f90:2149-2186 */
static double mp5_left(const double s[5]) {
  const double tmp1 = 1.0 / 60.0, tmp2 = 4.0 / 3.0, alpha = 4.0, eps1 = 1.0e-10;
  double u1 = tmp1 * (2.0 * s[4] - 13.0 * s[3] + 47.0 * s[2] + 27.0 * s[1] - 3.0 * s[0]);
  double ump = s[2] + dmm(s[1] - s[2], alpha * (s[2] - s[3]));
  if ((u1 - s[2]) * (u1 - ump) <= eps1) return u1;
  double dim1 = s[0] - 2.0 * s[1] + s[2];
  double di1 = s[1] - 2.0 * s[2] + s[3];
  double dip1 = s[2] - 2.0 * s[3] + s[4];
  double dm4iph = dm4(4.0 * di1 - dip1, 4.0 * dip1 - di1, di1, dip1);
  double dm4imh = dm4(4.0 * di1 - dim1, 4.0 * dim1 - di1, di1, dim1);
  double ul = s[2] + alpha * (s[2] - s[3]);
  double av = 0.5 * (s[2] + s[1]);
  double md = av + 0.5 * dm4imh;
  double lc = s[2] + 0.5 * (s[2] - s[3]) - tmp2 * dm4iph;
  double m1 = FMIN(FMIN(s[2], s[1]), md);
  double m2 = FMIN(FMIN(s[2], ul), lc);
  double m3 = FMAX(FMAX(s[2], s[1]), md);
  double m4 = FMAX(FMAX(s[2], ul), lc);
  double umax = FMAX(m1, m2);
  double umin = FMIN(m3, m4);
  return u1 + dmm(umin - u1, umax - u1);
}

/* ------------------------------------------------------------------ */
/* WENO5: rec.f90:2373-2642 (13./12. is a single-precision literal)     */
/* ------------------------------------------------------------------ */
static double weno_right(const double s[5]) {
  const double c1312 = (double)(13.0f / 12.0f), c14 = (double)(1.0f / 4.0f), epsi = 1.0e-6;
  const double cd0 = 1.0 / 10.0, cd1 = 6.0 / 10.0, cd2 = 3.0 / 10.0;
  const double c00 = 2.0 / 6.0, c01 = -7.0 / 6.0, c02 = 11.0 / 6.0, c10 = -1.0 / 6.0, c11 = 5.0 / 6.0,
               c12 = 2.0 / 6.0, c20 = 2.0 / 6.0, c21 = 5.0 / 6.0, c22 = -1.0 / 6.0;
  double a, b;
  a = s[0] - 2.0 * s[1] + s[2]; b = s[0] - 4.0 * s[1] + 3.0 * s[2];
  double b0 = c1312 * (a * a) + c14 * (b * b) + epsi;
  a = s[1] - 2.0 * s[2] + s[3]; b = s[1] - s[3];
  double b1 = c1312 * (a * a) + c14 * (b * b) + epsi;
  a = s[2] - 2.0 * s[3] + s[4]; b = 3.0 * s[2] - 4.0 * s[3] + s[4];
  double b2 = c1312 * (a * a) + c14 * (b * b) + epsi;
  double w0 = cd0 / (b0 * b0), w1 = cd1 / (b1 * b1), w2 = cd2 / (b2 * b2);
  double wt = w0 + w1 + w2;
  double o0 = w0 / wt, o1 = w1 / wt, o2 = w2 / wt;
  double r0 = c00 * s[0] + c01 * s[1] + c02 * s[2];
  double r1 = c10 * s[1] + c11 * s[2] + c12 * s[3];
  double r2 = c20 * s[2] + c21 * s[3] + c22 * s[4];
  return o0 * r0 + o1 * r1 + o2 * r2;
}
static double weno_left(const double s[5]) {
  const double c1312 = (double)(13.0f / 12.0f), c14 = (double)(1.0f / 4.0f), epsi = 1.0e-6;
  const double cd0 = 1.0 / 10.0, cd1 = 6.0 / 10.0, cd2 = 3.0 / 10.0;
  const double c00 = 2.0 / 6.0, c01 = -7.0 / 6.0, c02 = 11.0 / 6.0, c10 = -1.0 / 6.0, c11 = 5.0 / 6.0,
               c12 = 2.0 / 6.0, c20 = 2.0 / 6.0, c21 = 5.0 / 6.0, c22 = -1.0 / 6.0;
  double a, b;
  a = s[4] - 2.0 * s[3] + s[2]; b = s[4] - 4.0 * s[3] + 3.0 * s[2];
  double b0 = c1312 * (a * a) + c14 * (b * b) + epsi;
  a = s[3] - 2.0 * s[2] + s[1]; b = s[3] - s[1];
  double b1 = c1312 * (a * a) + c14 * (b * b) + epsi;
  a = s[2] - 2.0 * s[1] + s[0]; b = 3.0 * s[2] - 4.0 * s[1] + s[0];
  double b2 = c1312 * (a * a) + c14 * (b * b) + epsi;
  double w0 = cd0 / (b0 * b0), w1 = cd1 / (b1 * b1), w2 = cd2 / (b2 * b2);
  double wt = w0 + w1 + w2;
  double o0 = w0 / wt, o1 = w1 / wt, o2 = w2 / wt;
  double r0 = c00 * s[4] + c01 * s[3] + c02 * s[2];
  double r1 = c10 * s[3] + c11 * s[2] + c12 * s[1];
  double r2 = c20 * s[2] + c21 * s[1] + c22 * s[0];
  return o0 * r0 + o1 * r1 + o2 * r2;
}

static void five_point(const double *uri, double *uriir, double *urijr, double *urikr,
                       double *uriil, double *urijl, double *urikl, int scheme) {
  for (int k = 3; k <= NK - 2; k++)
    for (int j = 3; j <= NJ - 2; j++)
      for (int i = 3; i <= NI - 2; i++)
        for (int n = 1; n <= 9; n++) {
          double si[5], sj[5], sk[5];
          for (int q = 0; q < 5; q++) {
            si[q] = V(uri, n, i - 2 + q, j, k);
            sj[q] = V(uri, n, i, j - 2 + q, k);
            sk[q] = V(uri, n, i, j, k - 2 + q);
          }
          if (scheme == 6) {
            V(uriil, n, i, j, k) = mp5_right(si); V(uriir, n, i - 1, j, k) = mp5_left(si);
            V(urijl, n, i, j, k) = mp5_right(sj); V(urijr, n, i, j - 1, k) = mp5_left(sj);
            V(urikl, n, i, j, k) = mp5_right(sk); V(urikr, n, i, j, k - 1) = mp5_left(sk);
          } else {
            V(uriil, n, i, j, k) = weno_right(si); V(uriir, n, i - 1, j, k) = weno_left(si);
            V(urijl, n, i, j, k) = weno_right(sj); V(urijr, n, i, j - 1, k) = weno_left(sj);
            V(urikl, n, i, j, k) = weno_right(sk); V(urikr, n, i, j, k - 1) = weno_left(sk);
          }
        }
}

/* ------------------------------------------------------------------ */
/* mclim: rec.f90:208-367 (irec=1), MC-limited linear reconstruction.  */
/* Needs the cell-face positions x1a..x3a (orc_set_faces).             */
/* ------------------------------------------------------------------ */
static const double *FX[3] = {0, 0, 0};

void orc_set_faces(const double *x1a, const double *x2a, const double *x3a) {
  FX[0] = x1a; FX[1] = x2a; FX[2] = x3a;
}

static void mclim(const double *uri, const double *x1, const double *x2, const double *x3,
                  double *uriir, double *urijr, double *urikr, double *uriil, double *urijl, double *urikl) {
  const double small = 1.0e-12;
  const double *xc_[3] = {x1, x2, x3};
  double *lft[3] = {uriil, urijl, urikl}, *rgt[3] = {uriir, urijr, urikr};
  for (int k = 2; k <= NK - 1; k++)
    for (int j = 2; j <= NJ - 1; j++)
      for (int i = 2; i <= NI - 1; i++)
        for (int n = 1; n <= 9; n++) {
          const int id[3] = {i, j, k};
          for (int d = 0; d < 3; d++) {
            const int p = id[d];
            const double *x = xc_[d], *xa = FX[d];
            const double xp = x[p], xc = x[p - 1], xm = x[p - 2], xap = xa[p - 1], xam = xa[p - 2];
            const int dp[3] = {d == 0, d == 1, d == 2};
            const double u = V(uri, n, i, j, k);
            const double up = V(uri, n, i + dp[0], j + dp[1], k + dp[2]);
            const double um = V(uri, n, i - dp[0], j - dp[1], k - dp[2]);
            const double sp = (up - u) * (1.0 / (xp - xc + small));
            const double sm = (u - um) * (1.0 / (xc - xm + small));
            const double tmp = (sp < 0.0) ? -1.0 : 1.0;
            const double dk = tmp * FMAX(0.0, FMIN(FMIN(2.0 * fabs(sp), tmp * 2.0 * sm), tmp * 0.5 * (sp + sm)));
            V(lft[d], n, i, j, k) = u + dk * (xap - xc);
            V(rgt[d], n, i - dp[0], j - dp[1], k - dp[2]) = u - dk * (xc - xam);
          }
        }
}

/* ------------------------------------------------------------------ */
/* rec2: rec.f90:2-142                                                 */
/* ------------------------------------------------------------------ */
void orc_rec2(const double *uri, const double *x1, const double *x2, const double *x3,
              const double *dx1b, const double *dx2b, const double *dx3b,
              double *uriir, double *urijr, double *urikr,
              double *uriil, double *urijl, double *urikl,
              double *uuir, double *uujr, double *uukr,
              double *uuil, double *uujl, double *uukl,
              double *wwir, double *wwjr, double *wwkr,
              double *wwil, double *wwjl, double *wwkl,
              const double *gcovi, const double *gcovj, const double *gcovk,
              const double *gconi, const double *gconj, const double *gconk,
              const double *detgi, const double *detgj, const double *detgk, int nm0) {
  const size_t N = (size_t)NI * NJ * NK;
  if (P.irec == 5) {
    double *dma = (double *)malloc(sizeof(double) * 27 * N);
    double *tmp = (double *)malloc(sizeof(double) * 6 * N);
    ppm_interp(uri, dx1b, dx2b, dx3b, uriir, urijr, urikr, uriil, urijl, urikl, dma, dma + 9 * N, dma + 18 * N);
    ppm_detect(uri, x1, x2, x3, dx1b, dx2b, dx3b, uriir, urijr, urikr, uriil, urijl, urikl, dma,
               dma + 9 * N, dma + 18 * N, tmp, tmp + 3 * N);
    ppm_flaten(uri, uriir, urijr, urikr, uriil, urijl, urikl, tmp, tmp + 3 * N);
    ppm_monoto(uri, uriir, urijr, urikr, uriil, urijl, urikl);
    free(dma); free(tmp);
  } else if (P.irec == 6 || P.irec == 7) {
    five_point(uri, uriir, urijr, urikr, uriil, urijl, urikl, P.irec);
  } else if (P.irec == 1 && FX[0]) {
    mclim(uri, x1, x2, x3, uriir, urijr, urikr, uriil, urijl, urikl);
  } else {
    abort();
  }
  face_cons_flux(uriir, uuir, wwir, gcovi, gconi, detgi, 1, nm0);
  face_cons_flux(urijr, uujr, wwjr, gcovj, gconj, detgj, 2, nm0);
  face_cons_flux(urikr, uukr, wwkr, gcovk, gconk, detgk, 3, nm0);
  face_cons_flux(uriil, uuil, wwil, gcovi, gconi, detgi, 1, nm0);
  face_cons_flux(urijl, uujl, wwjl, gcovj, gconj, detgj, 2, nm0);
  face_cons_flux(urikl, uukl, wwkl, gcovk, gconk, detgk, 3, nm0);
}

/* ------------------------------------------------------------------ */
/* calcha4i/j/k: calcha.f90:370-906 (face speeds, icha 0/1/2)           */
/* ------------------------------------------------------------------ */
/* The speed pair of direction d before the lapse/shift correction:
   icha 0 calcha.f90:489-494 (Gammie), 1 :496-509 (Leismann), 2 :511-527
   (Del Zanna); gii = gcon3a(d,d), vb = vcov.bcon (calcha.f90:478). */
static void speed_pair(double vd, double vsq, double omsq, double cssq, double vb, double rohb,
                       double gfl, double gii, double *p, double *m) {
  if (P.icha == 0) {
    *p = vd + sqrt(omsq);
    *m = vd - sqrt(omsq);
    return;
  }
  double f1 = vd * (1.0 - omsq), f2, f3;
  if (P.icha == 1) {
    double r1 = cssq * (vb * vb) * (1.0 / (rohb * (gfl * gfl)));
    f2 = 1.0 - vsq * omsq - r1;
    f3 = ((vsq - 1.0) * omsq + r1) * ((gii * vsq - vd * vd) * omsq + vd * vd - gii + r1);
  } else {
    f2 = 1.0 - vsq * omsq;
    f3 = omsq * (1.0 - vsq) * ((1.0 - vsq * omsq) * gii - (1.0 - omsq) * vd * vd);
  }
  *p = (f1 + sqrt(f3)) * (1.0 / f2);
  *m = (f1 - sqrt(f3)) * (1.0 / f2);
}

static void face_speed(const double *w, const double *g, const double *gi, int d,
                       double *vfp, double *vfm) {
  prel_t s;
  prelude(w, g, gi, &s);
  double alpha1 = 1.0 / sqrt(-GM(gi, 0, 0));
  double beta_d = alpha1 * alpha1 * GM(gi, 0, d);
  double g3[9], gi3[9];
  cal3met(g, gi, g3, gi3);
  double vcon[4], vcov[4];
  vcon[1] = s.util[1] * (1.0 / s.gfl);
  vcon[2] = s.util[2] * (1.0 / s.gfl);
  vcon[3] = s.util[3] * (1.0 / s.gfl);
  lower1(vcon, vcov, g3);
  double vsq = 0.0;
  for (int m = 1; m <= 3; m++)
    for (int n = 1; n <= 3; n++) vsq = vsq + g3[(m - 1) + 3 * (n - 1)] * vcon[m] * vcon[n];
  double cssq, vasq;
  /* calcha.f90:480,648,826 and :486,654,832: the three directions write these differently */
  if (d == 3) cssq = P.gam * s.pr * (1.0 / s.roh);
  else cssq = P.gam * s.pr / s.roh;
  if (d == 2) vasq = s.bbsq / (s.roh + s.bbsq);
  else vasq = s.bbsq * (1.0 / (s.roh + s.bbsq));
  double omsq = vasq + cssq * (1.0 - vasq);
  double vb = vcov[1] * s.bcon[1] + vcov[2] * s.bcon[2] + vcov[3] * s.bcon[3];
  double p, m;
  speed_pair(vcon[d], vsq, omsq, cssq, vb, s.roh + s.bbsq, s.gfl, gi3[(d - 1) * 4], &p, &m);
  p = alpha1 * p - beta_d;
  m = alpha1 * m - beta_d;
  if (fabs(p) > 1.0) p = 1.0;
  if (fabs(m) > 1.0) m = -1.0;
  *vfp = p; *vfm = m;
}

/* calcha: calcha.f90:2-126 */
void orc_calcha(const double *uriir, const double *urijr, const double *urikr,
                const double *uriil, const double *urijl, const double *urikl,
                const double *gcovi, const double *gcovj, const double *gcovk,
                const double *gconi, const double *gconj, const double *gconk,
                double *cmaxi, double *cmini, double *cmaxj, double *cminj,
                double *cmaxk, double *cmink, int nm0) {
  if (P.icha < 0 || P.icha > 2) abort();
  for (int k = nm0; k <= NK - nm0; k++)
    for (int j = nm0; j <= NJ - nm0; j++)
      for (int i = nm0; i <= NI - nm0; i++) {
        double pr_, mr_, pl_, ml_;
        face_speed(&V(uriir, 1, i, j, k), G16(gcovi, i, j, k), G16(gconi, i, j, k), 1, &pr_, &mr_);
        face_speed(&V(uriil, 1, i, j, k), G16(gcovi, i, j, k), G16(gconi, i, j, k), 1, &pl_, &ml_);
        S3(cmaxi, i, j, k) = FMAX(FMAX(0.0, pr_), pl_);
        S3(cmini, i, j, k) = FMIN(FMIN(0.0, mr_), ml_);
        face_speed(&V(urijr, 1, i, j, k), G16(gcovj, i, j, k), G16(gconj, i, j, k), 2, &pr_, &mr_);
        face_speed(&V(urijl, 1, i, j, k), G16(gcovj, i, j, k), G16(gconj, i, j, k), 2, &pl_, &ml_);
        S3(cmaxj, i, j, k) = FMAX(FMAX(0.0, pr_), pl_);
        S3(cminj, i, j, k) = FMIN(FMIN(0.0, mr_), ml_);
        face_speed(&V(urikr, 1, i, j, k), G16(gcovk, i, j, k), G16(gconk, i, j, k), 3, &pr_, &mr_);
        face_speed(&V(urikl, 1, i, j, k), G16(gcovk, i, j, k), G16(gconk, i, j, k), 3, &pl_, &ml_);
        S3(cmaxk, i, j, k) = FMAX(FMAX(0.0, pr_), pl_);
        /* calcha.f90:107 takes the right state's plus speed here */
        S3(cmink, i, j, k) = FMIN(FMIN(0.0, pr_), ml_);
      }
}

/* hlle: hll.f90:47-192 */
void orc_hll(double *ww, const double *uuir, const double *uujr, const double *uukr,
             const double *uuil, const double *uujl, const double *uukl,
             const double *wwir, const double *wwjr, const double *wwkr,
             const double *wwil, const double *wwjl, const double *wwkl,
             const double *cmaxi, const double *cmini, const double *cmaxj,
             const double *cminj, const double *cmaxk, const double *cmink, int nm0) {
  const double *UR[3] = {uuir, uujr, uukr}, *UL[3] = {uuil, uujl, uukl};
  const double *FR[3] = {wwir, wwjr, wwkr}, *FL[3] = {wwil, wwjl, wwkl};
  const double *CX[3] = {cmaxi, cmaxj, cmaxk}, *CN[3] = {cmini, cminj, cmink};
  for (int k = nm0; k <= NK - nm0; k++)
    for (int j = nm0; j <= NJ - nm0; j++)
      for (int i = nm0; i <= NI - nm0; i++)
        for (int n = 1; n <= 9; n++)
          for (int d = 0; d < 3; d++) {
            double cmin = S3(CN[d], i, j, k), cmax = S3(CX[d], i, j, k);
            double *out = &WW(ww, d + 1, n, i, j, k);
            if (cmin > 0.0) *out = V(FL[d], n, i, j, k);
            else if (cmin <= 0.0 && cmax > 0.0) {
              double c1 = FMAX(fabs(cmax), fabs(cmin));
              *out = 0.5 * (V(FL[d], n, i, j, k) + V(FR[d], n, i, j, k) - c1 * (V(UR[d], n, i, j, k) - V(UL[d], n, i, j, k)));
            } else if (cmax <= 0.0) *out = V(FR[d], n, i, j, k);
            if (fabs(*out) < 1.0e-40) *out = 0.0;
          }
}

/* fluxct: ct.f90:30-118 */
void orc_ct(double *ww, int nm0, int nm1) {
  if (P.ict != 1) abort();
  const size_t N = (size_t)NI * NJ * NK;
  double *f = (double *)malloc(sizeof(double) * 6 * N);
  double *fyx = f, *fzx = f + N, *fxy = f + 2 * N, *fzy = f + 3 * N, *fxz = f + 4 * N, *fyz = f + 5 * N;
  for (int k = nm0; k <= NK - nm0; k++)
    for (int j = nm0; j <= NJ - nm0; j++)
      for (int i = nm0; i <= NI - nm0; i++) {
        S3(fyx, i, j, k) = WW(ww, 2, 7, i, j, k);
        S3(fzx, i, j, k) = WW(ww, 3, 7, i, j, k);
        S3(fxy, i, j, k) = WW(ww, 1, 8, i, j, k);
        S3(fzy, i, j, k) = WW(ww, 3, 8, i, j, k);
        S3(fxz, i, j, k) = WW(ww, 1, 9, i, j, k);
        S3(fyz, i, j, k) = WW(ww, 2, 9, i, j, k);
      }
  const double c8 = 1.0 / 8.0;
  for (int k = nm1; k <= NK - nm1; k++)
    for (int j = nm1; j <= NJ - nm1; j++)
      for (int i = nm1; i <= NI - nm1; i++) {
#define F(a, I, J, K) S3(a, I, J, K)
        WW(ww, 2, 7, i, j, k) = c8 * (2.0 * F(fyx, i, j, k) + F(fyx, i + 1, j, k) + F(fyx, i - 1, j, k) - F(fxy, i, j, k) -
                                      F(fxy, i, j + 1, k) - F(fxy, i - 1, j, k) - F(fxy, i - 1, j + 1, k));
        WW(ww, 3, 7, i, j, k) = c8 * (2.0 * F(fzx, i, j, k) + F(fzx, i + 1, j, k) + F(fzx, i - 1, j, k) - F(fxz, i, j, k) -
                                      F(fxz, i, j, k + 1) - F(fxz, i - 1, j, k) - F(fxz, i - 1, j, k + 1));
        WW(ww, 1, 8, i, j, k) = c8 * (2.0 * F(fxy, i, j, k) + F(fxy, i, j + 1, k) + F(fxy, i, j - 1, k) - F(fyx, i, j, k) -
                                      F(fyx, i + 1, j, k) - F(fyx, i, j - 1, k) - F(fyx, i + 1, j - 1, k));
        WW(ww, 3, 8, i, j, k) = c8 * (2.0 * F(fzy, i, j, k) + F(fzy, i, j + 1, k) + F(fzy, i, j - 1, k) - F(fyz, i, j, k) -
                                      F(fyz, i, j, k + 1) - F(fyz, i, j - 1, k) - F(fyz, i, j - 1, k + 1));
        WW(ww, 1, 9, i, j, k) = c8 * (2.0 * F(fxz, i, j, k) + F(fxz, i, j, k + 1) + F(fxz, i, j, k - 1) - F(fzx, i, j, k) -
                                      F(fzx, i + 1, j, k) - F(fzx, i, j, k - 1) - F(fzx, i + 1, j, k - 1));
        WW(ww, 2, 9, i, j, k) = c8 * (2.0 * F(fyz, i, j, k) + F(fyz, i, j, k + 1) + F(fyz, i, j, k - 1) - F(fzy, i, j, k) -
                                      F(fzy, i, j + 1, k) - F(fzy, i, j, k - 1) - F(fzy, i, j + 1, k - 1));
#undef F
      }
  free(f);
}

/* caltenr: calsf.f90:6-104 */
void orc_caltenr(const double *uri, double *tenr, const double *gcov, const double *gcon) {
  for (int k = 1; k <= NK; k++)
    for (int j = 1; j <= NJ; j++)
      for (int i = 1; i <= NI; i++) {
        prel_t s;
        prelude(&V(uri, 1, i, j, k), G16(gcov, i, j, k), G16(gcon, i, j, k), &s);
        double *t = &tenr[16 * CELL(i, j, k)];
        for (int n = 0; n < 4; n++)
          for (int m = 0; m < 4; m++) {
            double dl = (m == n) ? 1.0 : 0.0;
            GM(t, m, n) = (s.roh + s.bbsq) * s.ucon[m] * s.ucov[n] + (s.pr + 0.5 * s.bbsq) * dl - s.bbcon[m] * s.bbcov[n];
          }
      }
}

/* calsf: calsf.f90:107-151 */
void orc_calsf(double *sf, const double *tenr, const double *christ, const double *detg) {
  for (int k = 1; k <= NK; k++)
    for (int j = 1; j <= NJ; j++)
      for (int i = 1; i <= NI; i++) {
        const double *t = &tenr[16 * CELL(i, j, k)];
        const double *c = &christ[64 * CELL(i, j, k)];
        double a = 0.0, b = 0.0, cc = 0.0, d = 0.0;
        for (int n = 0; n < 4; n++)
          for (int m = 0; m < 4; m++) {
            double tv = GM(t, m, n);
            a = a + tv * c[n + 4 * 1 + 16 * m];
            b = b + tv * c[n + 4 * 2 + 16 * m];
            cc = cc + tv * c[n + 4 * 3 + 16 * m];
            d = d + tv * c[n + 4 * 0 + 16 * m];
          }
        double dg = S3(detg, i, j, k);
        SF(sf, 2, i, j, k) = dg * a;
        SF(sf, 3, i, j, k) = dg * b;
        SF(sf, 4, i, j, k) = dg * cc;
        SF(sf, 5, i, j, k) = dg * d;
      }
}

/* rkt.f90 */
#define RK_LOOP for (int k = 1 + nm1; k <= NK - nm1; k++) for (int j = 1 + nm1; j <= NJ - nm1; j++) for (int i = 1 + nm1; i <= NI - nm1; i++) for (int m = 1; m <= 9; m++)
#define DIV1 (WW(ww, 1, m, i, j, k) - WW(ww, 1, m, i - 1, j, k))
#define DIV2 (WW(ww, 2, m, i, j, k) - WW(ww, 2, m, i, j - 1, k))
#define DIV3 (WW(ww, 3, m, i, j, k) - WW(ww, 3, m, i, j, k - 1))
void orc_rk2fst(double *uh, const double *ww, const double *uu, int nm1, const double *a1, const double *a2, const double *a3) {
  RK_LOOP V(uh, m, i, j, k) = V(uu, m, i, j, k) - a1[i - 1] * DIV1 - a2[j - 1] * DIV2 - a3[k - 1] * DIV3;
}
void orc_rk2snd(double *uu, const double *ww, const double *uo, const double *uh, int nm1, const double *a1, const double *a2, const double *a3) {
  RK_LOOP V(uu, m, i, j, k) = (V(uo, m, i, j, k) + V(uh, m, i, j, k)) * 0.5 + (-a1[i - 1] * DIV1 - a2[j - 1] * DIV2 - a3[k - 1] * DIV3) * 0.5;
}
void orc_rk3snd(double *uu, const double *ww, const double *uo, const double *uh, int nm1, const double *a1, const double *a2, const double *a3) {
  RK_LOOP V(uu, m, i, j, k) = (3.0 * V(uo, m, i, j, k) + V(uh, m, i, j, k)) * 0.25 + (-a1[i - 1] * DIV1 - a2[j - 1] * DIV2 - a3[k - 1] * DIV3) * 0.25;
}
void orc_rk3trd(double *uu, const double *ww, const double *uo, const double *us, int nm1, const double *a1, const double *a2, const double *a3) {
  const double third = (double)(1.0f / 3.0f), twothird = (double)(2.0f / 3.0f); /* rkt.f90:169,172 */
  RK_LOOP V(uu, m, i, j, k) = (V(uo, m, i, j, k) + 2.0 * V(us, m, i, j, k)) * third + (-a1[i - 1] * DIV1 - a2[j - 1] * DIV2 - a3[k - 1] * DIV3) * twothird;
}
/* rk2adsff (which=0, w=1), rk2adsfs (1, 0.5), rk3adsfs (2, 0.25), rk3adsft (3, 2./3. float): rkt.f90:183-320 */
void orc_rkadsf(double *uu, const double *sf, double dt, int nm1, int which) {
  for (int k = 1 + nm1; k <= NK - nm1; k++)
    for (int j = 1 + nm1; j <= NJ - nm1; j++)
      for (int i = 1 + nm1; i <= NI - nm1; i++)
        for (int n = 2; n <= 5; n++) {
          double s = SF(sf, n, i, j, k) * dt;
          if (which == 0) V(uu, n, i, j, k) = V(uu, n, i, j, k) + s;
          else if (which == 1) V(uu, n, i, j, k) = V(uu, n, i, j, k) + s * 0.5;
          else if (which == 2) V(uu, n, i, j, k) = V(uu, n, i, j, k) + s * 0.25;
          else V(uu, n, i, j, k) = V(uu, n, i, j, k) + s * (double)(2.0f / 3.0f);
        }
}
#undef RK_LOOP
#undef DIV1
#undef DIV2
#undef DIV3

/* ------------------------------------------------------------------ */
/* recov1a: recov.f90:578-1228                                         */
/* ------------------------------------------------------------------ */
void orc_recov1a(double *uu, double *uri, const double *gcov, const double *gcon,
                 const double *detg, const double *x1, int nm1) {
  const double safty = 1.0e-4, saftybig = 1.0e5;
  double dv = 1.0e-2;
  const double gam = P.gam;
  const double gam1 = gam / (gam - 1.0);
  const double gam2 = 1.0 / gam1;
  const int iter = P.iter;
  double wg = 0.0; /* recov.f90:632: carried from cell to cell when no root is positive */
  for (int k = 1 + nm1; k <= NK - nm1; k++)
    for (int j = 1 + nm1; j <= NJ - nm1; j++)
      for (int i = 1 + nm1; i <= NI - nm1; i++) {
        double delta = 1.0, deltaend = 1.0e-7;
        int irecov = 0;
        const double *g = G16(gcov, i, j, k), *gi = G16(gcon, i, j, k);
        double alpha1 = 1.0 / sqrt(-GM(gi, 0, 0));
        double detg1 = S3(detg, i, j, k);
        double *u = &V(uu, 1, i, j, k), *w = &V(uri, 1, i, j, k);
        double dro = alpha1 * u[0] * (1.0 / detg1);
        double qcov[4], qcon[4], bcon[4], bcov[4], oncov[4], oncon[4];
        qcov[0] = alpha1 * (u[4] - u[0]) * (1.0 / detg1);
        qcov[1] = alpha1 * u[1] * (1.0 / detg1);
        qcov[2] = alpha1 * u[2] * (1.0 / detg1);
        qcov[3] = alpha1 * u[3] * (1.0 / detg1);
        double gss = alpha1 * u[5] * (1.0 / detg1);
        bcon[0] = 0.0;
        bcon[1] = alpha1 * u[6] * (1.0 / detg1);
        bcon[2] = alpha1 * u[7] * (1.0 / detg1);
        bcon[3] = alpha1 * u[8] * (1.0 / detg1);
        double ro, pp, util[4], util1[4];
        util[1] = w[1]; util[2] = w[2]; util[3] = w[3];
        double ro1 = w[0], pp1 = w[4];
        util1[1] = w[1]; util1[2] = w[2]; util1[3] = w[3];
        lower(bcon, bcov, g);
        upper(qcov, qcon, gi);
        double bsq = bcon[1] * bcov[1] + bcon[2] * bcov[2] + bcon[3] * bcov[3];
        double qdotb = qcov[0] * bcon[0] + qcov[1] * bcon[1] + qcov[2] * bcon[2] + qcov[3] * bcon[3];
        double qdotbsq = qdotb * qdotb;
        oncov[0] = -1.0 * alpha1; oncov[1] = 0.0; oncov[2] = 0.0; oncov[3] = 0.0;
        upper(oncov, oncon, gi);
        double qdotn = qcon[0] * oncov[0];
        double qsq = qcov[0] * qcon[0] + qcov[1] * qcon[1] + qcov[2] * qcon[2] + qcov[3] * qcon[3];
        double qtsq = qsq + (qdotn * qdotn);
        double utsq = 0.0;
        for (int m = 1; m <= 3; m++)
          for (int n = 1; n <= 3; n++) utsq = utsq + GM(g, m, n) * util[m] * util[n];
        if (utsq < 0.0 && fabs(utsq) <= safty) utsq = fabs(utsq);
        if (fabs(utsq) > 1.0e4) utsq = 1.0e4;
        double gfsq = 1.0 + utsq;
        double gf = sqrt(gfsq);
        double vsq = (gfsq - 1.0) * (1.0 / gfsq);
        double vsq1 = vsq;
        double t1a = 4.0 - vsq1;
        double t1b = 4.0 * (bsq + qdotn);
        double t1c = qtsq + bsq * bsq + 2.0 * bsq * qdotn;
        double t1d = (t1b * t1b) - 4.0 * t1a * t1c;
        if (t1d <= 0.0) t1d = fabs(t1d);
        double wgp = (-t1b + sqrt(t1d)) * (1.0 / (2.0 * t1a));
        double wgm = (-t1b - sqrt(t1d)) * (1.0 / (2.0 * t1a));
        if (wgp > 0.0) wg = wgp;
        else if (wgm >= 0.0) wg = wgm;
        else {
          CNT.recov_wg_carry++;
          if (DBG_CARRY && DBG_N < DBG_CAP) DBG_CARRY[DBG_N] = (long long)CELL(i, j, k);
          DBG_N++;
        }
        /* recov.f90:823-850: "goto 95" adjustment loops */
        int guard = 0;
        for (; guard < 100000; guard++) {
          pp = gam2 * (wg * gf - dro * sqrt(gf));
          double t3a = (wg * wg * wg) * (wg + 2.0 * bsq) - qdotbsq * (2.0 * wg + bsq);
          double t3b = (wg * wg) * (qtsq - bsq * bsq);
          double t3c = (wg * wg * wg) + (wg * wg) * (0.5 * bsq + qdotn - pp) - 0.5 * qdotbsq;
          if (t3a <= t3b) { wg = fabs(1.5 * wg); continue; }
          if (t3c > 0.0) { wg = fabs((double)0.9f * wg); continue; }
          break;
        }
        if (guard >= 100000) CNT.guard++;
        double wsq = wg * wg;
        double xsq = (bsq + wg) * (bsq + wg);
        vsq = fabs((wsq * qtsq + qdotbsq * (bsq + 2.0 * wg)) * (1.0 / (wsq * xsq)));
        if (vsq > 1.0) { CNT.recov_vsq_gt1++; vsq = 1.0 - dv; }
        double vsqo = vsq;
        double xxn = wg, yyn = vsq;
        double w3, gtmp, ptmp;
        for (int nnn = 1; nnn <= iter; nnn++) {
          if (!(fabs(delta) >= deltaend && irecov == 0)) break; /* remaining trips do nothing */
          if (nnn == iter) { irecov = 1; break; }
          wg = xxn; vsq = yyn;
          wsq = wg * wg; w3 = wsq * wg;
          gtmp = 1.0 - vsq;
          if (gtmp < 0.0 || gtmp < safty) { gtmp = safty; vsq = 1.0 - gtmp; }
          ptmp = gam2 * (wg * gtmp - dro * sqrt(gtmp));
          double f1 = qtsq - vsq * (wg + bsq) * (wg + bsq) + (qdotbsq * (bsq + 2.0 * wg)) * (1.0 / wsq);
          double g1 = -qdotn - 0.5 * bsq * (1.0 + vsq) + 0.5 * (qdotbsq * (1.0 / wsq)) - wg + ptmp;
          double dpdw = gam2 * gtmp;
          double dpdvsq = -gam2 * wg + 0.5 * gam2 * dro * (1.0 / (sqrt(gtmp)));
          double dfx = -2.0 * (wg + bsq) * (vsq + qdotbsq * (1.0 / w3));
          double dfy = -((wg + bsq) * (wg + bsq));
          double dgx = -1.0 - (qdotbsq * (1.0 / w3)) + dpdw;
          double dgy = -0.5 * bsq + dpdvsq;
          double det = dfx * dgy - dfy * dgx;
          if (det == 0.0) { CNT.recov_det0++; det = 1.0; }
          double dx = (-dgy * f1 + dfy * g1) * (1.0 / det);
          double dy = (dgx * f1 - dfx * g1) * (1.0 / det);
          delta = 0.5 * (fabs(dx) + fabs(dy));
          xxn = wg + dx;
          yyn = vsq + dy;
          if (xxn < 0.0) xxn = fabs(xxn);
          if (xxn > saftybig) xxn = saftybig;
          if (yyn < 0.0) yyn = fabs(yyn);
          if (fabs(yyn) > 1.0) { dv = 1.0e-2; yyn = 1.0 - dv; }
        }
        if (irecov == 0) {
          wg = xxn; vsq = yyn;
          gtmp = 1.0 - vsq;
          if (gtmp < 0.0 || gtmp < safty) { gtmp = safty; vsq = 1.0 - gtmp; }
          gf = 1.0 / sqrt(gtmp);
          ro = dro * (1.0 / gf);
          double w2 = wg * (gtmp);
          pp = gam2 * (w2 - ro);
          if (ro < 0.0 || pp < 0.0) irecov = 1;
          if (irecov == 0) {
            for (int m = 1; m <= 3; m++) {
              double qt = qcon[m] + oncon[m] * qdotn;
              util[m] = (gf * (1.0 / (wg + bsq))) * (qt + qdotb * bcon[m] * (1.0 / wg));
            }
          }
          wg = xxn; vsq = yyn;
        }
        if (irecov == 1) {
          /* entropy inversion: recov.f90:975-1147 */
          CNT.recov_fallback++;
          delta = 1.0; deltaend = 1.0e-6; irecov = 0;
          xxn = vsqo;
          for (int nnn = 1; nnn <= iter; nnn++) {
            if (!(fabs(delta) >= deltaend && irecov == 0)) break;
            if (nnn == iter) { irecov = 1; break; }
            vsq = xxn;
            gtmp = 1.0 - vsq;
            if (gtmp < 0.0 || gtmp < safty) { gtmp = safty; vsq = 1.0 - gtmp; }
            gf = 1.0 / sqrt(gtmp);
            gfsq = gf * gf;
            ro = dro * (1.0 / gf);
            pp = (gss / dro) * (pow(ro, gam));
            double roh = ro + gam1 * pp;
            wg = fabs(gfsq * roh);
            wsq = wg * wg; w3 = wsq * wg;
            double dgfdvsq = 0.5 * pow(gtmp, -1.5);
            double dwdvsq = dro * dgfdvsq + gam1 * gss * (pow(dro, gam - 1.0)) * (2.0 - gam) * (pow(gf, 1.0 - gam)) * dgfdvsq;
            double f = qtsq - vsq * (wg + bsq) * (wg + bsq) + (qdotbsq * (bsq + 2.0 * wg)) / wsq;
            double a = -((wg + bsq) * (wg + bsq));
            double b = -2.0 * vsq * (wg + bsq);
            double c = -2.0 * qdotbsq * (bsq + 2.0 * wg) * (1.0 / w3);
            double d = 2.0 * qdotbsq * (1.0 / wsq);
            double df = a + (b + c + d) * dwdvsq;
            delta = f / df;
            xxn = vsq - delta;
            if (xxn < 0.0) xxn = fabs(xxn);
            if (fabs(xxn) >= 1.0) xxn = 1.0 - dv;
          }
          if (irecov == 0) {
            vsq = xxn;
            gtmp = 1.0 - vsq;
            if (gtmp < 0.0 || gtmp < safty) gtmp = safty;
            gf = 1.0 / sqrt(gtmp);
            gfsq = gf * gf;
            ro = dro * (1.0 / gf);
            pp = (gss * (1.0 / gf)) * (pow(ro, gam - 1.0));
            double roh = ro + gam1 * pp;
            if (ro < 0.0) { CNT.recov_neg_rho++; ro = P.dmin * pow(x1[i - 1], -1.5); }
            if (pp < 0.0) { CNT.recov_neg_p++; pp = (gam - 1.0) * P.pmin * pow(x1[i - 1], -2.5); }
            wg = fabs(gfsq * roh);
            for (int m = 1; m <= 3; m++) {
              double qt = qcon[m] + oncon[m] * qdotn;
              util[m] = (gf / (wg + bsq)) * (qt + qdotb * bcon[m] * (1.0 / wg));
            }
          }
        }
        double ucon[4];
        if (irecov == 0) {
          w[0] = ro; w[1] = util[1]; w[2] = util[2]; w[3] = util[3]; w[4] = pp;
          w[5] = ro * pp / pow(ro, gam - 1.0);
          w[6] = u[6] * (1.0 / detg1); w[7] = u[7] * (1.0 / detg1); w[8] = u[8] * (1.0 / detg1);
          calucon(util, ucon, gf, gi);
          u[5] = detg1 * w[5] * ucon[0];
        } else {
          CNT.recov_fail++;
          w[0] = ro1; w[1] = util1[1]; w[2] = util1[2]; w[3] = util1[3]; w[4] = pp1;
          w[5] = ro1 * pp1 / pow(ro1, gam);
          w[6] = u[6] * (1.0 / detg1); w[7] = u[7] * (1.0 / detg1); w[8] = u[8] * (1.0 / detg1);
          calucon(util, ucon, gf, gi);
          u[5] = detg1 * ro1 * w[5] * ucon[0];
        }
      }
}

/* ------------------------------------------------------------------ */
/* recov2: recov.f90:1919-2454 (Mignone & McKinney 2007: one Newton-     */
/* Raphson unknown W' = W - D, gamma-law EoS)                            */
/* ------------------------------------------------------------------ */
void orc_recov2(double *uu, double *uri, const double *gcov, const double *gcon,
                const double *detg, const double *x1, int nm1) {
  const double safty = 1.0e-5, dv = 1.0e-4;  /* recov.f90:1962-1964 */
  const double gam = P.gam;
  const double gam1 = gam / (gam - 1.0);
  const double gam2 = 1.0 / gam1;
  const int iter = P.iter;
  double wg = 0.0; /* recov.f90:1968: carried from cell to cell when no root is admissible */
  for (int k = 1 + nm1; k <= NK - nm1; k++)
    for (int j = 1 + nm1; j <= NJ - nm1; j++)
      for (int i = 1 + nm1; i <= NI - nm1; i++) {
        double delta = 1.0;
        const double deltaend = 1.0e-8;
        int irecov = 0;
        const double *g = G16(gcov, i, j, k), *gi = G16(gcon, i, j, k);
        double alpha1 = 1.0 / sqrt(-GM(gi, 0, 0));
        double detg1 = S3(detg, i, j, k);
        double *u = &V(uu, 1, i, j, k), *w = &V(uri, 1, i, j, k);
        double dro = alpha1 * u[0] / detg1;
        double qcov[4], qcon[4], bcon[4], bcov[4], oncov[4], oncon[4];
        qcov[0] = alpha1 * (u[4] - u[0]) * (1.0 / detg1);
        qcov[1] = alpha1 * u[1] * (1.0 / detg1);
        qcov[2] = alpha1 * u[2] * (1.0 / detg1);
        qcov[3] = alpha1 * u[3] * (1.0 / detg1);
        bcon[0] = 0.0;
        bcon[1] = alpha1 * u[6] * (1.0 / detg1);
        bcon[2] = alpha1 * u[7] * (1.0 / detg1);
        bcon[3] = alpha1 * u[8] * (1.0 / detg1);
        double util[4], ro = 0.0, pp = 0.0;
        util[1] = w[1]; util[2] = w[2]; util[3] = w[3];
        const double ro1 = w[0], pp1 = w[4];
        const double util1[4] = {0.0, w[1], w[2], w[3]};
        lower(bcon, bcov, g);
        upper(qcov, qcon, gi);
        double bsq = bcon[1] * bcov[1] + bcon[2] * bcov[2] + bcon[3] * bcov[3];
        double qdotb = qcov[0] * bcon[0] + qcov[1] * bcon[1] + qcov[2] * bcon[2] + qcov[3] * bcon[3];
        double qdotbsq = qdotb * qdotb;
        oncov[0] = -1.0 * alpha1; oncov[1] = 0.0; oncov[2] = 0.0; oncov[3] = 0.0;
        upper(oncov, oncon, gi);
        double qdotn = qcon[0] * oncov[0];
        double qsq = qcov[0] * qcon[0] + qcov[1] * qcon[1] + qcov[2] * qcon[2] + qcov[3] * qcon[3];
        double qtsq = qsq + (qdotn * qdotn);
        double utsq = 0.0;
        for (int m = 1; m <= 3; m++)
          for (int n = 1; n <= 3; n++) utsq = utsq + GM(g, m, n) * util[m] * util[n];
        if (utsq < 0.0 && fabs(utsq) <= safty) utsq = fabs(utsq);
        if (utsq < 0.0 || fabs(utsq) > 1.0e4) utsq = 1.0e4;
        double gfsq = 1.0 + utsq;
        double gf = sqrt(gfsq);
        double vsq = (gfsq - 1.0) * (1.0 / gfsq);
        /* initial W: the larger admissible root of the quadratic (recov.f90:2125-2146) */
        double t1a = 4.0 - vsq;
        double t1b = 4.0 * (bsq + qdotn);
        double t1c = qtsq + bsq * bsq + 2.0 * bsq * qdotn;
        double t1d = (t1b * t1b) - 4.0 * t1a * t1c;
        if (t1d <= 0.0) t1d = fabs(t1d);
        double wgp = (-t1b + sqrt(t1d)) * (1.0 / (2.0 * t1a));
        double wgm = (-t1b - sqrt(t1d)) * (1.0 / (2.0 * t1a));
        if (wgp > 0.0) wg = wgp;
        else if (wgm >= 0.0) wg = wgm;
        else {
          CNT.recov_wg_carry++;
          if (DBG_CARRY && DBG_N < DBG_CAP) DBG_CARRY[DBG_N] = (long long)CELL(i, j, k);
          DBG_N++;
        }
        /* recov.f90:2150-2172: "goto 95" adjustment loops */
        int guard = 0;
        for (; guard < 100000; guard++) {
          pp = gam2 * (wg * gf - dro * sqrt(gf));
          double t3a = (wg * wg * wg) * (wg + 2.0 * bsq) - qdotbsq * (2.0 * wg + bsq);
          double t3b = (wg * wg) * (qtsq - bsq * bsq);
          double t3c = (wg * wg * wg) + (wg * wg) * (0.5 * bsq + qdotn - pp) - 0.5 * qdotbsq;
          if (t3a <= t3b) { wg = fabs(1.5 * wg); continue; }
          if (t3c > 0.0) { wg = fabs((double)0.9f * wg); continue; }
          break;
        }
        if (guard >= 100000) CNT.guard++;
        double wgd = wg - dro;
        double xxn = wgd;
        /* Newton-Raphson on W' (recov.f90:2180-2300) */
        for (int nnn = 1; nnn <= iter; nnn++) {
          if (!(fabs(delta) >= deltaend && irecov == 0)) break; /* remaining trips do nothing */
          if (nnn == iter) { irecov = 1; break; }
          wgd = xxn;
          wg = wgd + dro;
          double wsq = wg * wg;
          double w3 = wsq * wg;
          double xsq = (bsq + wg) * (bsq + wg);
          vsq = fabs((wsq * qtsq + qdotbsq * (bsq + 2.0 * wg)) * (1.0 / (wsq * xsq)));
          if (fabs(vsq) > 1.0) { CNT.recov_vsq_gt1++; vsq = 1.0 - dv; }
          if (vsq < 0.0) vsq = fabs(vsq);
          double gtmp = 1.0 - vsq;
          if (gtmp < 0.0 || gtmp < safty) { gtmp = safty; vsq = 1.0 - gtmp; }
          double usq = (1.0 - gtmp) / (gtmp);
          gf = sqrt(1.0 + usq);
          gfsq = gf * gf;
          ro = dro / gf;
          double xi = (wg - dro * gf) * (1.0 / gfsq);
          pp = gam2 * xi;
          const double dpdxi = gam2, dpdro = 0.0;
          double t1a_ = (bsq + wg);
          double t1b_ = (bsq + wg) * (bsq + wg) * (bsq + wg);
          double t1c_ = (qtsq == 0.0) ? 0.0 : -2.0 * qtsq * (1.0 / t1b_);
          double t1d_ = (qdotbsq == 0.0) ? 0.0 : -2.0 * qdotbsq * (3.0 * wg * t1a_ + bsq * bsq) * (1.0 / (t1b_ * w3));
          double dvdwd = t1c_ + t1d_;
          double dxidwd = (1.0 / gfsq) - 0.5 * gf * (dro + 2.0 * gf * xi) * dvdwd;
          double drodwd = -0.5 * dro * gf * dvdwd;
          double dpdwd = dpdxi * dxidwd + dpdro * drodwd;
          double t5a = bsq * qtsq - qdotbsq;
          double t5b = bsq + wgd + dro;
          double t5c = 0.0, t5d = 0.0;
          if (t5a != 0.0) {
            t5c = 0.5 * t5a * (1.0 / (t5b * t5b));
            t5d = t5a * (1.0 / (t5b * t5b * t5b));
          }
          double f = -(qdotn + dro) - wgd + pp - 0.5 * bsq - t5c;
          double df = -1.0 + dpdwd + t5d;
          delta = f / df;
          xxn = wgd - delta;
          if (xxn < 0.0) xxn = fabs(xxn);
          if (xxn > 1.0e8) xxn = 1.0e8;
        }
        if (irecov == 0) {
          /* recov.f90:2305-2370 */
          wgd = xxn;
          wg = wgd + dro;
          double wsq = wg * wg;
          double xsq = (bsq + wg) * (bsq + wg);
          vsq = fabs((wsq * qtsq + qdotbsq * (bsq + 2.0 * wg)) * (1.0 / (wsq * xsq)));
          if (vsq < 0.0) vsq = fabs(vsq);
          if (vsq > 1.0) { CNT.recov_vsq_gt1++; vsq = 1.0 - dv; }
          double gtmp = 1.0 - vsq;
          if (gtmp < safty) gtmp = safty;
          double usq = (1.0 - gtmp) / gtmp;
          gf = sqrt(1.0 + usq);
          gfsq = gf * gf;
          ro = dro / gf;
          double xi = wg * (1.0 / gfsq) - ro;
          pp = gam2 * xi;
          if (ro < 0.0) { CNT.recov_neg_rho++; ro = P.dmin * pow(x1[i - 1], -1.5); }
          if (pp < 0.0) { CNT.recov_neg_p++; pp = P.pmin * (gam - 1.0) * pow(x1[i - 1], -2.5); }
          for (int m = 1; m <= 3; m++) {
            double qt = qcon[m] + oncon[m] * qdotn;
            util[m] = (gf * (1.0 / (wg + bsq))) * (qt + qdotb * bcon[m] * (1.0 / wg));
          }
          w[0] = ro; w[1] = util[1]; w[2] = util[2]; w[3] = util[3]; w[4] = pp;
          w[5] = pp / pow(ro, gam - 1.0);
        } else {
          CNT.recov_fail++;
          w[0] = ro1; w[1] = util1[1]; w[2] = util1[2]; w[3] = util1[3]; w[4] = pp1;
          w[5] = pp1 / pow(ro1, gam - 1.0);
        }
        w[6] = u[6] * (1.0 / detg1); w[7] = u[7] * (1.0 / detg1); w[8] = u[8] * (1.0 / detg1);
      }
}

/* recov: recov.f90:2-47, the inversion iwvec selects */
void orc_recov(double *uu, double *uri, const double *gcov, const double *gcon,
               const double *detg, const double *x1, int nm1) {
  if (P.iwvec == 1) orc_recov1a(uu, uri, gcov, gcon, detg, x1, nm1);
  else if (P.iwvec == 2) orc_recov2(uu, uri, gcov, gcon, detg, x1, nm1);
  else abort();
}

/* ------------------------------------------------------------------ */
/* bnd4 = bndpefb4 + bndryfb4: bnd.f90:4-1059, single-rank branches     */
/* ------------------------------------------------------------------ */
static void bnd_side(double *uu, const double *uo, int m, int dir, int outer, int type, int nm1) {
  /* dir 1:i 2:j 3:k; applies over the full extent of the other two axes */
  int nmax = dir == 1 ? NI : dir == 2 ? NJ : NK;
  int na = dir == 1 ? NJ : NI, nb = dir == 3 ? NJ : NK;
#define AT(a, p, A, B) (dir == 1 ? &V(a, m, p, A, B) : dir == 2 ? &V(a, m, A, p, B) : &V(a, m, A, B, p))
#define AT2(a, p, A, B) (dir == 1 ? &V(a, 2, p, A, B) : dir == 2 ? &V(a, 2, A, p, B) : &V(a, 2, A, B, p))
  if (type == 1) return;
  for (int B = 1; B <= nb; B++)
    for (int A = 1; A <= na; A++) {
      if (!outer) {
        switch (type) {
          case 2: for (int n = 1; n <= nm1; n++) *AT(uu, n, A, B) = *AT(uo, n, A, B); break;
          case 3: for (int n = nm1; n >= 1; n--) *AT(uu, n, A, B) = (4.0 * *AT(uu, n + 1, A, B) - *AT(uu, n + 2, A, B)) / 3.0; break;
          case 4: for (int n = nm1; n >= 1; n--) *AT(uu, n, A, B) = 2.0 * *AT(uu, n + 1, A, B) - *AT(uu, n + 2, A, B); break;
          case 5: for (int n = nm1; n >= 1; n--) *AT(uu, n, A, B) = *AT(uu, 1 + nm1, A, B); break;
          case 7: for (int n = nm1; n >= 1; n--) *AT(uu, n, A, B) = *AT(uo, n, A, B) + *AT(uu, n + 1, A, B) - *AT(uo, n + 1, A, B); break;
          case 12: for (int n = 1; n <= nm1; n++) *AT(uu, n, A, B) = -*AT(uu, 2 * nm1 - n + 1, A, B); break;
          case 13: for (int n = 1; n <= nm1; n++) *AT(uu, n, A, B) = 0.0; break;
          case 14: for (int n = 1; n <= nm1; n++) *AT(uu, n, A, B) = *AT(uu, 2 * nm1 - n + 1, A, B); break;
          default: abort();
        }
      } else {
        switch (type) {
          case 2: for (int n = 0; n <= nm1 - 1; n++) *AT(uu, nmax - n, A, B) = *AT(uo, nmax - n, A, B); break;
          case 3: for (int n = nm1 - 1; n >= 0; n--) *AT(uu, nmax - n, A, B) = (4.0 * *AT(uu, nmax - n - 1, A, B) - *AT(uu, nmax - n - 2, A, B)) / 3.0; break;
          case 4: for (int n = nm1 - 1; n >= 0; n--) *AT(uu, nmax - n, A, B) = 2.0 * *AT(uu, nmax - n - 1, A, B) - *AT(uu, nmax - n - 2, A, B); break;
          case 5: for (int n = nm1 - 1; n >= 0; n--) *AT(uu, nmax - n, A, B) = *AT(uu, nmax - nm1, A, B); break;
          case 7: for (int n = nm1 - 1; n >= 0; n--) *AT(uu, nmax - n, A, B) = *AT(uo, nmax - n, A, B) + *AT(uu, nmax - n - 1, A, B) - *AT(uo, nmax - n - 1, A, B); break;
          case 12: for (int n = 0; n <= nm1 - 1; n++) *AT(uu, nmax - n, A, B) = -*AT(uu, nmax - 2 * nm1 + 1 + n, A, B); break;
          case 13: for (int n = 0; n <= nm1 - 1; n++) *AT(uu, nmax - n, A, B) = 0.0; break;
          case 14: for (int n = 0; n <= nm1 - 1; n++) *AT(uu, nmax - n, A, B) = *AT(uu, nmax - 2 * nm1 + 1 + n, A, B); break;
          case 18:
            if (dir != 1) abort(); /* bnd.f90:598-610, x1 outer only */
            for (int n = nm1 - 1; n >= 0; n--) {
              if (*AT2(uu, nmax - nm1, A, B) < 0.0) *AT(uu, nmax - n, A, B) = 0.0;
              else *AT(uu, nmax - n, A, B) = *AT(uu, nmax - nm1, A, B);
            }
            break;
          default: abort();
        }
      }
    }
#undef AT
#undef AT2
}

void orc_bnd4(double *uu, const double *uo, int nm1) {
  /* bndpefb4: bnd.f90:25-381, periodic copies when one rank spans the axis */
  int ib0 = 1 + nm1, ib1 = NI - nm1;
  for (int k = 1 + nm1; k <= NK - nm1; k++)
    for (int j = 1 + nm1; j <= NJ - nm1; j++)
      for (int i = 1; i <= nm1; i++)
        for (int m = 1; m <= 9; m++)
          if (P.iboux1in[m - 1] == 1 || P.iboux1ot[m - 1] == 1) {
            V(uu, m, ib0 - i, j, k) = V(uu, m, ib1 + 1 - i, j, k);
            V(uu, m, ib1 + i, j, k) = V(uu, m, ib0 - 1 + i, j, k);
          }
  ib0 = 1 + nm1; ib1 = NJ - nm1;
  for (int k = 1 + nm1; k <= NK - nm1; k++)
    for (int j = 1; j <= nm1; j++)
      for (int i = 1 + nm1; i <= NI - nm1; i++)
        for (int m = 1; m <= 9; m++)
          if (P.iboux2in[m - 1] == 1 || P.iboux2ot[m - 1] == 1) {
            V(uu, m, i, ib0 - j, k) = V(uu, m, i, ib1 + 1 - j, k);
            V(uu, m, i, ib1 + j, k) = V(uu, m, i, ib0 - 1 + j, k);
          }
  ib0 = 1 + nm1; ib1 = NK - nm1;
  for (int k = 1; k <= nm1; k++)
    for (int j = 1 + nm1; j <= NJ - nm1; j++)
      for (int i = 1 + nm1; i <= NI - nm1; i++)
        for (int m = 1; m <= 9; m++)
          if (P.iboux3in[m - 1] == 1 || P.iboux3ot[m - 1] == 1) {
            V(uu, m, i, j, ib0 - k) = V(uu, m, i, j, ib1 + 1 - k);
            V(uu, m, i, j, ib1 + k) = V(uu, m, i, j, ib0 - 1 + k);
          }
  /* bndryfb4: bnd.f90:384-1059, x1 then x2 then x3, variables outermost */
  for (int m = 1; m <= 9; m++) {
    bnd_side(uu, uo, m, 1, 0, P.iboux1in[m - 1], nm1);
    bnd_side(uu, uo, m, 1, 1, P.iboux1ot[m - 1], nm1);
  }
  for (int m = 1; m <= 9; m++) {
    bnd_side(uu, uo, m, 2, 0, P.iboux2in[m - 1], nm1);
    bnd_side(uu, uo, m, 2, 1, P.iboux2ot[m - 1], nm1);
  }
  for (int m = 1; m <= 9; m++) {
    bnd_side(uu, uo, m, 3, 0, P.iboux3in[m - 1], nm1);
    bnd_side(uu, uo, m, 3, 1, P.iboux3ot[m - 1], nm1);
  }
}

/* infchk: bnd.f90:1636-1824; the inner-boundary branch writes nothing */
void orc_infchk(double *uri, const double *gcov, const double *gcon, int nm1) {
  const double gflmax = 10.0;
  for (int k = 1; k <= NK; k++)
    for (int j = 1; j <= NJ; j++)
      for (int n = 0; n <= nm1 - 1; n++) {
        int i = NI - n;
        double util[4] = {0.0, V(uri, 2, i, j, k), V(uri, 3, i, j, k), V(uri, 4, i, j, k)};
        const double *g = G16(gcov, i, j, k), *gi = G16(gcon, i, j, k);
        double gfl = calgfl(util, g);
        double vel[4];
        vel[1] = util[1] * (1.0 / gfl);
        vel[2] = util[2] * (1.0 / gfl);
        vel[3] = util[3] * (1.0 / gfl);
        if (vel[1] < 0.0) {
          double alpha1 = 1.0 / sqrt(-GM(gi, 0, 0));
          double beta1 = alpha1 * alpha1 * GM(gi, 0, 1);
          vel[1] = beta1 * (1.0 / alpha1);
          double vsq = 0.0;
          for (int m = 1; m <= 3; m++)
            for (int nn = 1; nn <= 3; nn++) vsq = vsq + GM(g, m, nn) * vel[m] * vel[nn];
          if (vsq >= 1.0) vsq = 1.0 - 1.0 / (gflmax * gflmax);
          gfl = 1.0 / sqrt(1.0 - vsq);
          V(uri, 2, i, j, k) = vel[1] * gfl;
          V(uri, 3, i, j, k) = vel[2] * gfl;
          V(uri, 4, i, j, k) = vel[3] * gfl;
        }
      }
}

/* atmosphere1a: correction.f90:226-447 (its conserved-variable tail is dead code) */
void orc_atmosphere1a(double *uri, const double *gcov, const double *gcon, const double *x1) {
  const double epsilon = 1.0e-2, gflmax = 10.0, gam = P.gam;
  for (int k = 1; k <= NK; k++)
    for (int j = 1; j <= NJ; j++)
      for (int i = 1; i <= NI; i++) {
        double rr = x1[i - 1];
        double *w = &V(uri, 1, i, j, k);
        double de = w[0], pr = w[4];
        double util[4] = {0.0, w[1], w[2], w[3]};
        double bcon[4] = {0.0, w[6], w[7], w[8]};
        const double *g = G16(gcov, i, j, k);
        (void)gcon;
        double demin = P.dmin * pow(rr, -1.5);
        double prmin = (gam - 1.0) * P.pmin * pow(rr, -2.5);
        double dthe = 1.0 * (1.0 + epsilon) * demin;
        double dthe1 = 1.0 * (1.0 + epsilon) * prmin;
        double bsq = bcon[1] * bcon[1] + bcon[2] * bcon[2] + bcon[3] * bcon[3];
        if (de < dthe || (pr < dthe1 && bsq <= 1.0e-20)) {
          CNT.atm_floor++;
          de = demin; pr = prmin;
          double vel[4] = {0.0, 0.0, 0.0, 0.0};
          double vsq = 0.0;
          for (int m = 1; m <= 3; m++)
            for (int n = 1; n <= 3; n++) vsq = vsq + GM(g, m, n) * vel[m] * vel[n];
          if (vsq >= 1.0) vsq = 1.0 - 1.0 / (gflmax * gflmax);
          double gfl = 1.0 / sqrt(1.0 - vsq);
          util[1] = vel[1] * gfl; util[2] = vel[2] * gfl; util[3] = vel[3] * gfl;
        }
        w[0] = de; w[1] = util[1]; w[2] = util[2]; w[3] = util[3]; w[4] = pr; w[5] = 0.0;
        w[6] = bcon[1]; w[7] = bcon[2]; w[8] = bcon[3];
      }
}

/* pminmax2b: correction.f90:744-999 */
void orc_pminmax2b(double *uri, const double *gcov, const double *gcon, const double *x1) {
  const double gflmax = 20.0, gam = P.gam;
  for (int k = 1; k <= NK; k++)
    for (int j = 1; j <= NJ; j++)
      for (int i = 1; i <= NI; i++) {
        double rr = x1[i - 1];
        double *w = &V(uri, 1, i, j, k);
        double de = w[0], pr = w[4];
        double util[4] = {0.0, w[1], w[2], w[3]};
        double bcon[4] = {0.0, w[6], w[7], w[8]};
        double bcon1[4] = {0.0, bcon[1], bcon[2], bcon[3]};
        const double *g = G16(gcov, i, j, k), *gi = G16(gcon, i, j, k);
        double g3[9], gi3[9];
        cal3met(g, gi, g3, gi3);
        double gfl = calgfl(util, g);
        double vel[4];
        vel[1] = util[1] * (1.0 / gfl);
        vel[2] = util[2] * (1.0 / gfl);
        vel[3] = util[3] * (1.0 / gfl);
        double vsq = 0.0;
        for (int m = 1; m <= 3; m++)
          for (int n = 1; n <= 3; n++) vsq = vsq + g3[(m - 1) + 3 * (n - 1)] * vel[m] * vel[n];
        if (vsq >= 1.0 || gfl > gflmax) vsq = 1.0 - 1.0 / (gflmax * gflmax);
        gfl = 1.0 / sqrt(1.0 - vsq);
        util[1] = vel[1] * gfl; util[2] = vel[2] * gfl; util[3] = vel[3] * gfl;
        gfl = calgfl(util, g);
        double ucon[4], ucov[4], bbcon[4], bbcov[4], bcov1[4];
        calucon(util, ucon, gfl, gi);
        lower(ucon, ucov, g);
        calbbcon(bcon, bbcon, ucov, ucon);
        lower(bbcon, bbcov, g);
        double bbsq = calbbsq(bbcon, bbcov);
        lower(bcon1, bcov1, g);
        double bsq = bcon1[1] * bcov1[1] + bcon1[2] * bcov1[2] + bcon1[3] * bcov1[3];
        int scaled = 0;
        int guard = 0;
        for (; guard < 100000; guard++) {
          double roh = de + (gam / (gam - 1.0)) * pr;
          double t1a = roh * gfl * gfl - pr;
          double t1b = fabs((double)0.05f * bsq);
          if (t1a <= t1b) { de = (double)1.05f * de; pr = (double)1.05f * pr; scaled = 1; continue; }
          break;
        }
        if (guard >= 100000) CNT.guard++;
        if (scaled) CNT.pmm_scale++;
        if (de != w[0]) w[0] = de;
        else if (pr != w[4]) w[4] = pr;
        double demin = P.dmin * pow(rr, -1.5);
        double prmin = (gam - 1.0) * P.pmin * pow(rr, -2.5);
        const double bbromax = 50.0, bbprmax = 2.0e3, prromax = 1.0e6;
        int hit = 0;
        if (pr < prmin) { w[4] = prmin; hit = 1; }
        if (pr > P.pmax) { w[4] = P.pmax; hit = 1; }
        if (de < demin) { w[0] = demin; hit = 1; }
        double prro = w[4] * (1.0 / w[0]);
        if (prro > prromax) { w[0] = w[4] / prromax; hit = 1; }
        double bbro = bbsq * (1.0 / w[0]);
        double bbpr = bbsq * (1.0 / w[4]);
        if (bbro > bbromax) { w[0] = bbsq / bbromax; hit = 1; }
        if (bbpr > bbprmax) { w[4] = bbsq / bbprmax; hit = 1; }
        if (hit) CNT.pmm_floor++;
      }
}

/* ------------------------------------------------------------------ */
/* cdtcfl4: cdtcfl.f90:2-129 with calcha4 (calcha.f90:129-367)          */
/* ------------------------------------------------------------------ */
static void centre_speeds(const double *w, const double *g, const double *gi, double vp[4], double vm[4]) {
  prel_t s;
  prelude(w, g, gi, &s);
  double alpha1 = 1.0 / sqrt(-GM(gi, 0, 0));
  double beta1[4];
  beta1[1] = alpha1 * alpha1 * GM(gi, 0, 1);
  beta1[2] = alpha1 * alpha1 * GM(gi, 0, 2);
  beta1[3] = alpha1 * alpha1 * GM(gi, 0, 3);
  double g3[9], gi3[9];
  cal3met(g, gi, g3, gi3);
  double vcon[4], vcov[4];
  vcon[1] = s.util[1] * (1.0 / s.gfl);
  vcon[2] = s.util[2] * (1.0 / s.gfl);
  vcon[3] = s.util[3] * (1.0 / s.gfl);
  lower1(vcon, vcov, g3);
  double vsq = 0.0;
  for (int m = 1; m <= 3; m++)
    for (int n = 1; n <= 3; n++) vsq = vsq + g3[(m - 1) + 3 * (n - 1)] * vcon[m] * vcon[n];
  double cssq = P.gam * s.pr / s.roh;
  double vasq = s.bbsq * (1.0 / (s.roh + s.bbsq));
  double omsq = vasq + cssq * (1.0 - vasq);
  double vb = vcov[1] * s.bcon[1] + vcov[2] * s.bcon[2] + vcov[3] * s.bcon[3];
  for (int d = 1; d <= 3; d++)
    speed_pair(vcon[d], vsq, omsq, cssq, vb, s.roh + s.bbsq, s.gfl, gi3[(d - 1) * 4], &vp[d], &vm[d]);
  for (int d = 1; d <= 3; d++) {
    vp[d] = alpha1 * vp[d] - beta1[d];
    vm[d] = alpha1 * vm[d] - beta1[d];
  }
  for (int d = 1; d <= 3; d++) if (fabs(vp[d]) > 1.0) vp[d] = 1.0;
  for (int d = 1; d <= 3; d++) if (fabs(vm[d]) > 1.0) vm[d] = -1.0;
}

void orc_cdtcfl4(const double *uri, const double *gcov, const double *gcon, double *dtcfl,
                 const double *dx1, const double *dx2, const double *dx3,
                 double *dt1, double *dt2, double *dt3, int nm1) {
  if (P.icha < 0 || P.icha > 2) abort();
  const int nmax = nm1 * 2 + (nm1 - 1);
  const double small = 1.0e-12;
  double dtc = 1.0, d1 = 1.0, d2 = 1.0, d3 = 1.0;
  for (int k = 2; k <= NK - 1; k++)
    for (int j = 2; j <= NJ - 1; j++)
      for (int i = 2; i <= NI - 1; i++) {
        double vp[4], vm[4];
        centre_speeds(&V(uri, 1, i, j, k), G16(gcov, i, j, k), G16(gcon, i, j, k), vp, vm);
        double vf[4], dta[4];
        for (int d = 1; d <= 3; d++) vf[d] = FMAX(FMAX(0.0, fabs(vp[d])), fabs(vm[d]));
        int nd[4] = {0, NI, NJ, NK};
        const double *dx[4] = {0, dx1, dx2, dx3};
        int idx[4] = {0, i, j, k};
        for (int d = 1; d <= 3; d++) {
          if (nd[d] > nmax && vf[d] > 0.0) {
            if (vf[d] > 1.0) vf[d] = 1.0;
            dta[d] = fabs(dx[d][idx[d] - 1] / (vf[d] + small));
          } else dta[d] = 1.0e10;
        }
        dtc = FMIN(dtc, dta[1]); dtc = FMIN(dtc, dta[2]); dtc = FMIN(dtc, dta[3]);
        d1 = FMIN(d1, dta[1]); d2 = FMIN(d2, dta[2]); d3 = FMIN(d3, dta[3]);
      }
  *dtcfl = dtc; *dt1 = d1; *dt2 = d2; *dt3 = d3;
}

/* ------------------------------------------------------------------ */
/* One time step: main.f90:743-1149                                     */
/* ------------------------------------------------------------------ */
static double *W_uo, *W_urio, *W_uh, *W_us, *W_un, *W_ww, *W_sf, *W_tenr, *W_face, *W_c;
static size_t W_N;

void orc_free_work(void) {
  free(W_uo); free(W_urio); free(W_uh); free(W_us); free(W_un); free(W_ww); free(W_sf);
  free(W_tenr); free(W_face); free(W_c);
  W_uo = W_urio = W_uh = W_us = W_un = W_ww = W_sf = W_tenr = W_face = W_c = NULL;
  W_N = 0;
}

static void work(size_t N) {
  if (W_N == N) return;
  orc_free_work();
  W_N = N;
  W_uo = calloc(9 * N, 8); W_urio = calloc(9 * N, 8); W_uh = calloc(9 * N, 8);
  W_us = calloc(9 * N, 8); W_un = calloc(9 * N, 8); W_ww = calloc(27 * N, 8);
  W_sf = calloc(4 * N, 8); W_tenr = calloc(16 * N, 8); W_face = calloc(18 * 9 * N, 8);
  W_c = calloc(6 * N, 8);
}

static void stage(orc_state *s, int st, double dt, const double *a1, const double *a2, const double *a3,
                  int nm0, int nm1) {
  const size_t N = (size_t)NI * NJ * NK;
  double *f = W_face;
  double *uriir = f, *urijr = f + 9 * N, *urikr = f + 18 * N, *uriil = f + 27 * N, *urijl = f + 36 * N,
         *urikl = f + 45 * N, *uuir = f + 54 * N, *uujr = f + 63 * N, *uukr = f + 72 * N, *uuil = f + 81 * N,
         *uujl = f + 90 * N, *uukl = f + 99 * N, *wwir = f + 108 * N, *wwjr = f + 117 * N, *wwkr = f + 126 * N,
         *wwil = f + 135 * N, *wwjl = f + 144 * N, *wwkl = f + 153 * N;
  double *c = W_c;
  orc_rec2(s->uri, s->x1, s->x2, s->x3, s->dx1b, s->dx2b, s->dx3b, uriir, urijr, urikr, uriil, urijl, urikl,
           uuir, uujr, uukr, uuil, uujl, uukl, wwir, wwjr, wwkr, wwil, wwjl, wwkl, s->gcovi, s->gcovj,
           s->gcovk, s->gconi, s->gconj, s->gconk, s->detgi, s->detgj, s->detgk, nm0);
  orc_calcha(uriir, urijr, urikr, uriil, urijl, urikl, s->gcovi, s->gcovj, s->gcovk, s->gconi, s->gconj,
             s->gconk, c, c + N, c + 2 * N, c + 3 * N, c + 4 * N, c + 5 * N, nm0);
  orc_hll(W_ww, uuir, uujr, uukr, uuil, uujl, uukl, wwir, wwjr, wwkr, wwil, wwjl, wwkl, c, c + N, c + 2 * N,
          c + 3 * N, c + 4 * N, c + 5 * N, nm0);
  orc_ct(W_ww, nm0, nm1);
  orc_caltenr(s->uri, W_tenr, s->gcov, s->gcon);
  orc_calsf(W_sf, W_tenr, s->christ, s->detg);
  double *out;
  if (st == 1) {
    out = W_uh;
    orc_rk2fst(out, W_ww, s->uu, nm1, a1, a2, a3);
    orc_rkadsf(out, W_sf, dt, nm1, 0);
  } else if (st == 2) {
    out = W_us;
    if (P.irkt == 2) { orc_rk2snd(out, W_ww, W_uo, W_uh, nm1, a1, a2, a3); orc_rkadsf(out, W_sf, dt, nm1, 1); }
    else { orc_rk3snd(out, W_ww, W_uo, W_uh, nm1, a1, a2, a3); orc_rkadsf(out, W_sf, dt, nm1, 2); }
  } else {
    out = W_un;
    orc_rk3trd(out, W_ww, W_uo, W_us, nm1, a1, a2, a3);
    orc_rkadsf(out, W_sf, dt, nm1, 3);
  }
  orc_recov(out, s->uri, s->gcov, s->gcon, s->detg, s->x1, nm1);
  orc_bnd4(s->uri, W_urio, nm1);
  if (P.model >= 7 && P.model <= 20) orc_infchk(s->uri, s->gcov, s->gcon, nm1);
  if (P.model >= 7 && P.model <= 15) orc_atmosphere1a(s->uri, s->gcov, s->gcon, s->x1);
  orc_pminmax2b(s->uri, s->gcov, s->gcon, s->x1);
  orc_caluu1(s->uri, out, s->gcov, s->gcon, s->detg);
  memcpy(s->uu, out, sizeof(double) * 9 * N);
}

double orc_step(orc_state *s, double *time) {
  const size_t N = (size_t)NI * NJ * NK;
  work(N);
  int nm0 = (P.irec == 5) ? 4 : (P.irec == 1) ? 2 : 3; /* main.f90:179-211 */
  int nm1 = nm0 + 1;
  memcpy(W_uo, s->uu, sizeof(double) * 9 * N);
  memcpy(W_urio, s->uri, sizeof(double) * 9 * N);
  double dtcfl, d1, d2, d3;
  orc_cdtcfl4(s->uri, s->gcov, s->gcon, &dtcfl, s->dx1b, s->dx2b, s->dx3b, &d1, &d2, &d3, nm1);
  double dt = P.cfl * dtcfl;
  double *a1 = malloc(sizeof(double) * NI), *a2 = malloc(sizeof(double) * NJ), *a3 = malloc(sizeof(double) * NK);
  for (int i = 1; i <= NI - 1; i++) a1[i - 1] = dt / s->dx1b[i - 1];
  for (int j = 1; j <= NJ - 1; j++) a2[j - 1] = dt / s->dx2b[j - 1];
  for (int k = 1; k <= NK - 1; k++) a3[k - 1] = dt / s->dx3b[k - 1];
  *time = *time + dt;
  stage(s, 1, dt, a1, a2, a3, nm0, nm1);
  stage(s, 2, dt, a1, a2, a3, nm0, nm1);
  if (P.irkt == 3) stage(s, 3, dt, a1, a2, a3, nm0, nm1);
  free(a1); free(a2); free(a3);
  return dt;
}
