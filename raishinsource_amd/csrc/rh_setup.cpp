// Host-side set-up of a model: grid, metric, Christoffel symbols and the
// magnetised Fishbone-Moncrief torus.  These produce the inputs of the time
// step; they run once per model on the host, in the reference's layouts.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/raishin_hip.h"
#include "../../include/raishin_setup.h"
#include "rh_metric.hpp"
#include "rh_torus.hpp"

namespace {

constexpr double PI = 3.14159265358979323;  // pram.f90:31

inline size_t cell(const raishin_pram& p, int i, int j, int k) {  // 1-based
  return (size_t)(i - 1) + (size_t)p.imax * ((size_t)(j - 1) + (size_t)p.jmax * (size_t)(k - 1));
}
#define GM(g, m, n) (g)[(m) + 4 * (n)]

// sin and cos of the folded theta, clamped away from zero, as kshcov1 /
// kshcon1 take them (coord.f90:535-690)
void ks_trig(double x3, double& sth, double& cth) {
  const double small = 1.0e-12;
  const double theta = (x3 > PI) ? 2.0 * PI - x3 : x3;
  sth = std::sin(theta);
  if (std::fabs(sth) < small) sth = (sth > 0.0) ? small : -small;
  cth = std::cos(theta);
  if (std::fabs(cth) < small) cth = (cth > 0.0) ? small : -small;
}
rhmet::ThetaTrig theta_trig(double th) {
  const double delta = 1.0e-5;
  rhmet::ThetaTrig t;
  ks_trig(th, t.s, t.c);
  ks_trig(th + delta, t.sa, t.ca);
  ks_trig(th - delta, t.sb, t.cb);
  return t;
}
void metric_cov(const raishin_pram& p, double* g, double r, double th) {
  double s = 0.0, c = 0.0;
  if (p.metric == 303) ks_trig(th, s, c);
  rhmet::cov(p.metric, p.akm, g, r, s, c);
}
void metric_con(const raishin_pram& p, double* g, double r, double th) {
  double s = 0.0, c = 0.0;
  if (p.metric == 303) ks_trig(th, s, c);
  rhmet::con(p.metric, p.akm, g, r, s, c);
}

// callfish: mdgrmhd.f90:808-834 (angular momentum of the FM torus)
double callfish(double akm, double rmax) {
  const double t1a = ((akm * akm) - 2.0 * akm * std::sqrt(rmax) + (rmax * rmax));
  const double t1b = (-2.0 * akm * rmax) * t1a;
  const double t1c = std::sqrt(2.0 * akm * std::sqrt(rmax) + (-3.0 + rmax) * rmax);
  const double t1d = (akm + (-2.0 + rmax) * std::sqrt(rmax));
  const double t1e = ((rmax * rmax * rmax) + (akm * akm) * (2.0 + rmax));
  const double t1f = std::sqrt(1.0 + (2.0 * akm / std::pow(rmax, 1.5)) - (3.0 / rmax));
  const double t2a = (rmax * rmax * rmax) * std::sqrt(2.0 * akm * std::sqrt(rmax) + (-3.0 + rmax) * rmax);
  const double t2b = ((akm * akm) + (-2.0 + rmax) * rmax);
  return (t1a * ((t1b / t1c) + t1d * t1e / t1f)) / (t2a * t2b);
}

inline double splitmix_uniform(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

}  // namespace

namespace rhm {
// The torus constants both set-ups use (evaluated on the host, glibc): the
// Fishbone-Moncrief l at r_max (callfish) and the inner-edge terms of ln h.
rhtorus::Consts torus_consts(const raishin_pram& p, const raishin_torus& t) {
  const double gam = p.gam, akm = p.akm, kpol1 = 1.0e-3;
  const double rin = t.rin, rmax = t.rmax;
  const double thedge = (double)0.02f * PI, thedge1 = (double)0.98f * PI, small = 1.0e-12;
  const double clm = callfish(akm, rmax);
  const double delin = rin * rin - 2.0 * rin + akm * akm;
  const double aain = (rin * rin + akm * akm) * (rin * rin + akm * akm) - delin * akm * akm;
  const double sigin = rin * rin;
  const double t2a = 4.0 * (clm * clm * sigin * sigin) * delin * (1.0 / (aain * aain));
  const double t2c = sigin * delin * (1.0 / aain);
  const double t2d = 0.5 * std::log((1.0 + std::sqrt(1.0 + t2a)) * (1.0 / t2c));
  const double t2e = 0.5 * std::sqrt(1.0 + t2a);
  const double t2f = 2.0 * akm * rin * clm * (1.0 / aain);
  return rhtorus::Consts{gam, akm, kpol1, rin, clm, t2d, t2e, t2f, thedge, thedge1, small};
}
}  // namespace rhm
using rhm::torus_consts;

static int mdtorus1_impl(const raishin_pram* pp, const raishin_torus* tp, double* uri, const double* gcov,
                         const double* gcon, const double* detg, const double* x1, const double* x2,
                         const double* x3, bool slice) {
  // the metric arrays hold every cell, or (slice) the one phi row of an
  // axisymmetric metric: (4,4,imax,kmax) and (imax,kmax)
  auto mq = [&](int i, int j, int k) -> size_t {
    return slice ? (size_t)(i - 1) + (size_t)pp->imax * (size_t)(k - 1) : cell(*pp, i, j, k);
  };
  // Fishbone-Moncrief torus with a single poloidal loop, after mdtorus1
  // (mdgrmhd.f90:879-1654): atmosphere, torus, density normalisation, seeded
  // pressure noise, A_phi = max(rho/rho_max - 0.2, 0), field scaled to beta_min.
  const raishin_pram& p = *pp;
  const raishin_torus& t = *tp;
  (void)x2;
  const double gam = p.gam;
  const int ni = p.imax, nj = p.jmax, nk = p.kmax;
  const size_t N = (size_t)ni * nj * nk;
  std::vector<double> alnh(N, 0.0);
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) {
        double* w = &uri[9 * cell(p, i, j, k)];
        const double rr = x1[i - 1];
        const double de = p.dmin * std::pow(rr, -1.5), pr = p.pmin * (gam - 1.0) * std::pow(rr, -2.5);
        w[0] = de; w[1] = w[2] = w[3] = 0.0; w[4] = pr; w[5] = pr * (1.0 / std::pow(de, gam - 1.0));
        w[6] = w[7] = w[8] = 0.0;
      }
  if (p.metric != 303) return 0;
  const rhtorus::Consts tc = torus_consts(p, t);
  double rhomax = 0.0;
#pragma omp parallel for schedule(dynamic) reduction(max : rhomax)
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) {
        const double rr = x1[i - 1], th = x3[k - 1];
        double sth = std::sin(th);
        if (std::fabs(sth) < tc.small) sth = (sth >= 0.0) ? tc.small : -tc.small;
        double s_ks, c_ks;
        ks_trig(th, s_ks, c_ks);
        const rhtorus::Column col = rhtorus::column(tc, rr, th, sth, std::cos(th), s_ks, c_ks);
        alnh[cell(p, i, j, k)] = col.al;
        if (col.in) {
          double* w = &uri[9 * cell(p, i, j, k)];
          w[0] = col.de; w[1] = col.u[1]; w[2] = col.u[2]; w[3] = col.u[3]; w[4] = col.pr;
          w[5] = col.pr * (1.0 / std::pow(col.de, gam - 1.0));
          if (col.de > rhomax) rhomax = col.de;
        }
      }
  if (rhomax <= 0.0) return -1;
  // normalise density to 1 and perturb the pressure.  The noise is drawn in the
  // reference's (k,j,i) order from a seeded splitmix64 stream; with one interior
  // phi cell the model is axisymmetric and one value is drawn per (i,k).
  const int nm0 = (p.irec == 5) ? 4 : (p.irec == 1 || p.irec == 2 || p.irec == 3) ? 2 : 3;  // main.f90:179-211
  const bool axisym = (nj - 2 * (nm0 + 1)) <= 1;
  uint64_t seed = (uint64_t)t.seed;
  std::vector<double> noise((size_t)ni * nk, 0.0);
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) {
        const double rr = x1[i - 1], th = x3[k - 1];
        const double rhomin1 = p.dmin * std::pow(rr, -1.5), prmin1 = p.pmin * (gam - 1.0) * std::pow(rr, -2.5);
        const double hm1 = std::exp(alnh[cell(p, i, j, k)]) - 1.0;
        if (!(hm1 > 0.0 && rr >= tc.rin && th >= tc.thedge && th <= tc.thedge1)) continue;
        double* w = &uri[9 * cell(p, i, j, k)];
        double de = w[0] * (1.0 / rhomax), pr = w[4] * (1.0 / rhomax);
        double u;
        if (axisym) {
          double& nz = noise[(size_t)(i - 1) + (size_t)ni * (k - 1)];
          if (j == 1) nz = splitmix_uniform(seed);
          u = nz;
        } else {
          u = splitmix_uniform(seed);
        }
        pr = pr * (1.0 + t.noise * (u - 0.5));
        if (de <= rhomin1 || pr <= prmin1) { de = rhomin1; w[1] = w[2] = w[3] = 0.0; pr = prmin1; }
        w[0] = de; w[4] = pr; w[5] = pr * (1.0 / std::pow(de, gam - 1.0));
      }
  // vector potential at cell corners and B = curl A / sqrt(-g): mdgrmhd.f90:1350-1460
  std::vector<double> pom(N, 0.0);
  for (int k = 2; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 2; i <= ni; i++) {
        const double rhoav = 0.25 * (uri[9 * cell(p, i, j, k)] + uri[9 * cell(p, i - 1, j, k)] +
                                     uri[9 * cell(p, i, j, k - 1)] + uri[9 * cell(p, i - 1, j, k - 1)]);
        const double pm = (rhoav != 0.0) ? rhoav - t.acut : 0.0;
        if (pm > 0.0) pom[cell(p, i, j, k)] = pm;
      }
  for (int j = 1; j <= nj; j++)
    for (int i = 1; i <= ni; i++) pom[cell(p, i, j, 1)] = pom[cell(p, i, j, 2)];
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++) pom[cell(p, 1, j, k)] = pom[cell(p, 2, j, k)];
  for (int k = 2; k <= nk - 1; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 2; i <= ni - 1; i++) {
        const size_t c = cell(p, i, j, k);
        const double ta = 0.5 * (pom[cell(p, i + 1, j, k + 1)] + pom[cell(p, i, j, k + 1)] - pom[cell(p, i + 1, j, k)] - pom[c]);
        const double dx3a = 0.5 * (x3[k] + x3[k - 1]) - 0.5 * (x3[k - 1] + x3[k - 2]);
        const double tb = 0.5 * (pom[cell(p, i + 1, j, k + 1)] + pom[cell(p, i + 1, j, k)] - pom[cell(p, i, j, k + 1)] - pom[c]);
        const double dx1a = 0.5 * (x1[i] + x1[i - 1]) - 0.5 * (x1[i - 1] + x1[i - 2]);
        uri[9 * c + 6] = (ta != 0.0) ? -ta * (1.0 / (dx3a * detg[mq(i, j, k)])) : 0.0;
        uri[9 * c + 8] = (tb != 0.0) ? tb * (1.0 / (dx1a * detg[mq(i, j, k)])) : 0.0;
      }
  for (int j = 1; j <= nj; j++)
    for (int i = 1; i <= ni; i++)
      for (int q : {6, 8}) {
        uri[9 * cell(p, i, j, 1) + q] = uri[9 * cell(p, i, j, 2) + q];
        uri[9 * cell(p, i, j, nk) + q] = uri[9 * cell(p, i, j, nk - 1) + q];
      }
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int q : {6, 8}) {
        uri[9 * cell(p, 1, j, k) + q] = uri[9 * cell(p, 2, j, k) + q];
        uri[9 * cell(p, ni, j, k) + q] = uri[9 * cell(p, ni - 1, j, k) + q];
      }
  // beta normalisation: B *= sqrt(beta_act / beta), beta_act = min 2p/b^2
  double bmin = 1.0e5;
#pragma omp parallel for schedule(static) reduction(min : bmin)
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) {
        const size_t c = cell(p, i, j, k), q = mq(i, j, k);
        const double b2 = rhtorus::bsq(&uri[9 * c], &gcov[16 * q], &gcon[16 * q]);
        const double pb = (b2 != 0.0) ? 2.0 * uri[9 * c + 4] / b2 : 1.0e5;
        if (pb < bmin) bmin = pb;
      }
  const double bnorm = (t.beta == 0.0) ? 0.0 : std::sqrt(bmin / t.beta);
  for (size_t c = 0; c < N; c++)
    for (int q = 6; q <= 8; q++) uri[9 * c + q] = uri[9 * c + q] * bnorm;
  return 0;
}

namespace rhm {
// Per k the sin/cos pairs the set-up's metric takes (folded, clamped): at
// x3(k), x3(k)+1e-5, x3(k)-1e-5 and x3a(k), 8 values per k.  The device table
// kernel starts from these, so its tables equal raishin_coord1's and
// raishin_calchrist's.
void theta_trig_table(const raishin_pram& p, const double* x3, const double* x3a, double* out) {
  for (int k = 0; k < p.kmax; k++) {
    const rhmet::ThetaTrig t = theta_trig(x3[k]);
    double* o = out + 8 * k;
    o[0] = t.s; o[1] = t.c; o[2] = t.sa; o[3] = t.ca; o[4] = t.sb; o[5] = t.cb;
    ks_trig(x3a[k], o[6], o[7]);
  }
}
}  // namespace rhm

extern "C" {

void raishin_grid(const raishin_pram* pp, const raishin_grid_ext* ext, double* x1, double* x2, double* x3,
                  double* dx1, double* dx2, double* dx3, double* dx1b, double* dx2b, double* dx3b, double* x1a,
                  double* x2a, double* x3a, double* x1b, double* x2b, double* x3b) {
  const raishin_pram& p = *pp;
  const int n[3] = {p.imax, p.jmax, p.kmax};
  const int ix[3] = {ext->ix1, 1, 1};
  const double lo[3] = {ext->xmin, ext->ymin, ext->zmin}, hi[3] = {ext->xmax, ext->ymax, ext->zmax};
  double* X[3] = {x1, x2, x3};
  double* DX[3] = {dx1, dx2, dx3};
  double* DXB[3] = {dx1b, dx2b, dx3b};
  double* XA[3] = {x1a, x2a, x3a};
  double* XB[3] = {x1b, x2b, x3b};
  for (int a = 0; a < 3; a++) {
    const int nn = n[a];
    double *x = X[a], *dx = DX[a], *xa = XA[a], *xb = XB[a], *dxb = DXB[a];
    if (ix[a] == 2) {
      // logarithmic radius: grid.f90:31-50; 1./float(imax-1) is single precision
      const double xx0 = 0.0;
      const double s0 = std::log(lo[a] - xx0), s1 = std::log(hi[a] - xx0);
      const double dxt = (s1 - s0) * (double)(1.0f / (float)(nn - 1));
      for (int i = 1; i <= nn; i++) {
        const double xt = s0 + (double)(float)(i - 1) * dxt;
        x[i - 1] = std::exp(xt) + xx0;
        xb[i - 1] = xt;
      }
      for (int i = 2; i <= nn; i++) dx[i - 1] = x[i - 1] - x[i - 2];
      dx[0] = dx[1];
    } else {
      // uniform: grid.f90:22-28, 146-152, 175-181
      for (int i = 1; i <= nn; i++) {
        dx[i - 1] = (hi[a] - lo[a]) / (double)(float)(nn - 1);
        x[i - 1] = lo[a] + dx[i - 1] * (double)(float)(i - 1);
        xb[i - 1] = x[i - 1];
      }
    }
    for (int i = 1; i <= nn - 1; i++) xa[i - 1] = 0.5 * (x[i - 1] + x[i]);
    xa[nn - 1] = xa[nn - 2] + dx[nn - 1];
    // dx*b(i) = x*a(i) - x*a(i-1), i = 2..n-1 kept (grid.f90:103-108 writes one past the end)
    for (int i = 2; i <= nn - 1; i++) dxb[i - 1] = xa[i - 1] - xa[i - 2];
    dxb[0] = (nn > 2) ? dxb[1] : xa[1] - xa[0];
  }
}

void raishin_coord1(const raishin_pram* pp, double* gcov, double* gcovi, double* gcovj, double* gcovk, double* gcon,
                    double* gconi, double* gconj, double* gconk, double* detg, double* detgi, double* detgj,
                    double* detgk, const double* x1, const double* x2, const double* x3, const double* x1a,
                    const double* x2a, const double* x3a) {
  // coord.f90:5-146 (metric 1 and 303) and caldetg (coord.f90:1016-1086)
  // The metrics built here (303, 1) have no phi dependence: each (i,k) is
  // computed once and copied to every j.  Columns k are independent.
  const raishin_pram& p = *pp;
  (void)x2; (void)x2a;
#pragma omp parallel for schedule(dynamic)
  for (int k = 1; k <= p.kmax; k++)
    for (int i = 1; i <= p.imax; i++) {
      double g[16], gi[16], gk[16], h[16], hi[16], hk[16];
      metric_cov(p, g, x1[i - 1], x3[k - 1]);
      metric_cov(p, gi, x1a[i - 1], x3[k - 1]);
      metric_cov(p, gk, x1[i - 1], x3a[k - 1]);
      metric_con(p, h, x1[i - 1], x3[k - 1]);
      metric_con(p, hi, x1a[i - 1], x3[k - 1]);
      metric_con(p, hk, x1[i - 1], x3a[k - 1]);
      for (int j = 1; j <= p.jmax; j++) {
        const size_t c = cell(p, i, j, k);
        std::memcpy(&gcov[16 * c], g, 128); std::memcpy(&gcovi[16 * c], gi, 128);
        std::memcpy(&gcovj[16 * c], g, 128); std::memcpy(&gcovk[16 * c], gk, 128);
        std::memcpy(&gcon[16 * c], h, 128); std::memcpy(&gconi[16 * c], hi, 128);
        std::memcpy(&gconj[16 * c], h, 128); std::memcpy(&gconk[16 * c], hk, 128);
      }
    }
  double* D[4] = {detg, detgi, detgj, detgk};
  const double* Gs[4] = {gcov, gcovi, gcovj, gcovk};
  const size_t N = (size_t)p.imax * p.jmax * p.kmax;
  for (int s = 0; s < 4; s++)
#pragma omp parallel for schedule(static)
    for (long long cc = 0; cc < (long long)N; cc++) D[s][cc] = rhmet::detg(&Gs[s][16 * cc]);
}

void raishin_calchrist(const raishin_pram* pp, double* christ, const double* x1, const double* x2, const double* x3) {
  // calchrist: coord.f90:1089-1321 (finite-difference metric derivatives, delta=1e-5)
  // Computed for j = 1 and copied to every j: the metrics built here have no
  // phi dependence, and the phi derivative's two metric evaluations are equal
  // (its difference is zero for every phi).
  const raishin_pram& p = *pp;
#pragma omp parallel for schedule(dynamic)
  for (int k = 1; k <= p.kmax; k++) {
    const rhmet::ThetaTrig tt = theta_trig(x3[k - 1]);
    for (int i = 1; i <= p.imax; i++) {
      double* ch = &christ[64 * cell(p, i, 1, k)];
      rhmet::christoffel(p.metric, p.akm, x1[i - 1], x3[k - 1], x2[0], tt, ch);
      for (int jj = 2; jj <= p.jmax; jj++) std::memcpy(&christ[64 * cell(p, i, jj, k)], ch, 64 * sizeof(double));
    }
  }
}

int raishin_mdtorus1(const raishin_pram* pp, const raishin_torus* tp, double* uri, const double* gcov,
                     const double* gcon, const double* detg, const double* x1, const double* x2, const double* x3) {
  return mdtorus1_impl(pp, tp, uri, gcov, gcon, detg, x1, x2, x3, false);
}

int raishin_mdtorus1_ik(const raishin_pram* pp, const raishin_torus* tp, double* uri, const double* gcov,
                        const double* gcon, const double* detg, const double* x1, const double* x2, const double* x3) {
  return mdtorus1_impl(pp, tp, uri, gcov, gcon, detg, x1, x2, x3, true);
}

}  // extern "C"

int raishin_mdshoctub(const raishin_pram* pp, const raishin_tube* t, double* uri) {
  // mdshoctub (mdgrmhd.f90:55-117) with the two states as parameters: cells
  // i <= imax/2 take the left state, the rest the right; flat metric, so
  // u~^i = gamma v^i.  uri(6) = p / rho**(gam-1) (mdgrmhd.f90:103).
  const raishin_pram& p = *pp;
  double w[2][9];
  for (int s = 0; s < 2; s++) {
    const double vsq = t->v[s][0] * t->v[s][0] + t->v[s][1] * t->v[s][1] + t->v[s][2] * t->v[s][2];
    if (!(t->de[s] > 0.0) || !(t->pr[s] > 0.0) || !(vsq < 1.0)) return 1;
    const double lor = 1.0 / std::sqrt(1.0 - vsq);
    w[s][0] = t->de[s];
    for (int a = 0; a < 3; a++) w[s][1 + a] = lor * t->v[s][a];
    w[s][4] = t->pr[s];
    w[s][5] = t->pr[s] * (1.0 / std::pow(t->de[s], p.gam - 1.0));
    for (int a = 0; a < 3; a++) w[s][6 + a] = t->b[s][a];
  }
  for (int k = 1; k <= p.kmax; k++)
    for (int j = 1; j <= p.jmax; j++)
      for (int i = 1; i <= p.imax; i++) {
        const double* ws = w[i <= p.imax / 2 ? 0 : 1];
        const size_t c = cell(p, i, j, k);
        for (int n = 0; n < 9; n++) uri[9 * c + n] = ws[n];
      }
  return 0;
}

