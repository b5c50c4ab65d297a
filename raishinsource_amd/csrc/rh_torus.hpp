// The magnetised Fishbone-Moncrief torus, per (r, theta) column, after
// mdgrmhd.f90:879-1654 (mdtorus1) with physfunc.f90:366-505 (transbl2ks) and
// coord.f90:369-450 (kercov1): shared by the host set-up (rh_setup.cpp) and
// the device set-up (rh_torus.hip).  The arithmetic is the same expression for
// expression; sin and cos come from the caller (the host), so the two differ
// only where the device's log, exp and pow round differently from glibc's.
#ifndef RH_TORUS_HPP
#define RH_TORUS_HPP

#include <cmath>

#include "rh_metric.hpp"

namespace rhtorus {

#define RHT_G(g, m, n) (g)[(m) + 4 * (n)]

// What a column needs besides its position: the FM constant l (callfish) and
// the inner-edge terms of ln h (mdgrmhd.f90), all computed once on the host.
struct Consts {
  double gam, akm, kpol1, rin, clm, t2d, t2e, t2f, thedge, thedge1, small;
};

struct Column {
  double al;        // ln h - ln h(r_in), 0 outside
  double de, pr;    // polytrope density and pressure (before normalisation)
  double u[4];      // KS u~^i (1..3)
  bool in;          // a torus cell
};

// kercov1 (coord.f90:369-450) from the sin/cos of the folded theta, clamped
RH_HD void bl_cov(double akm, double* g, double rr, double sth, double cth) {
  const double s2 = sth * sth;
  const double a2 = akm * akm, r2 = rr * rr;
  const double del = 1.0 - (2.0 / rr) + (a2 * (1.0 / r2));
  const double sig = 1.0 + a2 * cth * cth * (1.0 / r2);
  for (int e = 0; e < 16; e++) g[e] = 0.0;
  RHT_G(g, 0, 0) = -(1.0 - 2.0 * (1.0 / (rr * sig)));
  RHT_G(g, 0, 2) = -2.0 * akm * s2 * (1.0 / (rr * sig));
  RHT_G(g, 1, 1) = sig / del;
  RHT_G(g, 2, 0) = RHT_G(g, 0, 2);
  RHT_G(g, 2, 2) = r2 * s2 * (1.0 + (a2 * (1.0 / r2)) + 2.0 * a2 * s2 * (1.0 / (r2 * rr * sig)));
  RHT_G(g, 3, 3) = r2 * sig;
}

// transbl2ks (physfunc.f90:366-505, metric 303): BL u~ -> KS u~, with the
// folded, clamped sin/cos of theta
RH_HD void bl2ks(double akm, const double ut[4], double out[4], double rr, double sth, double cth) {
  const double rbh1 = 1.0 + sqrt(1.0 - akm * akm);
  double gi[16];
  rhmet::ks_con(akm, gi, rr, sth, cth);
  const double alpha1 = 1.0 / sqrt(-RHT_G(gi, 0, 0));
  const double beta1[4] = {0.0, alpha1 * alpha1 * RHT_G(gi, 0, 1), alpha1 * alpha1 * RHT_G(gi, 0, 2),
                           alpha1 * alpha1 * RHT_G(gi, 0, 3)};
  if (rr > rbh1) {
    double g[16];
    bl_cov(akm, g, rr, sth, cth);
    double ucon[4] = {0.0, ut[1], ut[2], ut[3]};
    const double aa = RHT_G(g, 0, 0);
    const double bb = 2.0 * (RHT_G(g, 0, 1) * ucon[1] + RHT_G(g, 0, 2) * ucon[2] + RHT_G(g, 0, 3) * ucon[3]);
    const double cc = 1.0 + RHT_G(g, 1, 1) * ucon[1] * ucon[1] + RHT_G(g, 2, 2) * ucon[2] * ucon[2] +
                      RHT_G(g, 3, 3) * ucon[3] * ucon[3] +
                      2.0 * (RHT_G(g, 1, 2) * ucon[1] * ucon[2] + RHT_G(g, 1, 3) * ucon[1] * ucon[3] +
                             RHT_G(g, 2, 3) * ucon[2] * ucon[3]);
    const double discr = bb * bb - 4.0 * aa * cc;
    ucon[0] = (-bb - sqrt(discr)) * (1.0 / (2.0 * aa));
    double tr[16] = {0.0};
    const double tmp2 = rr * rr - (2.0 * rr) + akm * akm;
    if (tmp2 > 0.0) {
      RHT_G(tr, 0, 0) = 1.0; RHT_G(tr, 0, 1) = 2.0 * rr * (1.0 / tmp2); RHT_G(tr, 1, 1) = 1.0;
      RHT_G(tr, 2, 2) = 1.0; RHT_G(tr, 2, 1) = akm * (1.0 / tmp2); RHT_G(tr, 3, 3) = 1.0;
    }
    double u1[4] = {0.0, 0.0, 0.0, 0.0};
    for (int n = 0; n < 4; n++)
      for (int m = 0; m < 4; m++) u1[m] = u1[m] + RHT_G(tr, m, n) * ucon[n];
    const double gfl = u1[0] * alpha1;
    for (int m = 1; m <= 3; m++) out[m] = u1[m] + gfl * beta1[m] * (1.0 / alpha1);
  } else {
    for (int m = 1; m <= 3; m++) out[m] = beta1[m] * (1.0 / alpha1);
  }
}

// One (r, theta) column: rr = x1(i), th = x3(k); sin_th (clamped) and cos_th
// of theta itself, s_ks / c_ks the folded, clamped pair the metrics take.
RH_HD Column column(const Consts& c, double rr, double th, double sin_th, double cos_th, double s_ks, double c_ks) {
  Column o;
  o.de = o.pr = 0.0;
  o.u[0] = o.u[1] = o.u[2] = o.u[3] = 0.0;
  const double akm = c.akm, clm = c.clm, gam = c.gam;
  const double sth = sin_th, cth = cos_th;
  const double del = (rr * rr) - (2.0 * rr) + (akm * akm);
  const double aa = (rr * rr + akm * akm) * (rr * rr + akm * akm) - del * akm * akm * sth * sth;
  const double sig = rr * rr + akm * akm * cth * cth;
  double al = 0.0;
  const bool box = rr >= c.rin && th >= c.thedge && th <= c.thedge1;
  if (box) {
    const double t1a = 4.0 * (clm * clm * sig * sig) * del * (1.0 / (aa * sth * aa * sth));
    const double t1c = sig * del * (1.0 / aa);
    const double t1d = 0.5 * log((1.0 + sqrt(1.0 + t1a)) * (1.0 / t1c));
    const double t1e = 0.5 * sqrt(1.0 + t1a);
    const double t1f = 2.0 * akm * rr * clm * (1.0 / aa);
    al = t1d - t1e - t1f - (c.t2d - c.t2e - c.t2f);
    if (al <= 0.0) al = 0.0;
  }
  o.al = al;
  const double hm1 = exp(al) - 1.0;
  o.in = hm1 > 0.0 && box;
  if (o.in) {
    o.de = pow(hm1 * (gam - 1.0) / (c.kpol1 * gam), 1.0 / (gam - 1.0));
    o.pr = c.kpol1 * pow(o.de, gam);
    const double expm2chi = sig * sig * del * (1.0 / (aa * aa * sth * sth));
    const double up1 = sqrt((-1.0 + sqrt(1.0 + 4.0 * clm * clm * expm2chi)) / 2.0);
    const double t3a = 2.0 * akm * rr * sqrt(1.0 + up1 * up1);
    const double t3b = sqrt(aa * sig * del);
    const double t3c = sqrt(sig * (1.0 / aa)) * up1 * (1.0 / sth);
    const double ut[4] = {0.0, 0.0, t3a * (1.0 / t3b) + t3c, 0.0};
    bl2ks(akm, ut, o.u, rr, s_ks, c_ks);
  }
  return o;
}

// b^2 of primitives w with the metric g, gi at the cell (physfunc.f90 helpers),
// for the beta normalisation
RH_HD double bsq(const double* w, const double* g, const double* gi) {
  double ut[4] = {0.0, w[1], w[2], w[3]};
  double utsq = 0.0;
  for (int m = 1; m <= 3; m++)
    for (int n = 1; n <= 3; n++) utsq = utsq + RHT_G(g, m, n) * ut[m] * ut[n];
  if (utsq < 0.0) utsq = (fabs(utsq) <= 1.0e-10) ? fabs(utsq) : 1.0e-10;
  if (fabs(utsq) > 5.0e2) utsq = 5.0e2;
  const double gfl = sqrt(1.0 + utsq);
  const double alpha1 = 1.0 / sqrt(-RHT_G(gi, 0, 0));
  double ucon[4], ucov[4], bb[4], bbcov[4];
  ucon[0] = gfl * (1.0 / alpha1);
  for (int m = 1; m <= 3; m++) ucon[m] = ut[m] - gfl * (alpha1 * alpha1 * RHT_G(gi, 0, m)) * (1.0 / alpha1);
  for (int m = 0; m < 4; m++)
    ucov[m] = RHT_G(g, m, 0) * ucon[0] + RHT_G(g, m, 1) * ucon[1] + RHT_G(g, m, 2) * ucon[2] + RHT_G(g, m, 3) * ucon[3];
  const double b[4] = {0.0, w[6], w[7], w[8]};
  bb[0] = b[1] * ucov[1] + b[2] * ucov[2] + b[3] * ucov[3];
  for (int m = 1; m <= 3; m++) bb[m] = (b[m] + bb[0] * ucon[m]) * (1.0 / ucon[0]);
  for (int m = 0; m < 4; m++)
    bbcov[m] = RHT_G(g, m, 0) * bb[0] + RHT_G(g, m, 1) * bb[1] + RHT_G(g, m, 2) * bb[2] + RHT_G(g, m, 3) * bb[3];
  return bb[0] * bbcov[0] + bb[1] * bbcov[1] + bb[2] * bbcov[2] + bb[3] * bbcov[3];
}

// The n-th value (n >= 1) of the seeded splitmix64 stream the noise draws
// from: the stream adds its increment once per draw, so any draw is reachable
// directly.
RH_HD double splitmix_nth(unsigned long long seed, unsigned long long n) {
  unsigned long long z = seed + n * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

#undef RHT_G

}  // namespace rhtorus

#endif
