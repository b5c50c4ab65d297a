// The axisymmetric metrics of the set-up (metric 303 Kerr-Schild, metric 1
// Minkowski) as arithmetic on (r, sin theta, cos theta): the host set-up and
// the device table kernel both call these, and only the host evaluates sin
// and cos, so the two produce the same bits.  +, -, *, / and sqrt are exact
// (correctly rounded) in both builds, which disable contraction.
#ifndef RH_METRIC_HPP
#define RH_METRIC_HPP

#ifdef __HIPCC__
#define RH_HD __host__ __device__ inline
#else
#define RH_HD inline
#endif

namespace rhmet {

#define RHM_G(g, m, n) (g)[(m) + 4 * (n)]

// kshcov1 (coord.f90:535-613) with sin/cos of the folded theta already taken
// and clamped away from zero
RH_HD void ks_cov(double akm, double* g, double rr, double sth, double cth) {
  const double s2 = sth * sth;
  const double a2 = akm * akm, r2 = rr * rr;
  const double del = r2 - (2.0 * rr) + a2;
  const double sig = r2 + a2 * cth * cth;
  const double aa = ((r2 + a2) * (r2 + a2)) - a2 * del * s2;
  for (int e = 0; e < 16; e++) g[e] = 0.0;
  RHM_G(g, 0, 0) = -1.0 + 2.0 * rr / sig;
  RHM_G(g, 0, 1) = 2.0 * rr / sig;
  RHM_G(g, 0, 2) = -2.0 * akm * rr * s2 * (1.0 / sig);
  RHM_G(g, 1, 0) = RHM_G(g, 0, 1);
  RHM_G(g, 1, 1) = 1.0 + 2.0 * rr / sig;
  RHM_G(g, 1, 2) = -1.0 * akm * s2 * (1.0 + 2.0 * rr / sig);
  RHM_G(g, 2, 0) = RHM_G(g, 0, 2);
  RHM_G(g, 2, 1) = RHM_G(g, 1, 2);
  RHM_G(g, 2, 2) = aa * s2 * (1.0 / sig);
  RHM_G(g, 3, 3) = sig;
}

// kshcon1 (coord.f90:616-690)
RH_HD void ks_con(double akm, double* g, double rr, double sth, double cth) {
  const double s2 = sth * sth;
  const double a2 = akm * akm, r2 = rr * rr;
  const double del = r2 - (2.0 * rr) + a2;
  const double sig = r2 + a2 * cth * cth;
  for (int e = 0; e < 16; e++) g[e] = 0.0;
  RHM_G(g, 0, 0) = -1.0 - 2.0 * rr / sig;
  RHM_G(g, 0, 1) = 2.0 * rr / sig;
  RHM_G(g, 1, 0) = RHM_G(g, 0, 1);
  RHM_G(g, 1, 1) = del / sig;
  RHM_G(g, 1, 2) = akm * (1.0 / sig);
  RHM_G(g, 2, 1) = RHM_G(g, 1, 2);
  RHM_G(g, 2, 2) = 1.0 / (sig * s2);
  RHM_G(g, 3, 3) = 1.0 / sig;
}

// mincov1 / mincon1: coord.f90:149-198
RH_HD void minkowski(double* g, double sgn00) {
  for (int e = 0; e < 16; e++) g[e] = 0.0;
  RHM_G(g, 0, 0) = sgn00; RHM_G(g, 1, 1) = 1.0; RHM_G(g, 2, 2) = 1.0; RHM_G(g, 3, 3) = 1.0;
}

RH_HD void cov(int metric, double akm, double* g, double r, double sth, double cth) {
  if (metric == 303) ks_cov(akm, g, r, sth, cth);
  else minkowski(g, -1.0);
}
RH_HD void con(int metric, double akm, double* g, double r, double sth, double cth) {
  if (metric == 303) ks_con(akm, g, r, sth, cth);
  else minkowski(g, -1.0);
}

// caldetg (coord.f90:1016-1086): sqrt(-det g) by cofactors of the first column
RH_HD double detg(const double* g) {
  double aa, bb, cc, pp, qq, rr, xx, yy, zz;
  aa = RHM_G(g, 1, 1); bb = RHM_G(g, 1, 2); cc = RHM_G(g, 1, 3); pp = RHM_G(g, 2, 1); qq = RHM_G(g, 2, 2);
  rr = RHM_G(g, 2, 3); xx = RHM_G(g, 3, 1); yy = RHM_G(g, 3, 2); zz = RHM_G(g, 3, 3);
  const double t1 = aa * qq * zz + bb * rr * xx + cc * pp * yy - cc * qq * xx - bb * pp * zz - aa * rr * yy;
  aa = RHM_G(g, 0, 1); bb = RHM_G(g, 0, 2); cc = RHM_G(g, 0, 3);
  const double t2 = aa * qq * zz + bb * rr * xx + cc * pp * yy - cc * qq * xx - bb * pp * zz - aa * rr * yy;
  pp = RHM_G(g, 1, 1); qq = RHM_G(g, 1, 2); rr = RHM_G(g, 1, 3);
  const double t3 = aa * qq * zz + bb * rr * xx + cc * pp * yy - cc * qq * xx - bb * pp * zz - aa * rr * yy;
  xx = RHM_G(g, 2, 1); yy = RHM_G(g, 2, 2); zz = RHM_G(g, 2, 3);
  const double t4 = aa * qq * zz + bb * rr * xx + cc * pp * yy - cc * qq * xx - bb * pp * zz - aa * rr * yy;
  return sqrt(-1.0 * (RHM_G(g, 0, 0) * t1 - RHM_G(g, 1, 0) * t2 + RHM_G(g, 2, 0) * t3 - RHM_G(g, 3, 0) * t4));
}

// The trigonometric inputs of one theta for one Christoffel evaluation.
struct ThetaTrig {
  double s, c;     // at theta
  double sa, ca;   // at theta + delta
  double sb, cb;   // at theta - delta
};

// calchrist (coord.f90:1089-1321) at one (r, theta): metric derivatives by
// central differences of width delta = 1e-5 (+1e-12 in the denominators),
// Gamma^l_{mn} = g^{lk} (d_n g_km + d_m g_kn - d_k g_mn)/2, written as
// ch[l + 4m + 16n].  The metric has no phi dependence, so the phi difference
// is zero; ph is the phi position it is taken at.
RH_HD void christoffel(int metric, double akm, double r, double th, double ph, const ThetaTrig& t, double* ch) {
  const double delta = 1.0e-5, small = 1.0e-12;
  const double ria = r + delta, rib = r - delta, pja = ph + delta, pjb = ph - delta, tka = th + delta,
               tkb = th - delta;
  double gia[16], gib[16], gja[16], gjb[16], gka[16], gkb[16], gi[16];
  cov(metric, akm, gia, ria, t.s, t.c);
  cov(metric, akm, gib, rib, t.s, t.c);
  cov(metric, akm, gja, r, t.s, t.c);
  cov(metric, akm, gjb, r, t.s, t.c);
  cov(metric, akm, gka, r, t.sa, t.ca);
  cov(metric, akm, gkb, r, t.sb, t.cb);
  con(metric, akm, gi, r, t.s, t.c);
  double dg[4][4][4];  // dg[l][m][n] = d g_lm / dx^n
  for (int m = 0; m < 4; m++)
    for (int l = 0; l < 4; l++) {
      dg[l][m][0] = 0.0;
      dg[l][m][1] = (RHM_G(gia, l, m) - RHM_G(gib, l, m)) / (ria - rib + small);
      dg[l][m][2] = (RHM_G(gja, l, m) - RHM_G(gjb, l, m)) / (pja - pjb + small);
      dg[l][m][3] = (RHM_G(gka, l, m) - RHM_G(gkb, l, m)) / (tka - tkb + small);
    }
  double t2[4][4][4];
  for (int n = 0; n < 4; n++)
    for (int m = 0; m < 4; m++)
      for (int l = 0; l < 4; l++) t2[l][m][n] = 0.5 * (dg[m][l][n] + dg[n][l][m] - dg[n][m][l]);
  for (int n = 0; n < 4; n++)
    for (int m = 0; m < 4; m++)
      for (int l = 0; l < 4; l++)
        ch[l + 4 * m + 16 * n] = RHM_G(gi, l, 0) * t2[0][m][n] + RHM_G(gi, l, 1) * t2[1][m][n] +
                                 RHM_G(gi, l, 2) * t2[2][m][n] + RHM_G(gi, l, 3) * t2[3][m][n];
}

#undef RHM_G

}  // namespace rhmet

#endif
