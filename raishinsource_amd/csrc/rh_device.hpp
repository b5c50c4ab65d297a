// Device-side pointwise physics of the RAISHIN GRMHD step (gfx950).
//
// Every function reproduces the reference arithmetic operation for operation
// (same association, same single-precision literals) so that, compiled with
// -ffp-contract=off, results are bit-identical to the Fortran reference.
// Fortran intrinsic semantics: MAX(a,b) = a>b?a:b, MIN(a,b) = a<b?a:b,
// SIGN(1d0,x) = copysign(1,x).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define RH_DEV __device__ __forceinline__

namespace rh {

template <class R> RH_DEV R fmaxf_(R a, R b) { return a > b ? a : b; }
template <class R> RH_DEV R fminf_(R a, R b) { return a < b ? a : b; }
// libm by arithmetic type (the float forms serve the mixed-precision kernels)
RH_DEV double absr(double x) { return fabs(x); }
RH_DEV float absr(float x) { return fabsf(x); }
RH_DEV double sqrtr(double x) { return sqrt(x); }
RH_DEV float sqrtr(float x) { return sqrtf(x); }
RH_DEV double powr(double x, double y) { return pow(x, y); }
// float (the mixed form's face states only): 2^(y log2 x) on the hardware's
// base-2 logarithm and exponential, about 1e-6 relative for the densities
// here; NaN for x < 0 and 0 for x = 0 (y > 0), as powf
RH_DEV float powr(float x, float y) { return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x)); }
RH_DEV double fsign1(double x) { return copysign(1.0, x); }
RH_DEV float fsign1(float x) { return copysignf(1.0f, x); }
// x*y + z: in double a product and a sum, each rounded, which is the
// reference's arithmetic with contraction off; in float (the mixed form,
// which has no reference bits to keep) one fused multiply-add.  Written so
// that the double form keeps the operand order of the expression it
// replaces (a sum's two operands commute exactly).
RH_DEV double mad_(double x, double y, double z) { return x * y + z; }
RH_DEV float mad_(float x, float y, float z) { return __builtin_fmaf(x, y, z); }

// Geometry of one model on the device.
struct Geo {
  int ni, nj, nk;        // imax, jmax, kmax (ghosts included)
  int nm0, nm1;          // ghost widths (main.f90:179-211)
  long long N;           // cells stored per plane: ni*njs*nk
  long long P2;          // ni*nk: plane size of the axisymmetric (i,k) tables
  // Stored phi rows.  njs = nj, jfix = 0: every row.  njs = 1, jfix > 0:
  // the model is axisymmetric with one interior phi cell and periodic phi,
  // so every phi row holds the same values at all times; row jfix alone is
  // stored and computed, and every j maps to it.
  int njs, jfix;
  int icha;              // wave-speed formula (calcha.f90): 0 Gammie, 1 Leismann, 2 Del Zanna
  int iwvec;             // primitive recovery (recov.f90:2-47): 1 recov1a, 2 recov2
  // workgroup size per tuned launch slot (TuneSlot in rh_kernels.hpp);
  // 0 means 256.  Chosen by raishin_model_autotune or set by the caller.
  unsigned short bs[8];
  // PPM sweeps (irec 5): interfaces per lane along j and k of the fused
  // sweep (0: the default), and ppm_split = 1 for the split form (face
  // values through HBM in their own kernel).  Chosen by the autotuner or
  // set by the caller (raishin_model_set_sweep_tiles); results never depend
  // on them.
  unsigned short seg[2];
  unsigned short ppm_split;
  // 1: the face flux of the normal field component is set to zero (not the
  // reference's arithmetic: its HLLE keeps a dissipative normal-B flux that
  // flux-CT leaves alone); flux-CT then keeps div B at round-off
  unsigned short divb0;
};

// One metric location (centre or a face) from the (i,k) tables:
// planes 0..15 gcov(m,n) at m+4n, 16..31 gcon, 32 sqrt(-g).
struct Met {
  const double* __restrict__ b;
  long long P2, q;
  RH_DEV double g(int m, int n) const { return b[(long long)(m + 4 * n) * P2 + q]; }
  RH_DEV double gi(int m, int n) const { return b[(long long)(16 + m + 4 * n) * P2 + q]; }
  RH_DEV double detg() const { return b[32LL * P2 + q]; }
};

// physfunc.f90 helpers -------------------------------------------------
// Templated on the arithmetic type R: double reproduces the reference; float
// is the mixed-precision face arithmetic (metric values rounded to float).

// calgfl: physfunc.f90:2-41
template <class R>
RH_DEV R calgfl(const R u1, const R u2, const R u3, const Met& M) {
  const R ut[4] = {R(0), u1, u2, u3};
  R utsq = R(0);
#pragma unroll
  for (int m = 1; m <= 3; m++)
#pragma unroll
    for (int n = 1; n <= 3; n++) utsq = mad_(R(M.g(m, n)) * ut[m], ut[n], utsq);
  if (utsq < R(0)) {
    if (absr(utsq) <= R(1.0e-10)) utsq = absr(utsq);
    else utsq = R(1.0e-10);
  }
  if (absr(utsq) > R(5.0e2)) utsq = R(5.0e2);
  return sqrtr(R(1) + utsq);
}

// calucon: physfunc.f90:44-74
template <class R>
RH_DEV void calucon(const R ut[4], R ucon[4], R gfl, const Met& M) {
  const R alpha1 = R(1) / sqrtr(-R(M.gi(0, 0)));
  const R b1 = alpha1 * alpha1 * R(M.gi(0, 1));
  const R b2 = alpha1 * alpha1 * R(M.gi(0, 2));
  const R b3 = alpha1 * alpha1 * R(M.gi(0, 3));
  ucon[0] = gfl * (R(1) / alpha1);
  ucon[1] = mad_(-(gfl * b1), R(1) / alpha1, ut[1]);
  ucon[2] = mad_(-(gfl * b2), R(1) / alpha1, ut[2]);
  ucon[3] = mad_(-(gfl * b3), R(1) / alpha1, ut[3]);
}

// lower: physfunc.f90:158-192
template <class R>
RH_DEV void lower(const R v[4], R w[4], const Met& M) {
#pragma unroll
  for (int m = 0; m < 4; m++)
    w[m] = mad_(R(M.g(m, 3)), v[3], mad_(R(M.g(m, 2)), v[2], mad_(R(M.g(m, 1)), v[1], R(M.g(m, 0)) * v[0])));
}
// upper: physfunc.f90:195-230
template <class R>
RH_DEV void upper(const R v[4], R w[4], const Met& M) {
#pragma unroll
  for (int m = 0; m < 4; m++)
    w[m] = mad_(R(M.gi(m, 3)), v[3], mad_(R(M.gi(m, 2)), v[2], mad_(R(M.gi(m, 1)), v[1], R(M.gi(m, 0)) * v[0])));
}

// calbbcon: physfunc.f90:77-95
template <class R>
RH_DEV void calbbcon(const R bcon[4], R bb[4], const R ucov[4], const R ucon[4]) {
  bb[0] = mad_(bcon[3], ucov[3], mad_(bcon[2], ucov[2], bcon[1] * ucov[1]));
  bb[1] = mad_(bb[0], ucon[1], bcon[1]) * (R(1) / ucon[0]);
  bb[2] = mad_(bb[0], ucon[2], bcon[2]) * (R(1) / ucon[0]);
  bb[3] = mad_(bb[0], ucon[3], bcon[3]) * (R(1) / ucon[0]);
}

// cal3met: coord.f90:970-1013; g3/gi3 hold gcov3a/gcon3a(m,n) at (m-1)+3(n-1)
template <class R>
RH_DEV void cal3met(const Met& M, R g3[9], R gi3[9]) {
  const R alpha1 = R(1) / sqrtr(-R(M.gi(0, 0)));
  R beta1[4];
  beta1[1] = alpha1 * alpha1 * R(M.gi(0, 1));
  beta1[2] = alpha1 * alpha1 * R(M.gi(0, 2));
  beta1[3] = alpha1 * alpha1 * R(M.gi(0, 3));
  R oncon[4], oncov[4];
  oncon[0] = R(1) / alpha1;
  oncon[1] = beta1[1] / alpha1;
  oncon[2] = beta1[2] / alpha1;
  oncon[3] = beta1[3] / alpha1;
  oncov[0] = -alpha1; oncov[1] = R(0); oncov[2] = R(0); oncov[3] = R(0);
#pragma unroll
  for (int m = 1; m <= 3; m++)
#pragma unroll
    for (int n = 1; n <= 3; n++) {
      g3[(m - 1) + 3 * (n - 1)] = mad_(oncov[m], oncov[n], R(M.g(m, n)));
      gi3[(m - 1) + 3 * (n - 1)] = mad_(oncon[m], oncon[n], R(M.gi(m, n)));
    }
}

// Shared prelude of caluu1/caluu3/calflx3*/caltenr/calcha4*:
// primitives -> enthalpy, Lorentz factor, u^mu, u_mu, b^mu, b_mu, b^2.
template <class R>
struct PrelT {
  R de, pr, roh, gfl, bbsq;
  R ut[4], bcon[4], ucon[4], ucov[4], bb[4], bbcov[4];
};
using Prel = PrelT<double>;

template <class R>
RH_DEV void prelude(const R w[9], R gam, const Met& M, PrelT<R>& s) {
  s.de = w[0];
  s.ut[0] = R(0); s.ut[1] = w[1]; s.ut[2] = w[2]; s.ut[3] = w[3];
  s.pr = w[4];
  s.bcon[0] = R(0); s.bcon[1] = w[6]; s.bcon[2] = w[7]; s.bcon[3] = w[8];
  s.roh = mad_(gam / (gam - R(1)), s.pr, s.de);   // caluu.f90:52 (ieos=0)
  s.gfl = calgfl<R>(w[1], w[2], w[3], M);
  calucon<R>(s.ut, s.ucon, s.gfl, M);
  lower<R>(s.ucon, s.ucov, M);
  calbbcon<R>(s.bcon, s.bb, s.ucov, s.ucon);
  lower<R>(s.bb, s.bbcov, M);
  s.bbsq = mad_(s.bb[3], s.bbcov[3], mad_(s.bb[2], s.bbcov[2], mad_(s.bb[1], s.bbcov[1], s.bb[0] * s.bbcov[0])));
}

// Row d of sqrtr(-g) * (rho u^d, T^d_i, T^d_t + rho u^d, rho s u^d, induction):
// d=0 is caluu1/caluu3 (caluu.f90:85-103), d=1..3 calflx3i/j/k (calflx.f90:468-507,
// 608-646, 744-782). powdg = de**gam computed once by the caller.
template <class R>
RH_DEV void tmunu_row(const PrelT<R>& s, int d, R detg, R powdg, R out[9]) {
  R delta[4] = {R(0), R(0), R(0), R(0)};
  delta[d] = R(1);
  const R rb = s.roh + s.bbsq, pb = mad_(R(0.5), s.bbsq, s.pr);
  const R ud = s.ucon[d];
  out[0] = s.de * ud;
#pragma unroll
  for (int q = 1; q <= 3; q++)
    out[q] = mad_(-s.bb[d], s.bbcov[q], mad_(rb * ud, s.ucov[q], pb * delta[q]));
  out[4] = mad_(s.de, ud, mad_(-s.bb[d], s.bbcov[0], mad_(rb * ud, s.ucov[0], pb * delta[0])));
  out[5] = s.de * (s.pr / powdg) * ud;
  if (d == 0) {
    out[6] = s.bcon[1]; out[7] = s.bcon[2]; out[8] = s.bcon[3];
  } else {
    const R r0 = R(1) / s.ucon[0];
#pragma unroll
    for (int c = 1; c <= 3; c++)
      out[5 + c] = (c == d) ? R(0) : mad_(-(s.ucon[c] * r0), s.bcon[d], (ud * r0) * s.bcon[c]);
  }
#pragma unroll
  for (int n = 0; n < 9; n++) out[n] = detg * out[n];
}

// The (f1 +- sqrt f3) / f2 speed pair of one direction d for the three
// formulas of calcha.f90 (icha 0: Gammie et al. 2003, vcon +- sqrt(omsq);
// icha 1: Leismann et al. 2005 with r1 = cs^2 (v.B)^2 / ((rho h + b^2) W^2);
// icha 2: Del Zanna et al. 2007), before the lapse/shift correction.
// gii = gcon3a(d,d); vb = v_i B^i; the evaluation order is the reference's
// (calcha.f90:250-310, 489-527 and the j/k twins).
template <class R, int ICHA>
RH_DEV void speed_pair(R vd, R vsq, R omsq, R cssq, R vb, R rohb, R gfl, R gii, R& p, R& m) {
  if constexpr (ICHA == 0) {
    p = vd + sqrtr(omsq);
    m = vd - sqrtr(omsq);
    return;
  }
  const R f1 = vd * (R(1) - omsq);
  R f2, f3;
  if constexpr (ICHA == 1) {
    const R r1 = cssq * (vb * vb) * (R(1) / (rohb * (gfl * gfl)));
    f2 = mad_(-vsq, omsq, R(1)) - r1;
    f3 = mad_(vsq - R(1), omsq, r1) * (mad_(mad_(-vd, vd, gii * vsq), omsq, vd * vd) - gii + r1);
  } else {
    f2 = mad_(-vsq, omsq, R(1));
    f3 = omsq * (R(1) - vsq) * mad_(-((R(1) - omsq) * vd), vd, mad_(-vsq, omsq, R(1)) * gii);
  }
  p = (f1 + sqrtr(f3)) * (R(1) / f2);
  m = (f1 - sqrtr(f3)) * (R(1) / f2);
}

// v_i B^i with v_i = gcov3a v^j (lower1, physfunc.f90:233-258)
template <class R>
RH_DEV R v_dot_b(const R* g3, const R vcon[4], const R bcon[4]) {
  R vcov[4];
#pragma unroll
  for (int m = 1; m <= 3; m++)
    vcov[m] = mad_(g3[(m - 1) + 6], vcon[3], mad_(g3[(m - 1) + 3], vcon[2], g3[(m - 1) + 0] * vcon[1]));
  return mad_(vcov[3], bcon[3], mad_(vcov[2], bcon[2], vcov[1] * bcon[1]));
}

// Fast magnetosonic speeds along direction d for a face state:
// calcha4i/j/k (calcha.f90:370-906).  The three directions write cs^2 and
// vA^2 differently (calcha.f90:480,648,826 and 486,654,832).
template <class R, int ICHA>
RH_DEV void face_speed(const PrelT<R>& s, int d, R gam, const Met& M, R& vfp, R& vfm) {
  const R alpha1 = R(1) / sqrtr(-R(M.gi(0, 0)));
  const R beta_d = alpha1 * alpha1 * R(M.gi(0, d));
  R g3[9], gi3[9];
  cal3met<R>(M, g3, gi3);
  R vcon[4];
  vcon[1] = s.ut[1] * (R(1) / s.gfl);
  vcon[2] = s.ut[2] * (R(1) / s.gfl);
  vcon[3] = s.ut[3] * (R(1) / s.gfl);
  R vsq = R(0);
#pragma unroll
  for (int m = 1; m <= 3; m++)
#pragma unroll
    for (int n = 1; n <= 3; n++) vsq = mad_(g3[(m - 1) + 3 * (n - 1)] * vcon[m], vcon[n], vsq);
  R cssq, vasq;
  if (d == 3) cssq = gam * s.pr * (R(1) / s.roh);
  else cssq = gam * s.pr / s.roh;
  if (d == 2) vasq = s.bbsq / (s.roh + s.bbsq);
  else vasq = s.bbsq * (R(1) / (s.roh + s.bbsq));
  const R omsq = mad_(cssq, R(1) - vasq, vasq);
  const R vb = ICHA == 1 ? v_dot_b<R>(g3, vcon, s.bcon) : R(0);
  R p, m;
  speed_pair<R, ICHA>(vcon[d], vsq, omsq, cssq, vb, s.roh + s.bbsq, s.gfl, gi3[(d - 1) * 4], p, m);
  p = mad_(alpha1, p, -beta_d);
  m = mad_(alpha1, m, -beta_d);
  if (absr(p) > R(1)) p = R(1);
  if (absr(m) > R(1)) m = R(-1);
  vfp = p; vfm = m;
}

// Centre speeds for the CFL condition: calcha4 (calcha.f90:129-367).
template <int ICHA>
RH_DEV void centre_speeds(const Prel& s, double gam, const Met& M, double vp[4], double vm[4]) {
  const double alpha1 = 1.0 / sqrtr(-M.gi(0, 0));
  double beta1[4];
  beta1[1] = alpha1 * alpha1 * M.gi(0, 1);
  beta1[2] = alpha1 * alpha1 * M.gi(0, 2);
  beta1[3] = alpha1 * alpha1 * M.gi(0, 3);
  double g3[9], gi3[9];
  cal3met(M, g3, gi3);
  double vcon[4];
  vcon[1] = s.ut[1] * (1.0 / s.gfl);
  vcon[2] = s.ut[2] * (1.0 / s.gfl);
  vcon[3] = s.ut[3] * (1.0 / s.gfl);
  double vsq = 0.0;
#pragma unroll
  for (int m = 1; m <= 3; m++)
#pragma unroll
    for (int n = 1; n <= 3; n++) vsq = vsq + g3[(m - 1) + 3 * (n - 1)] * vcon[m] * vcon[n];
  const double cssq = gam * s.pr / s.roh;
  const double vasq = s.bbsq * (1.0 / (s.roh + s.bbsq));
  const double omsq = vasq + cssq * (1.0 - vasq);
  const double vb = ICHA == 1 ? v_dot_b<double>(g3, vcon, s.bcon) : 0.0;
#pragma unroll
  for (int d = 1; d <= 3; d++)
    speed_pair<double, ICHA>(vcon[d], vsq, omsq, cssq, vb, s.roh + s.bbsq, s.gfl, gi3[(d - 1) * 4], vp[d], vm[d]);
#pragma unroll
  for (int d = 1; d <= 3; d++) {
    vp[d] = alpha1 * vp[d] - beta1[d];
    vm[d] = alpha1 * vm[d] - beta1[d];
    if (absr(vp[d]) > 1.0) vp[d] = 1.0;
    if (absr(vm[d]) > 1.0) vm[d] = -1.0;
  }
}

// Reconstruction ---------------------------------------------------------

template <class R>
RH_DEV R dmm(R x, R y) {
  return R(0.5) * (fsign1(x) + fsign1(y)) * fminf_(absr(x), absr(y));
}
template <class R>
RH_DEV R dm4(R w, R x, R y, R z) {
  return R(0.125) * (fsign1(w) + fsign1(x)) * absr(fsign1(w) + fsign1(y)) * (fsign1(w) + fsign1(z)) *
         fminf_(fminf_(fminf_(absr(w), absr(x)), absr(y)), absr(z));
}

// MP5 (Suresh & Huynh), rec.f90:2059-2343.  s[0..4] = u(c-2..c+2).
// Right edge of cell c (uri*l(c)): rec.f90:2113-2145.
template <class R>
RH_DEV R mp5_right(const R s[5]) {
  const R tmp1 = R(1.0) / R(60.0), tmp2 = R(4.0) / R(3.0), alpha = R(4.0), eps1 = R(1.0e-10);
  const R u1 = tmp1 * (R(2.0) * s[0] - R(13.0) * s[1] + R(47.0) * s[2] + R(27.0) * s[3] - R(3.0) * s[4]);
  const R ump = s[2] + dmm(s[3] - s[2], alpha * (s[2] - s[1]));
  if ((u1 - s[2]) * (u1 - ump) <= eps1) return u1;
  const R dim1 = s[0] - R(2.0) * s[1] + s[2];
  const R di1 = s[1] - R(2.0) * s[2] + s[3];
  const R dip1 = s[2] - R(2.0) * s[3] + s[4];
  const R dph = dm4(R(4.0) * di1 - dip1, R(4.0) * dip1 - di1, di1, dip1);
  const R dmh = dm4(R(4.0) * di1 - dim1, R(4.0) * dim1 - di1, di1, dim1);
  const R ul = s[2] + alpha * (s[2] - s[1]);
  const R av = R(0.5) * (s[2] + s[3]);
  const R md = av - R(0.5) * dph;
  const R lc = s[2] + R(0.5) * (s[2] - s[1]) + tmp2 * dmh;
  const R umax = fmaxf_(fminf_(fminf_(s[2], s[3]), md), fminf_(fminf_(s[2], ul), lc));
  const R umin = fminf_(fmaxf_(fmaxf_(s[2], s[3]), md), fmaxf_(fmaxf_(s[2], ul), lc));
  return u1 + dmm(umin - u1, umax - u1);
}
// Left edge of cell c (uri*r(c-1)): rec.f90:2149-2186.
template <class R>
RH_DEV R mp5_left(const R s[5]) {
  const R tmp1 = R(1.0) / R(60.0), tmp2 = R(4.0) / R(3.0), alpha = R(4.0), eps1 = R(1.0e-10);
  const R u1 = tmp1 * (R(2.0) * s[4] - R(13.0) * s[3] + R(47.0) * s[2] + R(27.0) * s[1] - R(3.0) * s[0]);
  const R ump = s[2] + dmm(s[1] - s[2], alpha * (s[2] - s[3]));
  if ((u1 - s[2]) * (u1 - ump) <= eps1) return u1;
  const R dim1 = s[0] - R(2.0) * s[1] + s[2];
  const R di1 = s[1] - R(2.0) * s[2] + s[3];
  const R dip1 = s[2] - R(2.0) * s[3] + s[4];
  const R dph = dm4(R(4.0) * di1 - dip1, R(4.0) * dip1 - di1, di1, dip1);
  const R dmh = dm4(R(4.0) * di1 - dim1, R(4.0) * dim1 - di1, di1, dim1);
  const R ul = s[2] + alpha * (s[2] - s[3]);
  const R av = R(0.5) * (s[2] + s[1]);
  const R md = av + R(0.5) * dmh;
  const R lc = s[2] + R(0.5) * (s[2] - s[3]) - tmp2 * dph;
  const R umax = fmaxf_(fminf_(fminf_(s[2], s[1]), md), fminf_(fminf_(s[2], ul), lc));
  const R umin = fminf_(fmaxf_(fmaxf_(s[2], s[1]), md), fmaxf_(fmaxf_(s[2], ul), lc));
  return u1 + dmm(umin - u1, umax - u1);
}

// WENO5 (Jiang & Shu), rec.f90:2373-2642; 13./12. is a single-precision literal.
template <class R>
RH_DEV R weno_edge(R a0, R a1, R a2, R a3, R a4) {
  // a0..a4 = stencil ordered from the far upwind side to the far downwind side
  const R c1312 = R(13.0f / 12.0f), c14 = R(0.25), epsi = R(1.0e-6);
  const R cd0 = R(1.0) / R(10.0), cd1 = R(6.0) / R(10.0), cd2 = R(3.0) / R(10.0);
  const R c00 = R(2.0) / R(6.0), c01 = -R(7.0) / R(6.0), c02 = R(11.0) / R(6.0), c10 = -R(1.0) / R(6.0), c11 = R(5.0) / R(6.0),
               c12 = R(2.0) / R(6.0), c20 = R(2.0) / R(6.0), c21 = R(5.0) / R(6.0), c22 = -R(1.0) / R(6.0);
  R a, b;
  a = a0 - R(2.0) * a1 + a2; b = a0 - R(4.0) * a1 + R(3.0) * a2;
  const R b0 = c1312 * (a * a) + c14 * (b * b) + epsi;
  a = a1 - R(2.0) * a2 + a3; b = a1 - a3;
  const R b1 = c1312 * (a * a) + c14 * (b * b) + epsi;
  a = a2 - R(2.0) * a3 + a4; b = R(3.0) * a2 - R(4.0) * a3 + a4;
  const R b2 = c1312 * (a * a) + c14 * (b * b) + epsi;
  const R w0 = cd0 / (b0 * b0), w1 = cd1 / (b1 * b1), w2 = cd2 / (b2 * b2);
  const R wt = w0 + w1 + w2;
  const R o0 = w0 / wt, o1 = w1 / wt, o2 = w2 / wt;
  const R r0 = c00 * a0 + c01 * a1 + c02 * a2;
  const R r1 = c10 * a1 + c11 * a2 + c12 * a3;
  const R r2 = c20 * a2 + c21 * a3 + c22 * a4;
  return o0 * r0 + o1 * r1 + o2 * r2;
}

}  // namespace rh
