// HIP kernels of the RAISHIN GRMHD time step for gfx950 (FP64 path).
//
// Device layout (structure of arrays, i fastest):
//   state planes   A[n*N + c],  c = (i-1) + ni*((j-1) + nj*(k-1)),  n = 0..8
//   fluxes         F[d][n*N + c] holds ww(d,n,i,j,k) = flux at i+1/2 (j+1/2, k+1/2)
//   metric tables  (i,k) planes: the Kerr-Schild metric has no phi dependence,
//                  so each location set is 33 planes of ni*nk values and the
//                  Christoffel symbols 64 planes.
// Each kernel cites the reference routines it restates.
#include <type_traits>
#include "rh_device.hpp"
#include "rh_metric.hpp"
#include "rh_kernels.hpp"

namespace rh {

RH_DEV long long cidx(int i, int j, int k, const Geo& G) {
  const long long jr = G.jfix ? 0 : (long long)(j - 1);
  return (long long)(i - 1) + (long long)G.ni * (jr + (long long)G.njs * (long long)(k - 1));
}
// j range of a box: the stored row of a collapsed model, else [lo, hi]
RH_DEV int jlo(const Geo& G, int lo) { return G.jfix ? G.jfix : lo; }
RH_DEV int jhi(const Geo& G, int hi) { return G.jfix ? G.jfix : hi; }
// strides of a step in j and in k in the stored planes
RH_DEV long long jstride(const Geo& G) { return G.jfix ? 0LL : (long long)G.ni; }
RH_DEV long long kstride(const Geo& G) { return (long long)G.ni * G.njs; }
RH_DEV long long q2(int i, int k, const Geo& G) { return (long long)(i - 1) + (long long)G.ni * (long long)(k - 1); }

// 1-based (i,j,k) of linear thread t over a box [lo_a, hi_a] per axis; false if outside.
RH_DEV bool box_index(long long t, int ilo, int ihi, int jlo, int jhi, int klo, int khi, int& i, int& j, int& k) {
  const long long ci = ihi - ilo + 1, cj = jhi - jlo + 1, ck = khi - klo + 1;
  if (ci <= 0 || cj <= 0 || ck <= 0 || t >= ci * cj * ck) return false;
  i = ilo + (int)(t % ci);
  const long long r = t / ci;
  j = jlo + (int)(r % cj);
  k = klo + (int)(r / cj);
  return true;
}

// Workgroup b is dispatched to XCD b % 8.  Renumbering the workgroups so that
// each XCD takes one contiguous eighth of the launch keeps neighbouring rows
// of a stencil in the same L2.  A bijection on [0, gridDim.x).
RH_DEV long long xcd_block() {
  const unsigned nb = gridDim.x, b = blockIdx.x, x = b % 8u, idx = b / 8u, q = nb / 8u, r = nb % 8u;
  return x < r ? (long long)x * (q + 1) + idx : (long long)r * (q + 1) + (long long)(x - r) * q + idx;
}
RH_DEV long long xcd_thread() { return xcd_block() * blockDim.x + threadIdx.x; }

// As box_index with the j and k axes swapped in the ordering (i fastest, then
// k, then j): consecutive workgroups step along k, the k sweep's stencil axis.
RH_DEV bool box_index_ikj(long long t, int ilo, int ihi, int jlo, int jhi, int klo, int khi, int& i, int& j, int& k) {
  const long long ci = ihi - ilo + 1, cj = jhi - jlo + 1, ck = khi - klo + 1;
  if (ci <= 0 || cj <= 0 || ck <= 0 || t >= ci * cj * ck) return false;
  i = ilo + (int)(t % ci);
  const long long r = t / ci;
  k = klo + (int)(r % ck);
  j = jlo + (int)(r / ck);
  return true;
}

// Waves along the D axis (D = 2 or 3): a lane's i within a 64-wide chunk
// of the row, consecutive waves on consecutive j (D=2) or k (D=3) rows, so
// the cell's index along D is the same in every lane of a wave.  Covers
// 64*ceil(ci/64)*cj*ck indices; lanes past the row's end return false.
template <int D>
RH_DEV bool box_index_wave(long long t, int ilo, int ihi, int jlo, int jhi, int klo, int khi, int& i, int& j,
                           int& k) {
  const long long ci = ihi - ilo + 1, cj = jhi - jlo + 1, ck = khi - klo + 1;
  if (ci <= 0 || cj <= 0 || ck <= 0) return false;
  const long long nic = (ci + 63) / 64;
  const long long w = t >> 6;
  long long ic;
  if (D == 2) {
    j = jlo + (int)(w % cj);
    const long long r = w / cj;
    ic = r % nic;
    const long long kk = r / nic;
    if (kk >= ck) return false;
    k = klo + (int)kk;
  } else {
    k = klo + (int)(w % ck);
    const long long r = w / ck;
    ic = r % nic;
    const long long jj = r / nic;
    if (jj >= cj) return false;
    j = jlo + (int)jj;
  }
  i = ilo + (int)(ic * 64 + (t & 63));
  return i <= ihi;
}

// Counter slots (raishin_counters field order).
enum { C_FALLBACK = 0, C_FAIL, C_NEGRHO, C_NEGP, C_VSQGT1, C_DET0, C_WGCARRY, C_ATM, C_PMMSCALE, C_PMMFLOOR,
       C_GUARD, C_NCOUNT };

// Event counts without atomics: each wavefront of a launch owns the row
// cpart[w*C_NCOUNT ..] (w = its index in the model's part of the launch, from
// the workgroup's index blk there); the lowest flagged lane
// adds the wave's popcount.  Launches are stream-ordered, so a row has one
// writer at a time; k_sum_counters folds the rows when counters are read.
// (Device-scope atomics on a few addresses from every wave serialise across
// the XCDs: they cost more than the kernels that raise the events.)
RH_DEV void count_wave(unsigned long long* cpart, long long blk, int slot, bool flag,
                       unsigned long long weight = 1) {
  const unsigned long long b = __ballot(flag);
  if (b && (threadIdx.x % warpSize) == (unsigned)(__ffsll((long long)b) - 1)) {
    const long long w = (blk * blockDim.x + threadIdx.x) / warpSize;
    cpart[w * C_NCOUNT + slot] += (unsigned long long)__popcll(b) * weight;
  }
}

// cnt[s] += sum over rows of cpart[.][s], then the rows are cleared.  One
// workgroup of 1024.
__global__ void __launch_bounds__(1024) k_sum_counters(unsigned long long* __restrict__ cpart, long long rows,
                                                       unsigned long long* __restrict__ cnt) {
  __shared__ unsigned long long part[C_NCOUNT][1024 / 64];
  unsigned long long acc[C_NCOUNT];
#pragma unroll
  for (int s = 0; s < C_NCOUNT; s++) acc[s] = 0;
  for (long long r = threadIdx.x; r < rows; r += blockDim.x)
#pragma unroll
    for (int s = 0; s < C_NCOUNT; s++) {
      acc[s] += cpart[r * C_NCOUNT + s];
      cpart[r * C_NCOUNT + s] = 0;
    }
  const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
#pragma unroll
  for (int s = 0; s < C_NCOUNT; s++) {
    unsigned long long v = acc[s];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    if (lane == 0) part[s][wv] = v;
  }
  __syncthreads();
  if (threadIdx.x < C_NCOUNT) {
    unsigned long long t = 0;
    for (int q = 0; q < (int)(blockDim.x / 64); q++) t += part[threadIdx.x][q];
    cnt[threadIdx.x] += t;
  }
}

// ------------------------------------------------------------------------
// PPM reconstruction of one cell (rec.f90:1261-2056), direction D.
// Hands each variable's left edge l (uri*r(c-1)) and right edge r (uri*l(c))
// to store(n, l, r), one variable at a time.
// p is the 1-based index of the cell along D; the loads are along stride s.
// dxs[d] holds the dummy dx*b(d) of rec2 (= actual dx*b(d-1), rec.f90:58).
// ------------------------------------------------------------------------
// Shape coefficients of the cell: flattening f2 (rec.f90:1736-1858) and the
// contact-steepening eta of the density (rec.f90:1537-1629).  R is the
// arithmetic type (double: the reference's; float: the mixed-precision form).
template <int D, class R>
RH_DEV void ppm_shape(const double* __restrict__ W, long long N, long long c, long long s, int p,
                      const double* __restrict__ dxs, const double* __restrict__ xs, R gam, R& f2_out, R& eta_out) {
  const R dm2 = R(dxs[p - 2]), dm1 = R(dxs[p - 1]), d0 = R(dxs[p]), dp1 = R(dxs[p + 1]), dp2 = R(dxs[p + 2]);
  // --- flattening coefficient f2 of this cell ---
  const int VEL = D;  // velocity component along D is variable D+1 (0-based D)
  const double* P5 = W + 4 * N + c;
  const double* PV = W + (long long)VEL * N + c;
  R f1[3];
#pragma unroll
  for (int q = -1; q <= 1; q++) {
    const R pp1 = R(P5[(q + 1) * s]), pm1 = R(P5[(q - 1) * s]), pp2 = R(P5[(q + 2) * s]), pm2 = R(P5[(q - 2) * s]);
    const R t3a = absr(pp1 - pm1);
    const R t1 = R(1) * fminf_(pp1, pm1) - t3a;
    const R t2 = R(PV[(q + 1) * s]) - R(PV[(q - 1) * s]);
    const R omg1 = (t1 < R(0) && t2 < R(0)) ? R(1) : R(0);
    const R t3b = (pp2 - pm2);
    R t4;
    if (t3b == R(0)) t4 = (t3a == R(0)) ? R(-0.52) : R(1.0) - R(0.52);
    else t4 = (t3a / t3b) - R(0.52);
    f1[q + 1] = fminf_(R(1), omg1 * fmaxf_(R(0), t4 * R(10.0)));
  }
  const R p5p = R(P5[s]), p5m = R(P5[-s]);
  const R f1b = (p5p - p5m < R(0)) ? f1[2] : f1[0];
  f2_out = fmaxf_(f1[1], f1b);

  // --- contact steepening parameter eta (density only) ---
  R eta;
  {
    const double* P1 = W + c;
    const R um2 = R(P1[-2 * s]), um1 = R(P1[-s]), u0 = R(P1[0]), up1 = R(P1[s]), up2 = R(P1[2 * s]);
    const R ddm = (R(1) / (dm2 + dm1 + d0)) * ((u0 - um1) / (d0 + dm1) - (um1 - um2) / (dm1 + dm2));
    const R ddp = (R(1) / (d0 + dp1 + dp2)) * ((up2 - up1) / (dp2 + dp1) - (up1 - u0) / (dp1 + d0));
    const R t1 = ddp * ddm;
    const R t2 = mad_(-R(0.1), fminf_(absr(up1), absr(um1)), absr(up1 - um1));
    R etatil;
    if (t1 >= R(0) && t2 <= R(0)) etatil = R(0);
    else {
      const R xm = R(xs[p - 2]), x0 = R(xs[p - 1]), xp = R(xs[p]);
      const R t2a = -(ddp - ddm) / (xp - xm);
      const R dl = x0 - xm, dr = xp - x0;
      const R t2b = mad_(dr * dr, dr, dl * dl * dl) / (up1 - um1);
      etatil = t2a * t2b;
    }
    eta = fmaxf_(R(0), fminf_(R(5.0) * (etatil - R(0.05)), R(1)));
    const R t3 = absr(up1 - um1) / fminf_(up1, um1);
    const R t4 = absr(p5p - p5m) / fminf_(p5p, p5m);
    if (gam * R(1) * t3 < t4) eta = R(0);
  }
  eta_out = eta;
}

// interp's width-only coefficients of a cell at index p (ppm_coefficients on
// the host, same expressions and order): t1a,t1b,t1c of dma at q = p-1, p,
// p+1 (rec.f90:1353-1355) and the four of equation (51) at q+1/2 for q = p
// and p-1 (rec.f90:1427-1436).  The same for every variable: loaded once.
template <class R>
struct PpmCoef {
  R ta[3], tb[3], tc[3];
  R e2[2], e3[2], e4[2], e5[2];
};

template <class R>
RH_DEV void ppm_coef(const double* __restrict__ pc, long long npc, int p, PpmCoef<R>& k) {
#pragma unroll
  for (int o = 0; o < 3; o++) {
    const int q = p - 1 + o;
    k.ta[o] = R(pc[q]); k.tb[o] = R(pc[npc + q]); k.tc[o] = R(pc[2 * npc + q]);
  }
#pragma unroll
  for (int o = 0; o < 2; o++) {
    const int q = p - o;
    k.e2[o] = R(pc[3 * npc + q]); k.e3[o] = R(pc[4 * npc + q]);
    k.e4[o] = R(pc[5 * npc + q]); k.e5[o] = R(pc[6 * npc + q]);
  }
}

// Interface values of variable n of the cell from its 5-point stencil
// u = (u(p-2) .. u(p+2)): interp (rec.f90:1309-1476), steepening of the
// density (rec.f90:1661-1677), flattening and monoto (rec.f90:1887-2056).
// l = uri*r(c-1), r = uri*l(c).
template <int D, class R>
RH_DEV void ppm_var(const R u[5], const PpmCoef<R>& k, int n, R f2, R eta, R& l_out, R& r_out) {
  auto dma_of = [&](R um, R u0, R up, int o) -> R {
    const R t1a = k.ta[o], t1b = k.tb[o], t1c = k.tc[o];
    const R sp1 = (up - u0);
    const R sm1 = (u0 - um);
    const R spm1 = sp1 * sm1;
    const R da = t1a * mad_(t1b, sp1, t1c * sm1);
    if (spm1 <= R(0)) return R(0);
    const R t1 = (da < R(0)) ? R(-1) : R(1);
    const R t1d = fminf_(absr(sm1), absr(sp1));
    if (D == 3) return t1 * fminf_(absr(da), t1d);   // rec.f90:1414
    return t1 * fminf_(absr(da), R(2.0) * t1d);
  };
  auto eq51 = [&](R uq, R uq1, R dmq, R dmq1, int o) -> R {
    // uq + t2 + t3 - e4 dmq1 + e5 dmq with t2 = e2 (uq1 - uq), t3 = e3 (uq1 - uq)
    const R du = uq1 - uq;
    return mad_(k.e5[o], dmq, mad_(-k.e4[o], dmq1, mad_(k.e3[o], du, mad_(k.e2[o], du, uq))));
  };
  const R um2 = u[0], um1 = u[1], u0 = u[2], up1 = u[3], up2 = u[4];
  const R dmm1 = dma_of(um2, um1, u0, 0);
  const R dm0 = dma_of(um1, u0, up1, 1);
  const R dmp1 = dma_of(u0, up1, up2, 2);
  R r = eq51(u0, up1, dm0, dmp1, 0);          // value at p+1/2
  R l = eq51(um1, u0, dmm1, dm0, 1);          // value at p-1/2
  if (n == 0) {
    r = mad_(r, R(1) - eta, mad_(-R(0.5), dmp1, up1) * eta);
    l = mad_(l, R(1) - eta, mad_(R(0.5), dmm1, um1) * eta);
  }
  r = mad_(u0, f2, r * (R(1) - f2));
  l = mad_(u0, f2, l * (R(1) - f2));
  if ((r - u0) * (l - u0) >= R(0)) { r = u0; l = u0; }            // equation (73)
  const R t3 = R(6.0) * (r - l) * mad_(-R(0.5), r + l, u0);        // equations (74), (75)
  const R t4 = (r - l) * (r - l);
  if (t3 > t4) l = mad_(R(3.0), u0, -(R(2.0) * r));
  if (-t3 > t4) r = mad_(R(3.0), u0, -(R(2.0) * l));
  l_out = l;
  r_out = r;
}

// ------------------------------------------------------------------------
// PPM face values of every cell once, along D: one thread per cell computes
// f2 and eta, then each variable's edges in turn, writing E indexed by
// interface: planes 0..8 at c-s the left edge of cell c (uri*r(c-1)), planes
// 9..17 at c its right edge (uri*l(c)).  Cells nm0..n-nm0+1 along D, the
// sweep's box across.  R: arithmetic and storage type of the edges.
// ------------------------------------------------------------------------
template <int D, class R>
RH_DEV void b_ppm_edges(const Geo& G, const EdgeArgs& A, long long blk) {
  const double* __restrict__ W = A.W;
  const double* __restrict__ dxs = A.dxs;
  const double* __restrict__ xs = A.xs;
  const double* __restrict__ pc = A.pc;
  const double gam = A.gam;
  R* __restrict__ E = (R*)A.E;
  int i, j, k;
  const long long t = blk * blockDim.x + threadIdx.x;
  const int hi = G.ni - G.nm0 + (D == 1), hj = jhi(G, G.nj - G.nm0 + (D == 2)), hk = G.nk - G.nm0 + (D == 3);
  const int lj = jlo(G, G.nm0);
  const bool in = (D == 1) ? box_index(t, G.nm0, hi, lj, hj, G.nm0, hk, i, j, k)
                           : box_index_wave<D>(t, G.nm0, hi, lj, hj, G.nm0, hk, i, j, k);
  if (!in) return;
  const long long N = G.N;
  const long long s = (D == 1) ? 1LL : (D == 2) ? jstride(G) : kstride(G);
  const long long c = cidx(i, j, k, G);
  // along j and k every lane of a wave has the same index along D: its
  // coefficients and widths are loaded once per wave, into scalar registers
  const int p = (D == 1) ? i : __builtin_amdgcn_readfirstlane((D == 2) ? j : k);
  const long long npc = ((D == 1) ? G.ni : (D == 2) ? G.nj : G.nk) + 2;
  PpmCoef<R> kc;
  ppm_coef<R>(pc, npc, p, kc);
  R f2, eta;
  ppm_shape<D, R>(W, N, c, s, p, dxs, xs, R(gam), f2, eta);
  // one variable at a time, the next variable's stencil loads in flight
  // while the current one is computed
  R cur[5], nxt[5];
#pragma unroll
  for (int o = 0; o < 5; o++) cur[o] = R(W[c + (o - 2) * s]);
#pragma unroll 1
  for (int n = 0; n < 9; n++) {
    if (n < 8) {
      const double* Pn = W + (long long)(n + 1) * N + c;
#pragma unroll
      for (int o = 0; o < 5; o++) nxt[o] = R(Pn[(o - 2) * s]);
    }
    R l, r;
    ppm_var<D, R>(cur, kc, n, f2, eta, l, r);
    E[(long long)n * N + c - s] = l;   // stored at the interface it belongs to (c-1/2)
    E[(long long)(9 + n) * N + c] = r;
#pragma unroll
    for (int o = 0; o < 5; o++) cur[o] = nxt[o];
  }
}

// MC-limited linear reconstruction of every cell along D (mclim,
// rec.f90:208-367, irec=1): slopes to both neighbours over the centre
// distances (+1e-12), the MC limiter, edges at the cell's faces; stored like
// the PPM edges.  A.pc holds the face positions x*a of the direction.
template <int D, class R>
RH_DEV void b_mc_edges(const Geo& G, const EdgeArgs& A, long long blk) {
  const double* __restrict__ W = A.W;
  const double* __restrict__ xs = A.xs;
  const double* __restrict__ xa = A.pc;
  R* __restrict__ E = (R*)A.E;
  int i, j, k;
  const long long t = blk * blockDim.x + threadIdx.x;
  const int hi = G.ni - G.nm0 + (D == 1), hj = jhi(G, G.nj - G.nm0 + (D == 2)), hk = G.nk - G.nm0 + (D == 3);
  const int lj = jlo(G, G.nm0);
  const bool in = (D == 3) ? box_index_ikj(t, G.nm0, hi, lj, hj, G.nm0, hk, i, j, k)
                           : box_index(t, G.nm0, hi, lj, hj, G.nm0, hk, i, j, k);
  if (!in) return;
  const long long N = G.N;
  const long long s = (D == 1) ? 1LL : (D == 2) ? jstride(G) : kstride(G);
  const long long c = cidx(i, j, k, G);
  const int p = (D == 1) ? i : (D == 2) ? j : k;
  const R small = R(1.0e-12);
  const R xp = R(xs[p]), xc = R(xs[p - 1]), xm = R(xs[p - 2]), xap = R(xa[p - 1]), xam = R(xa[p - 2]);
  const R ip = R(1) / (xp - xc + small), im = R(1) / (xc - xm + small);
#pragma unroll 3
  for (int n = 0; n < 9; n++) {
    const double* Pn = W + (long long)n * N + c;
    const R u = R(Pn[0]), up = R(Pn[s]), um = R(Pn[-s]);
    const R sp = (up - u) * ip;
    const R sm = (u - um) * im;
    const R tmp = (sp < R(0)) ? R(-1) : R(1);
    const R dk = tmp * fmaxf_(R(0), fminf_(fminf_(R(2) * absr(sp), tmp * R(2) * sm), tmp * R(0.5) * (sp + sm)));
    E[(long long)n * N + c - s] = u - dk * (xc - xam);   // left edge, at the interface c-1/2
    E[(long long)(9 + n) * N + c] = u + dk * (xap - xc); // right edge
  }
}

// Face physics and HLLE of interface c+1/2 along D from its two face states
// (wl: the left cell's right edge, wr: the right cell's left edge):
// caluu3 -> calflx3 -> calcha -> hlle, written to F (and the transverse B
// fluxes to their CTsrc planes).
template <int D, class R, int ICHA>
RH_DEV void sweep_face(const Geo& G, R* __restrict__ F, R* __restrict__ CTsrc, long long c, int i, int j, int k,
                       const Met& M, R gam, const R wl[9], const R wr[9]) {
  const long long N = G.N;
  const R detg = R(M.detg());
  R Ul[9], Fl[9], vpl, vml;
  {
    PrelT<R> sl;
    prelude<R>(wl, gam, M, sl);
    const R pw = powr(sl.de, gam);
    tmunu_row<R>(sl, 0, detg, pw, Ul);
    tmunu_row<R>(sl, D, detg, pw, Fl);
    face_speed<R, ICHA>(sl, D, gam, M, vpl, vml);
  }
  R Ur[9], Fr[9], vpr, vmr;
  {
    PrelT<R> sr;
    prelude<R>(wr, gam, M, sr);
    const R pw = powr(sr.de, gam);
    tmunu_row<R>(sr, 0, detg, pw, Ur);
    tmunu_row<R>(sr, D, detg, pw, Fr);
    face_speed<R, ICHA>(sr, D, gam, M, vpr, vmr);
  }
  const R cmax = fmaxf_(fmaxf_(R(0), vpr), vpl);
  const R cmin = (D == 3) ? fminf_(fminf_(R(0), vpr), vml)   // calcha.f90:107
                          : fminf_(fminf_(R(0), vmr), vml);
  R* Fo = F + c;
  // transverse B fluxes of this direction and their CTsrc planes
  // (0 ww(2,7) 1 ww(3,7) 2 ww(1,8) 3 ww(3,8) 4 ww(1,9) 5 ww(2,9))
  constexpr int nA = (D == 1) ? 7 : 6, nB = (D == 3) ? 7 : 8;
  constexpr int pA = (D == 1) ? 2 : (D == 2) ? 0 : 1, pB = (D == 1) ? 4 : (D == 2) ? 5 : 3;
  const bool in_ct = i >= G.nm1 && i <= G.ni - G.nm1 && j >= G.nm1 && j <= G.nj - G.nm1 && k >= G.nm1 &&
                     k <= G.nk - G.nm1;
#pragma unroll
  for (int n = 0; n < 9; n++) {
    double o;
    if (cmin > R(0)) o = double(Fl[n]);
    else if (cmin <= R(0) && cmax > R(0)) {
      const R c1 = fmaxf_(absr(cmax), absr(cmin));
      o = double(R(0.5) * mad_(-c1, Ur[n] - Ul[n], Fl[n] + Fr[n]));
    } else if (cmax <= R(0)) o = double(Fr[n]);
    else o = Fo[(long long)n * N];  // unordered (NaN) speeds: hlle keeps ww, then flushes it
    if (absr(o) < 1.0e-40) o = 0.0;
    // the labelled non-reference form: no flux of the normal field component
    // (the induction flux itself is zero; HLLE's dissipation term is what
    // the reference keeps, hll.f90:109-111), so flux-CT keeps div B
    if (n == 5 + D && G.divb0) o = 0.0;
    if (n == nA || n == nB) {
      CTsrc[(long long)(n == nA ? pA : pB) * N + c] = R(o);
      if (!in_ct) Fo[(long long)n * N] = R(o);  // inside the CT box fluxct writes this slot
    } else {
      Fo[(long long)n * N] = R(o);
    }
  }
}

// ------------------------------------------------------------------------
// Direction sweep: reconstruction + face fluxes + wave speeds + HLLE.
// rec2 (rec.f90:2-142) -> caluu3/calflx3* -> calcha (calcha.f90:2-126) -> hlle
// (hll.f90:47-192), fused per interface.  One thread per interface c+1/2.
// ------------------------------------------------------------------------
// F holds ww(D,:,:,:,:) exactly as the reference's array: the last value that
// hlle or ct wrote to each slot.  The two transverse B fluxes of direction D
// also go to their plane of CTsrc (the copies ct.f90:54-67 takes before
// averaging); inside the CT box fluxct then writes F itself.
//
// PPM's face values come from k_ppm_edges (E); MP5 and WENO5 reconstruct here.
template <int D, int IREC, class R, int ICHA>
RH_DEV void b_sweep(const Geo& G, const SweepArgs& A, long long blk) {
  const double* __restrict__ W = A.W;
  const double* __restrict__ metF = A.metF;
  const R* __restrict__ E = (const R*)A.E;
  R* __restrict__ F = (R*)A.F;
  R* __restrict__ CTsrc = (R*)A.CTsrc;
  const R gam = R(A.gam);
  int i, j, k;
  const long long t = blk * blockDim.x + threadIdx.x;
  const int lj = jlo(G, G.nm0), hj = jhi(G, G.nj - G.nm0);
  // k sweeps that read a stencil along k (MP5, WENO5) walk k before j so
  // neighbouring workgroups share the stencil planes in L2; with the face
  // values precomputed (PPM, MC) the i,j,k walk keeps each (i,k) metric
  // entry in L2 across the j rows instead.
  constexpr bool kfirst = D == 3 && (IREC == 6 || IREC == 7);
  const bool in = kfirst ? box_index_ikj(t, G.nm0, G.ni - G.nm0, lj, hj, G.nm0, G.nk - G.nm0, i, j, k)
                         : box_index(t, G.nm0, G.ni - G.nm0, lj, hj, G.nm0, G.nk - G.nm0, i, j, k);
  if (!in) return;
  const long long N = G.N;
  const long long s = (D == 1) ? 1LL : (D == 2) ? jstride(G) : kstride(G);
  const long long c = cidx(i, j, k, G);
  R wl[9], wr[9];
  if (IREC == 5 || IREC == 1) {
    // right edge of cell c (uri*l(c)) and left edge of cell c+1 (uri*r(c)),
    // both stored at interface c+1/2's index
#pragma unroll
    for (int n = 0; n < 9; n++) { wl[n] = E[(long long)(9 + n) * N + c]; wr[n] = E[(long long)n * N + c]; }
  } else {
#pragma unroll
    for (int n = 0; n < 9; n++) {
      const double* Pn = W + (long long)n * N + c;
      const R a = R(Pn[-2 * s]), b = R(Pn[-s]), cc = R(Pn[0]), d = R(Pn[s]), e = R(Pn[2 * s]), f = R(Pn[3 * s]);
      if (IREC == 6) {
        const R sl[5] = {a, b, cc, d, e};
        const R sr[5] = {b, cc, d, e, f};
        wl[n] = mp5_right<R>(sl);
        wr[n] = mp5_left<R>(sr);
      } else {
        wl[n] = weno_edge<R>(a, b, cc, d, e);
        wr[n] = weno_edge<R>(f, e, d, cc, b);
      }
    }
  }
  const Met M{metF, G.P2, q2(i, k, G)};
  sweep_face<D, R, ICHA>(G, F, CTsrc, c, i, j, k, M, gam, wl, wr);
}

// ------------------------------------------------------------------------
// PPM direction sweep with its face values computed in the same kernel:
// each cell's edges (interp, detect, flaten, monoto; rec.f90:1261-2056) are
// written to LDS and the interface physics (caluu3 -> calflx3 -> calcha ->
// hlle) reads them there, so the 18 edge planes never reach HBM.
//  D = 1: a wave takes 64 consecutive cells of an i row; lane l's interface
//         (the right face of its cell) takes the left edges of lane l+1's
//         cell from LDS.  Consecutive waves of a row overlap by one cell.
//  D = 2, 3: a wave holds 64 / A.seg consecutive i by A.seg consecutive
//         cells along D; lane l's interface takes the left edges of the
//         cell A.seg lanes further along D.  Waves overlap by one cell along D.
// The arithmetic is that of k_ppm_edges followed by k_sweep (the same
// functions), so the fluxes are the same bit for bit.
// ------------------------------------------------------------------------
// Edges of the cell at c (index p along D): left edges to lds[n*ld], right
// edges to lds[(9+n)*ld].
template <int D, class R>
RH_DEV void ppm_edges_lds(const double* __restrict__ W, long long N, long long c, long long s, int p,
                          const double* __restrict__ dxs, const double* __restrict__ xs,
                          const double* __restrict__ pc, long long npc, R gam, R* lds, int ld) {
  PpmCoef<R> kc;
  ppm_coef<R>(pc, npc, p, kc);
  R f2, eta;
  ppm_shape<D, R>(W, N, c, s, p, dxs, xs, gam, f2, eta);
  R cur[5], nxt[5];
#pragma unroll
  for (int o = 0; o < 5; o++) cur[o] = R(W[c + (o - 2) * s]);
#pragma unroll 1
  for (int n = 0; n < 9; n++) {
    if (n < 8) {
      const double* Pn = W + (long long)(n + 1) * N + c;
#pragma unroll
      for (int o = 0; o < 5; o++) nxt[o] = R(Pn[(o - 2) * s]);
    }
    R l, r;
    ppm_var<D, R>(cur, kc, n, f2, eta, l, r);
    lds[n * ld] = l;
    lds[(9 + n) * ld] = r;
#pragma unroll
    for (int o = 0; o < 5; o++) cur[o] = nxt[o];
  }
}

template <int D, class R, int ICHA>
RH_DEV void b_psweep(const Geo& G, const PSweepArgs& A, long long blk) {
  extern __shared__ unsigned char rh_dyn_lds[];  // 18 planes of blockDim.x values of R
  R* const L = reinterpret_cast<R*>(rh_dyn_lds);
  const int bt = blockDim.x, tid = threadIdx.x, lane = tid & 63;
  const long long N = G.N;
  const double* __restrict__ W = A.W;
  R* __restrict__ F = (R*)A.F;
  R* __restrict__ CTsrc = (R*)A.CTsrc;
  const R gam = R(A.gam);
  const int lj = jlo(G, G.nm0), hj = jhi(G, G.nj - G.nm0);
  const long long wv = (blk * bt + tid) >> 6;
  if (D == 1) {
    const long long nic = (G.ni - 2 * G.nm0 + 1 + 62) / 63, cj = hj - lj + 1, ck = G.nk - 2 * G.nm0 + 1;
    const long long ic = wv % nic, r = wv / nic;
    const int j = lj + (int)(r % cj);
    const long long kk = r / cj;
    const int k = G.nm0 + (int)kk;
    const int i = G.nm0 + (int)(ic * 63) + lane;
    const bool cell_in = kk < ck && i <= G.ni - G.nm0 + 1;
    const long long c = cell_in ? cidx(i, j, k, G) : 0;
    if (cell_in) ppm_edges_lds<1, R>(W, N, c, 1, i, A.dxs, A.xs, A.pc, G.ni + 2, gam, L + tid, bt);
    __syncthreads();  // every thread of the block reaches it: no early exit above
    if (cell_in && lane < 63 && i <= G.ni - G.nm0) {
      R wl[9], wr[9];
#pragma unroll
      for (int n = 0; n < 9; n++) { wl[n] = L[(9 + n) * bt + tid]; wr[n] = L[n * bt + tid + 1]; }
      const Met M{A.metF, G.P2, q2(i, k, G)};
      sweep_face<1, R, ICHA>(G, F, CTsrc, c, i, j, k, M, gam, wl, wr);
    }
    return;
  }
  // D = 2, 3: a wave holds TI consecutive i (fastest) by TP = 64 / TI
  // consecutive cells along D, so each load instruction reads TP runs of TI
  // contiguous values; lane l's interface takes the left edges of lane
  // l + TI's cell.  A wave's TP cells give TP - 1 interfaces (consecutive
  // waves along D overlap by one cell).  A collapsed model's one phi row is
  // its own neighbour along j (stride 0): there a wave holds 64 i of that
  // row, and each interface takes both edges of its own cell.
  const bool self = D == 2 && G.jfix;
  const int TP = self ? 1 : A.seg, TI = 64 / TP, IPW = self ? 1 : TP - 1;
  const int il = lane % TI, pl = lane / TI;
  const int plo = (D == 2) ? lj : G.nm0, phi = (D == 2) ? hj : G.nk - G.nm0;
  const int tlo = (D == 2) ? G.nm0 : lj, thi = (D == 2) ? G.nk - G.nm0 : hj;
  const long long nig = (G.ni - 2 * G.nm0 + 1 + TI - 1) / TI;
  const long long ndc = (phi - plo + 1 + IPW - 1) / IPW, ct = thi - tlo + 1;
  // consecutive waves along i (a workgroup's waves share cache lines), then
  // along D, then across (measured fastest for both directions: along k,
  // walking j before the chunks keeps the face metric in L2 but loses more
  // to the stencil rows)
  const long long ig = wv % nig, r = wv / nig, dc = r % ndc, tt = r / ndc;
  const int i = G.nm0 + (int)ig * TI + il;
  const int p = plo + (int)dc * IPW + pl;
  const int t = tlo + (int)tt;
  const bool cell_in = tt < ct && dc < ndc && i <= G.ni - G.nm0 && p <= phi + (self ? 0 : 1);
  const long long s = (D == 2) ? jstride(G) : kstride(G);
  const long long npc = ((D == 2) ? G.nj : G.nk) + 2;
  const long long c = cell_in ? ((D == 2) ? cidx(i, p, t, G) : cidx(i, t, p, G)) : 0;
  if (cell_in) ppm_edges_lds<D, R>(W, N, c, s, p, A.dxs, A.xs, A.pc, npc, gam, L + tid, bt);
  __syncthreads();  // every thread of the block reaches it: no early exit above
  if (cell_in && (self || pl < TP - 1) && p <= phi) {
    const int nb = self ? tid : tid + TI;
    R wl[9], wr[9];
#pragma unroll
    for (int n = 0; n < 9; n++) { wl[n] = L[(9 + n) * bt + tid]; wr[n] = L[n * bt + nb]; }
    const int j = (D == 2) ? p : t, k = (D == 2) ? t : p;
    const Met M{A.metF, G.P2, q2(i, k, G)};
    sweep_face<D, R, ICHA>(G, F, CTsrc, c, i, j, k, M, gam, wl, wr);
  }
}

// ------------------------------------------------------------------------
// Flux-CT (Toth 2000): fluxct, ct.f90:30-118.  Reads the six transverse B
// fluxes from CTsrc (planes 0 ww(2,7) 1 ww(3,7) 2 ww(1,8) 3 ww(3,8) 4 ww(1,9)
// 5 ww(2,9), as the sweeps left them) and writes the averaged values into
// the flux array in place, like ct.f90:98-109.  FT: the stored type (float
// in the mixed form); the averages are formed in double.
// ------------------------------------------------------------------------
template <class FT>
RH_DEV void b_fluxct(const Geo& G, const CtArgs& A, long long blk) {
  const FT* __restrict__ CTsrc = (const FT*)A.CTsrc;
  FT* __restrict__ F1 = (FT*)A.F1;
  FT* __restrict__ F2 = (FT*)A.F2;
  FT* __restrict__ F3 = (FT*)A.F3;
  int i, j, k;
  const long long t = blk * blockDim.x + threadIdx.x;
  if (!box_index(t, G.nm1, G.ni - G.nm1, jlo(G, G.nm1), jhi(G, G.nj - G.nm1), G.nm1, G.nk - G.nm1, i, j, k)) return;
  const long long N = G.N, c = cidx(i, j, k, G);
  const long long si = 1, sj = jstride(G), sk = kstride(G);
  const FT* fyx = CTsrc + 0 * N;  // ww(2,7)
  const FT* fzx = CTsrc + 1 * N;  // ww(3,7)
  const FT* fxy = CTsrc + 2 * N;  // ww(1,8)
  const FT* fzy = CTsrc + 3 * N;  // ww(3,8)
  const FT* fxz = CTsrc + 4 * N;  // ww(1,9)
  const FT* fyz = CTsrc + 5 * N;  // ww(2,9)
  const double c8 = 1.0 / 8.0;
  // 2.0 * x leads each sum, so every term is added in double
  double v[6];
  v[0] = c8 * (2.0 * fyx[c] + fyx[c + si] + fyx[c - si] - fxy[c] - fxy[c + sj] - fxy[c - si] - fxy[c - si + sj]);
  v[1] = c8 * (2.0 * fzx[c] + fzx[c + si] + fzx[c - si] - fxz[c] - fxz[c + sk] - fxz[c - si] - fxz[c - si + sk]);
  v[2] = c8 * (2.0 * fxy[c] + fxy[c + sj] + fxy[c - sj] - fyx[c] - fyx[c + si] - fyx[c - sj] - fyx[c + si - sj]);
  v[3] = c8 * (2.0 * fzy[c] + fzy[c + sj] + fzy[c - sj] - fyz[c] - fyz[c + sk] - fyz[c - sj] - fyz[c - sj + sk]);
  v[4] = c8 * (2.0 * fxz[c] + fxz[c + sk] + fxz[c - sk] - fzx[c] - fzx[c + si] - fzx[c - sk] - fzx[c + si - sk]);
  v[5] = c8 * (2.0 * fyz[c] + fyz[c + sk] + fyz[c - sk] - fzy[c] - fzy[c + sj] - fzy[c - sk] - fzy[c + sj - sk]);
  F2[6 * N + c] = FT(v[0]);
  F3[6 * N + c] = FT(v[1]);
  F1[7 * N + c] = FT(v[2]);
  F3[7 * N + c] = FT(v[3]);
  F1[8 * N + c] = FT(v[4]);
  F2[8 * N + c] = FT(v[5]);
  // the mixed form keeps the CT update of B in FP64: the update reads these
  // (the float copies above are what a later sweep keeps where its wave
  // speeds are unordered, as ww keeps what ct wrote)
  if (sizeof(FT) == 4 && A.CTd) {
#pragma unroll
    for (int q = 0; q < 6; q++) A.CTd[(long long)q * N + c] = v[q];
  }
}

// ------------------------------------------------------------------------
// Recovery of primitives, recov1a (recov.f90:578-1228), one interior cell.
// u[9] = stage conserved (after RK + sources); wold = previous primitives
// (initial guess and fallback).  Writes w[9].  The initial W of a cell whose
// quadratic has no non-negative root is inherited from the previous cell in
// the reference's loop order; that cell is flagged in *carry.
// ------------------------------------------------------------------------
struct RecovOut { double w[9]; int fallback, fail, negrho, negp, vsqgt1, det0, guard, indep; };

RH_DEV void recov_cell(const double u[9], const double wold[9], const Met& M, double gam, double dmin,
                       double pmin, int iter, double x1pow15, double x1pow25, double wg_in, bool& carry,
                       RecovOut& o, double& wg_out) {
  const double safty = 1.0e-4, saftybig = 1.0e5, dv = 1.0e-2;
  const double gam1 = gam / (gam - 1.0);
  const double gam2 = 1.0 / gam1;
  o.fallback = o.fail = o.negrho = o.negp = o.vsqgt1 = o.det0 = o.guard = 0;
  double delta = 1.0, deltaend = 1.0e-7;
  int irecov = 0;
  const double alpha1 = 1.0 / sqrt(-M.gi(0, 0));
  const double detg1 = M.detg();
  const double dro = alpha1 * u[0] * (1.0 / detg1);
  double qcov[4], qcon[4], bcon[4], bcov[4], oncov[4], oncon[4];
  qcov[0] = alpha1 * (u[4] - u[0]) * (1.0 / detg1);
  qcov[1] = alpha1 * u[1] * (1.0 / detg1);
  qcov[2] = alpha1 * u[2] * (1.0 / detg1);
  qcov[3] = alpha1 * u[3] * (1.0 / detg1);
  const double gss = alpha1 * u[5] * (1.0 / detg1);
  bcon[0] = 0.0;
  bcon[1] = alpha1 * u[6] * (1.0 / detg1);
  bcon[2] = alpha1 * u[7] * (1.0 / detg1);
  bcon[3] = alpha1 * u[8] * (1.0 / detg1);
  double util[4] = {0.0, wold[1], wold[2], wold[3]};
  const double ro1 = wold[0], pp1 = wold[4];
  lower(bcon, bcov, M);
  upper(qcov, qcon, M);
  const double bsq = bcon[1] * bcov[1] + bcon[2] * bcov[2] + bcon[3] * bcov[3];
  const double qdotb = qcov[0] * bcon[0] + qcov[1] * bcon[1] + qcov[2] * bcon[2] + qcov[3] * bcon[3];
  const double qdotbsq = qdotb * qdotb;
  oncov[0] = -1.0 * alpha1; oncov[1] = 0.0; oncov[2] = 0.0; oncov[3] = 0.0;
  upper(oncov, oncon, M);
  const double qdotn = qcon[0] * oncov[0];
  const double qsq = qcov[0] * qcon[0] + qcov[1] * qcon[1] + qcov[2] * qcon[2] + qcov[3] * qcon[3];
  const double qtsq = qsq + (qdotn * qdotn);
  double utsq = 0.0;
#pragma unroll
  for (int m = 1; m <= 3; m++)
#pragma unroll
    for (int n = 1; n <= 3; n++) utsq = utsq + M.g(m, n) * util[m] * util[n];
  if (utsq < 0.0 && fabs(utsq) <= safty) utsq = fabs(utsq);
  if (fabs(utsq) > 1.0e4) utsq = 1.0e4;
  double gfsq = 1.0 + utsq;
  double gf = sqrt(gfsq);
  double vsq = (gfsq - 1.0) * (1.0 / gfsq);
  const double t1a = 4.0 - vsq;
  const double t1b = 4.0 * (bsq + qdotn);
  const double t1c = qtsq + bsq * bsq + 2.0 * bsq * qdotn;
  double t1d = (t1b * t1b) - 4.0 * t1a * t1c;
  if (t1d <= 0.0) t1d = fabs(t1d);
  const double wgp = (-t1b + sqrt(t1d)) * (1.0 / (2.0 * t1a));
  const double wgm = (-t1b - sqrt(t1d)) * (1.0 / (2.0 * t1a));
  double wg = wg_in;
  carry = false;
  // With a NaN among these the result and the final W are NaN whatever W the
  // cell starts from, so an inherited guess cannot change it.
  o.indep = isnan(dro) || isnan(qtsq) || isnan(qdotn) || isnan(qdotbsq) || isnan(bsq);
  if (wgp > 0.0) wg = wgp;
  else if (wgm >= 0.0) wg = wgm;
  else carry = true;
  double pp = 0.0, ro = 0.0;
  // recov.f90:823-850 ("goto 95"); bounded here, bound hits are counted
  int guard = 0;
  for (; guard < 100000; guard++) {
    pp = gam2 * (wg * gf - dro * sqrt(gf));
    const double t3a = (wg * wg * wg) * (wg + 2.0 * bsq) - qdotbsq * (2.0 * wg + bsq);
    const double t3b = (wg * wg) * (qtsq - bsq * bsq);
    const double t3c = (wg * wg * wg) + (wg * wg) * (0.5 * bsq + qdotn - pp) - 0.5 * qdotbsq;
    if (t3a <= t3b) { wg = fabs(1.5 * wg); continue; }
    if (t3c > 0.0) { wg = fabs((double)0.9f * wg); continue; }
    break;
  }
  if (guard >= 100000) o.guard = 1;
  double wsq = wg * wg;
  const double xsq = (bsq + wg) * (bsq + wg);
  vsq = fabs((wsq * qtsq + qdotbsq * (bsq + 2.0 * wg)) * (1.0 / (wsq * xsq)));
  if (vsq > 1.0) { o.vsqgt1 = 1; vsq = 1.0 - dv; }
  const double vsqo = vsq;
  double xxn = wg, yyn = vsq, w3, gtmp;
  // 2D Newton-Raphson on (W, v^2): recov.f90:877-960.  Trips after convergence
  // change nothing in the reference, so the loop stops there.  A cell whose
  // iteration never meets the tolerance runs all iter trips in the reference;
  // such an iteration usually falls into a short exact cycle of states.  Once
  // a state repeats (Brent's cycle detection) while the tolerance still fails,
  // every later trip repeats the cycle, so whole periods are skipped (with
  // their det=0 events) and the last trips run as usual: the loop ends in the
  // state, and with the count, of the reference's full run.
  unsigned long long cyx = (unsigned long long)__double_as_longlong(xxn);
  unsigned long long cyy = (unsigned long long)__double_as_longlong(yyn);
  int cy_pow = 1, cy_len = 0, cy_det0 = 0;
  for (int nnn = 1; nnn <= iter; nnn++) {
    if (!(fabs(delta) >= deltaend && irecov == 0)) break;
    if (nnn == iter) { irecov = 1; break; }
    wg = xxn; vsq = yyn;
    wsq = wg * wg; w3 = wsq * wg;
    gtmp = 1.0 - vsq;
    if (gtmp < 0.0 || gtmp < safty) { gtmp = safty; vsq = 1.0 - gtmp; }
    const double ptmp = gam2 * (wg * gtmp - dro * sqrt(gtmp));
    const double f1 = qtsq - vsq * (wg + bsq) * (wg + bsq) + (qdotbsq * (bsq + 2.0 * wg)) * (1.0 / wsq);
    const double g1 = -qdotn - 0.5 * bsq * (1.0 + vsq) + 0.5 * (qdotbsq * (1.0 / wsq)) - wg + ptmp;
    const double dpdw = gam2 * gtmp;
    const double dpdvsq = -gam2 * wg + 0.5 * gam2 * dro * (1.0 / (sqrt(gtmp)));
    const double dfx = -2.0 * (wg + bsq) * (vsq + qdotbsq * (1.0 / w3));
    const double dfy = -((wg + bsq) * (wg + bsq));
    const double dgx = -1.0 - (qdotbsq * (1.0 / w3)) + dpdw;
    const double dgy = -0.5 * bsq + dpdvsq;
    double det = dfx * dgy - dfy * dgx;
    if (det == 0.0) { o.det0 += 1; det = 1.0; }
    const double dx = (-dgy * f1 + dfy * g1) * (1.0 / det);
    const double dy = (dgx * f1 - dfx * g1) * (1.0 / det);
    delta = 0.5 * (fabs(dx) + fabs(dy));
    xxn = wg + dx;
    yyn = vsq + dy;
    if (xxn < 0.0) xxn = fabs(xxn);
    if (xxn > saftybig) xxn = saftybig;
    if (yyn < 0.0) yyn = fabs(yyn);
    if (fabs(yyn) > 1.0) yyn = 1.0 - dv;
    const unsigned long long bx = (unsigned long long)__double_as_longlong(xxn);
    const unsigned long long by = (unsigned long long)__double_as_longlong(yyn);
    cy_len++;
    if (bx == cyx && by == cyy && fabs(delta) >= deltaend) {
      const int periods = (iter - 1 - nnn) / cy_len;
      nnn += periods * cy_len;
      o.det0 += periods * (o.det0 - cy_det0);
      cy_pow = 0x40000000;
    } else if (cy_len == cy_pow) {
      cyx = bx; cyy = by; cy_pow <<= 1; cy_len = 0; cy_det0 = o.det0;
    }
  }
  if (irecov == 0) {
    wg = xxn; vsq = yyn;
    gtmp = 1.0 - vsq;
    if (gtmp < 0.0 || gtmp < safty) { gtmp = safty; vsq = 1.0 - gtmp; }
    gf = 1.0 / sqrt(gtmp);
    ro = dro * (1.0 / gf);
    const double w2 = wg * (gtmp);
    pp = gam2 * (w2 - ro);
    if (ro < 0.0 || pp < 0.0) irecov = 1;
    if (irecov == 0) {
#pragma unroll
      for (int m = 1; m <= 3; m++) {
        const double qt = qcon[m] + oncon[m] * qdotn;
        util[m] = (gf * (1.0 / (wg + bsq))) * (qt + qdotb * bcon[m] * (1.0 / wg));
      }
    }
    wg = xxn;
  }
  if (irecov == 1) {
    // entropy-based 1D inversion on v^2: recov.f90:975-1147
    o.fallback = 1;
    delta = 1.0; deltaend = 1.0e-6; irecov = 0;
    xxn = vsqo;
    // same cycle skipping as the 2D iteration; the state is v^2 alone
    unsigned long long cyv = (unsigned long long)__double_as_longlong(xxn);
    int ey_pow = 1, ey_len = 0;
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
      const double roh = ro + gam1 * pp;
      wg = fabs(gfsq * roh);
      wsq = wg * wg; w3 = wsq * wg;
      const double dgfdvsq = 0.5 * pow(gtmp, -1.5);
      const double dwdvsq = dro * dgfdvsq + gam1 * gss * (pow(dro, gam - 1.0)) * (2.0 - gam) * (pow(gf, 1.0 - gam)) * dgfdvsq;
      const double f = qtsq - vsq * (wg + bsq) * (wg + bsq) + (qdotbsq * (bsq + 2.0 * wg)) / wsq;
      const double a = -((wg + bsq) * (wg + bsq));
      const double b = -2.0 * vsq * (wg + bsq);
      const double cc = -2.0 * qdotbsq * (bsq + 2.0 * wg) * (1.0 / w3);
      const double d = 2.0 * qdotbsq * (1.0 / wsq);
      const double df = a + (b + cc + d) * dwdvsq;
      delta = f / df;
      xxn = vsq - delta;
      if (xxn < 0.0) xxn = fabs(xxn);
      if (fabs(xxn) >= 1.0) xxn = 1.0 - dv;
      const unsigned long long bv = (unsigned long long)__double_as_longlong(xxn);
      ey_len++;
      if (bv == cyv && fabs(delta) >= deltaend) {
        nnn += ((iter - 1 - nnn) / ey_len) * ey_len;
        ey_pow = 0x40000000;
      } else if (ey_len == ey_pow) {
        cyv = bv; ey_pow <<= 1; ey_len = 0;
      }
    }
    if (irecov == 0) {
      vsq = xxn;
      gtmp = 1.0 - vsq;
      if (gtmp < 0.0 || gtmp < safty) gtmp = safty;
      gf = 1.0 / sqrt(gtmp);
      gfsq = gf * gf;
      ro = dro * (1.0 / gf);
      pp = (gss * (1.0 / gf)) * (pow(ro, gam - 1.0));
      const double roh = ro + gam1 * pp;
      if (ro < 0.0) { o.negrho = 1; ro = dmin * x1pow15; }
      if (pp < 0.0) { o.negp = 1; pp = (gam - 1.0) * pmin * x1pow25; }
      wg = fabs(gfsq * roh);
#pragma unroll
      for (int m = 1; m <= 3; m++) {
        const double qt = qcon[m] + oncon[m] * qdotn;
        util[m] = (gf / (wg + bsq)) * (qt + qdotb * bcon[m] * (1.0 / wg));
      }
    }
  }
  if (irecov == 0) {
    o.w[0] = ro; o.w[1] = util[1]; o.w[2] = util[2]; o.w[3] = util[3]; o.w[4] = pp;
    o.w[5] = ro * pp / pow(ro, gam - 1.0);
  } else {
    o.fail = 1;
    o.w[0] = ro1; o.w[1] = wold[1]; o.w[2] = wold[2]; o.w[3] = wold[3]; o.w[4] = pp1;
    o.w[5] = ro1 * pp1 / pow(ro1, gam);
  }
  o.w[6] = u[6] * (1.0 / detg1);
  o.w[7] = u[7] * (1.0 / detg1);
  o.w[8] = u[8] * (1.0 / detg1);
  wg_out = wg;
}

// ------------------------------------------------------------------------
// Recovery of primitives, recov2 (recov.f90:1919-2454: Mignone & McKinney
// 2007, one Newton-Raphson unknown W' = W - D, gamma-law EoS), one interior
// cell; the same interface as recov_cell.  The initial W of a cell whose
// quadratic has no admissible root is the previous cell's final W.  recov2
// prints "vsq > 1.0" at every trip that clamps v^2; vsqgt1 counts them.
// ------------------------------------------------------------------------
RH_DEV void recov2_cell(const double u[9], const double wold[9], const Met& M, double gam, double dmin,
                        double pmin, int iter, double x1pow15, double x1pow25, double wg_in, bool& carry,
                        RecovOut& o, double& wg_out) {
  const double safty = 1.0e-5, dv = 1.0e-4;
  const double gam1 = gam / (gam - 1.0);
  const double gam2 = 1.0 / gam1;
  o.fallback = o.fail = o.negrho = o.negp = o.vsqgt1 = o.det0 = o.guard = 0;
  double delta = 1.0;
  const double deltaend = 1.0e-8;
  int irecov = 0;
  const double alpha1 = 1.0 / sqrt(-M.gi(0, 0));
  const double detg1 = M.detg();
  const double dro = alpha1 * u[0] / detg1;
  double qcov[4], qcon[4], bcon[4], bcov[4], oncov[4], oncon[4];
  qcov[0] = alpha1 * (u[4] - u[0]) * (1.0 / detg1);
  qcov[1] = alpha1 * u[1] * (1.0 / detg1);
  qcov[2] = alpha1 * u[2] * (1.0 / detg1);
  qcov[3] = alpha1 * u[3] * (1.0 / detg1);
  bcon[0] = 0.0;
  bcon[1] = alpha1 * u[6] * (1.0 / detg1);
  bcon[2] = alpha1 * u[7] * (1.0 / detg1);
  bcon[3] = alpha1 * u[8] * (1.0 / detg1);
  double util[4] = {0.0, wold[1], wold[2], wold[3]};
  const double ro1 = wold[0], pp1 = wold[4];
  lower(bcon, bcov, M);
  upper(qcov, qcon, M);
  const double bsq = bcon[1] * bcov[1] + bcon[2] * bcov[2] + bcon[3] * bcov[3];
  const double qdotb = qcov[0] * bcon[0] + qcov[1] * bcon[1] + qcov[2] * bcon[2] + qcov[3] * bcon[3];
  const double qdotbsq = qdotb * qdotb;
  oncov[0] = -1.0 * alpha1; oncov[1] = 0.0; oncov[2] = 0.0; oncov[3] = 0.0;
  upper(oncov, oncon, M);
  const double qdotn = qcon[0] * oncov[0];
  const double qsq = qcov[0] * qcon[0] + qcov[1] * qcon[1] + qcov[2] * qcon[2] + qcov[3] * qcon[3];
  const double qtsq = qsq + (qdotn * qdotn);
  double utsq = 0.0;
#pragma unroll
  for (int m = 1; m <= 3; m++)
#pragma unroll
    for (int n = 1; n <= 3; n++) utsq = utsq + M.g(m, n) * util[m] * util[n];
  if (utsq < 0.0 && fabs(utsq) <= safty) utsq = fabs(utsq);
  if (utsq < 0.0 || fabs(utsq) > 1.0e4) utsq = 1.0e4;
  double gfsq = 1.0 + utsq;
  double gf = sqrt(gfsq);
  double vsq = (gfsq - 1.0) * (1.0 / gfsq);
  const double t1a = 4.0 - vsq;
  const double t1b = 4.0 * (bsq + qdotn);
  const double t1c = qtsq + bsq * bsq + 2.0 * bsq * qdotn;
  double t1d = (t1b * t1b) - 4.0 * t1a * t1c;
  if (t1d <= 0.0) t1d = fabs(t1d);
  const double wgp = (-t1b + sqrt(t1d)) * (1.0 / (2.0 * t1a));
  const double wgm = (-t1b - sqrt(t1d)) * (1.0 / (2.0 * t1a));
  double wg = wg_in;
  carry = false;
  // as in recov_cell: NaN inputs make the result independent of the guess
  o.indep = isnan(dro) || isnan(qtsq) || isnan(qdotn) || isnan(qdotbsq) || isnan(bsq);
  if (wgp > 0.0) wg = wgp;
  else if (wgm >= 0.0) wg = wgm;
  else carry = true;
  double pp = 0.0, ro = 0.0;
  // recov.f90:2150-2172 ("goto 95"); bounded here, bound hits are counted
  int guard = 0;
  for (; guard < 100000; guard++) {
    pp = gam2 * (wg * gf - dro * sqrt(gf));
    const double t3a = (wg * wg * wg) * (wg + 2.0 * bsq) - qdotbsq * (2.0 * wg + bsq);
    const double t3b = (wg * wg) * (qtsq - bsq * bsq);
    const double t3c = (wg * wg * wg) + (wg * wg) * (0.5 * bsq + qdotn - pp) - 0.5 * qdotbsq;
    if (t3a <= t3b) { wg = fabs(1.5 * wg); continue; }
    if (t3c > 0.0) { wg = fabs((double)0.9f * wg); continue; }
    break;
  }
  if (guard >= 100000) o.guard = 1;
  double wgd = wg - dro;
  double xxn = wgd;
  // Newton-Raphson on W' (recov.f90:2180-2300); a trip is a function of xxn
  // alone, so a repeated xxn while the tolerance fails is an exact cycle and
  // whole periods are skipped (with their "vsq > 1" events), as in recov_cell.
  unsigned long long cyx = (unsigned long long)__double_as_longlong(xxn);
  int cy_pow = 1, cy_len = 0, cy_vsq = 0;
  for (int nnn = 1; nnn <= iter; nnn++) {
    if (!(fabs(delta) >= deltaend && irecov == 0)) break;
    if (nnn == iter) { irecov = 1; break; }
    wgd = xxn;
    wg = wgd + dro;
    const double wsq = wg * wg;
    const double w3 = wsq * wg;
    const double xsq = (bsq + wg) * (bsq + wg);
    vsq = fabs((wsq * qtsq + qdotbsq * (bsq + 2.0 * wg)) * (1.0 / (wsq * xsq)));
    if (fabs(vsq) > 1.0) { o.vsqgt1 += 1; vsq = 1.0 - dv; }
    if (vsq < 0.0) vsq = fabs(vsq);
    double gtmp = 1.0 - vsq;
    if (gtmp < 0.0 || gtmp < safty) { gtmp = safty; vsq = 1.0 - gtmp; }
    const double usq = (1.0 - gtmp) / (gtmp);
    gf = sqrt(1.0 + usq);
    gfsq = gf * gf;
    ro = dro / gf;
    const double xi = (wg - dro * gf) * (1.0 / gfsq);
    pp = gam2 * xi;
    const double dpdxi = gam2, dpdro = 0.0;
    const double ta = (bsq + wg);
    const double tb = (bsq + wg) * (bsq + wg) * (bsq + wg);
    const double tc = (qtsq == 0.0) ? 0.0 : -2.0 * qtsq * (1.0 / tb);
    const double td = (qdotbsq == 0.0) ? 0.0 : -2.0 * qdotbsq * (3.0 * wg * ta + bsq * bsq) * (1.0 / (tb * w3));
    const double dvdwd = tc + td;
    const double dxidwd = (1.0 / gfsq) - 0.5 * gf * (dro + 2.0 * gf * xi) * dvdwd;
    const double drodwd = -0.5 * dro * gf * dvdwd;
    const double dpdwd = dpdxi * dxidwd + dpdro * drodwd;
    const double t5a = bsq * qtsq - qdotbsq;
    const double t5b = bsq + wgd + dro;
    double t5c = 0.0, t5d = 0.0;
    if (t5a != 0.0) {
      t5c = 0.5 * t5a * (1.0 / (t5b * t5b));
      t5d = t5a * (1.0 / (t5b * t5b * t5b));
    }
    const double f = -(qdotn + dro) - wgd + pp - 0.5 * bsq - t5c;
    const double df = -1.0 + dpdwd + t5d;
    delta = f / df;
    xxn = wgd - delta;
    if (xxn < 0.0) xxn = fabs(xxn);
    if (xxn > 1.0e8) xxn = 1.0e8;
    const unsigned long long bx = (unsigned long long)__double_as_longlong(xxn);
    cy_len++;
    if (bx == cyx && fabs(delta) >= deltaend) {
      const int periods = (iter - 1 - nnn) / cy_len;
      nnn += periods * cy_len;
      o.vsqgt1 += periods * (o.vsqgt1 - cy_vsq);
      cy_pow = 0x40000000;
    } else if (cy_len == cy_pow) {
      cyx = bx; cy_pow <<= 1; cy_len = 0; cy_vsq = o.vsqgt1;
    }
  }
  if (irecov == 0) {
    wgd = xxn;
    wg = wgd + dro;
    const double wsq = wg * wg;
    const double xsq = (bsq + wg) * (bsq + wg);
    vsq = fabs((wsq * qtsq + qdotbsq * (bsq + 2.0 * wg)) * (1.0 / (wsq * xsq)));
    if (vsq < 0.0) vsq = fabs(vsq);
    if (vsq > 1.0) { o.vsqgt1 += 1; vsq = 1.0 - dv; }
    double gtmp = 1.0 - vsq;
    if (gtmp < safty) gtmp = safty;
    const double usq = (1.0 - gtmp) / gtmp;
    gf = sqrt(1.0 + usq);
    gfsq = gf * gf;
    ro = dro / gf;
    const double xi = wg * (1.0 / gfsq) - ro;
    pp = gam2 * xi;
    if (ro < 0.0) { o.negrho = 1; ro = dmin * x1pow15; }
    if (pp < 0.0) { o.negp = 1; pp = pmin * (gam - 1.0) * x1pow25; }
#pragma unroll
    for (int m = 1; m <= 3; m++) {
      const double qt = qcon[m] + oncon[m] * qdotn;
      util[m] = (gf * (1.0 / (wg + bsq))) * (qt + qdotb * bcon[m] * (1.0 / wg));
    }
    o.w[0] = ro; o.w[1] = util[1]; o.w[2] = util[2]; o.w[3] = util[3]; o.w[4] = pp;
    o.w[5] = pp / pow(ro, gam - 1.0);
  } else {
    o.fail = 1;
    o.w[0] = ro1; o.w[1] = wold[1]; o.w[2] = wold[2]; o.w[3] = wold[3]; o.w[4] = pp1;
    o.w[5] = pp1 / pow(ro1, gam - 1.0);
  }
  o.w[6] = u[6] * (1.0 / detg1);
  o.w[7] = u[7] * (1.0 / detg1);
  o.w[8] = u[8] * (1.0 / detg1);
  wg_out = wg;
}

// one cell of the recovery iwvec selects
template <int IW>
RH_DEV void recov_any(const double u[9], const double wold[9], const Met& M, double gam, double dmin, double pmin,
                      int iter, double x1pow15, double x1pow25, double wg_in, bool& carry, RecovOut& o,
                      double& wg_out) {
  if constexpr (IW == 2) recov2_cell(u, wold, M, gam, dmin, pmin, iter, x1pow15, x1pow25, wg_in, carry, o, wg_out);
  else recov_cell(u, wold, M, gam, dmin, pmin, iter, x1pow15, x1pow25, wg_in, carry, o, wg_out);
}

// ------------------------------------------------------------------------
// Update of one interior cell: flux divergence (with CT'd B fluxes), the
// geometric source (caltenr + calsf, calsf.f90:6-151), the TVD-RK stage
// (rkt.f90:2-180) plus its source add (rkt.f90:183-320), then recov1a.
// STAGE: 1 rk2fst, 2 rk2snd, 3 rk3snd, 4 rk3trd.
// ------------------------------------------------------------------------
// Plane of Gamma^l_{ab} in the Christoffel tables (christ(l,a,b) at l+4a+16b).
// SYM: the tables are bit-for-bit symmetric in (a,b) (checked on the host),
// so (a,b) and (b,a) read one plane and the compiler loads it once.
template <bool SYM>
RH_DEV long long chr_plane(int l, int a, int b) {
  if (SYM && a > b) { const int t = a; a = b; b = t; }
  return (long long)(l + 4 * a + 16 * b);
}

template <int STAGE, bool SYM, class FT>
RH_DEV void b_rk(const Geo& G, const StageArgs& A, long long blk) {
  const FT* __restrict__ F1 = (const FT*)A.F1;
  const FT* __restrict__ F2 = (const FT*)A.F2;
  const FT* __restrict__ F3 = (const FT*)A.F3;
  int i, j, k;
  const long long t = blk * blockDim.x + threadIdx.x;
  if (!box_index(t, G.nm1 + 1, G.ni - G.nm1, jlo(G, G.nm1 + 1), jhi(G, G.nj - G.nm1), G.nm1 + 1, G.nk - G.nm1, i, j,
                 k))
    return;
  const long long N = G.N, c = cidx(i, j, k, G);
  const long long sj = jstride(G), sk = kstride(G);
  const double gam = A.gam;
  // dt and dt/dx (main.f90:760-780), recomputed per thread from the CFL minimum
  double dt, a1, a2, a3;
  if (A.ak1) {  // akap*b and dt handed in by the caller (main.f90:760-780 on the host)
    dt = A.dt_host;
    a1 = A.ak1[i - 1]; a2 = A.ak2[j - 1]; a3 = A.ak3[k - 1];
  } else {
    const double dtcfl = __longlong_as_double((long long)A.dtbits[0]);
    dt = A.cfl * dtcfl;
    a1 = dt / A.dx1b[i - 1]; a2 = dt / A.dx2b[j - 1]; a3 = dt / A.dx3b[k - 1];
  }
  const Met M{A.metC, G.P2, q2(i, k, G)};

  double w[9];
#pragma unroll
  for (int n = 0; n < 9; n++) w[n] = A.W[(long long)n * N + c];

  // geometric source: sf(2..5) = sqrt(-g) T^m_n Gamma^n_{(1,2,3,0) m}
  double sf[4];
  {
    Prel s;
    prelude(w, gam, M, s);
    double ta = 0.0, tb = 0.0, tc = 0.0, td = 0.0;
    const long long q = q2(i, k, G);
#pragma unroll
    for (int n = 0; n < 4; n++)
#pragma unroll
      for (int m = 0; m < 4; m++) {
        const double dl = (m == n) ? 1.0 : 0.0;
        const double tv = (s.roh + s.bbsq) * s.ucon[m] * s.ucov[n] + (s.pr + 0.5 * s.bbsq) * dl - s.bb[m] * s.bbcov[n];
        ta = ta + tv * A.chr[chr_plane<SYM>(n, 1, m) * G.P2 + q];
        tb = tb + tv * A.chr[chr_plane<SYM>(n, 2, m) * G.P2 + q];
        tc = tc + tv * A.chr[chr_plane<SYM>(n, 3, m) * G.P2 + q];
        td = td + tv * A.chr[chr_plane<SYM>(n, 0, m) * G.P2 + q];
      }
    const double dg = M.detg();
    sf[0] = dg * ta; sf[1] = dg * tb; sf[2] = dg * tc; sf[3] = dg * td;
  }

  // transverse-B fluxes from flux-CT's FP64 planes in the mixed form
  // (component m, direction d -> plane; -1: the sweep's own flux)
  constexpr int ctp[3][3] = {{-1, 0, 1}, {2, -1, 3}, {4, 5, -1}};
  const bool ctd = sizeof(FT) == 4 && A.CTd;
#pragma unroll
  for (int m = 0; m < 9; m++) {
    const long long o = (long long)m * N + c;
    auto fl = [&](const FT* Fd, int d, long long off) -> double {
      if (m >= 6 && ctd && ctp[m - 6][d] >= 0) return A.CTd[(long long)ctp[m - 6][d] * N + c - off];
      return double(Fd[o - off]);
    };
    const double d1 = fl(F1, 0, 0) - fl(F1, 0, 1), d2 = fl(F2, 1, 0) - fl(F2, 1, sj),
                 d3 = fl(F3, 2, 0) - fl(F3, 2, sk);
    double v;
    if (STAGE == 1) {
      v = A.Ua[o] - a1 * d1 - a2 * d2 - a3 * d3;
    } else if (STAGE == 2) {
      v = (A.Ua[o] + A.Ub[o]) * 0.5 + (-a1 * d1 - a2 * d2 - a3 * d3) * 0.5;
    } else if (STAGE == 3) {
      v = (3.0 * A.Ua[o] + A.Ub[o]) * 0.25 + (-a1 * d1 - a2 * d2 - a3 * d3) * 0.25;
    } else {
      v = (A.Ua[o] + 2.0 * A.Ub[o]) * (double)(1.0f / 3.0f) + (-a1 * d1 - a2 * d2 - a3 * d3) * (double)(2.0f / 3.0f);
    }
    if (m >= 1 && m <= 4) {
      const double sdt = sf[m - 1] * dt;
      if (STAGE == 1) v = v + sdt;
      else if (STAGE == 2) v = v + sdt * 0.5;
      else if (STAGE == 3) v = v + sdt * 0.25;
      else v = v + sdt * (double)(2.0f / 3.0f);
    }
    A.Ukeep[o] = v;
  }
}

// recov1a (recov.f90:578-1228) or recov2 (recov.f90:1919-2454) on every
// interior cell from the stage's RK result; cells whose initial W is
// inherited are flagged for k_recov_carry.
template <int IW>
RH_DEV void b_recov(const Geo& G, const StageArgs& A, long long blk) {
  int i, j, k;
  const long long t = blk * blockDim.x + threadIdx.x;
  if (!box_index(t, G.nm1 + 1, G.ni - G.nm1, jlo(G, G.nm1 + 1), jhi(G, G.nj - G.nm1), G.nm1 + 1, G.nk - G.nm1, i, j,
                 k))
    return;
  const long long N = G.N, c = cidx(i, j, k, G);
  const Met M{A.metC, G.P2, q2(i, k, G)};
  double w[9], u[9];
#pragma unroll
  for (int n = 0; n < 9; n++) {
    w[n] = A.W[(long long)n * N + c];
    u[n] = A.Ukeep[(long long)n * N + c];
  }
  RecovOut ro;
  bool carry;
  double wgout;
  recov_any<IW>(u, w, M, A.gam, A.dmin, A.pmin, A.iter, A.r15[i - 1], A.r25[i - 1], __builtin_nan(""), carry, ro,
                wgout);
  const bool dep = carry && !ro.indep;
#pragma unroll
  for (int n = 0; n < 9; n++) A.Wpre[(long long)n * N + c] = ro.w[n];
  A.wgout[c] = wgout;
  // a flagged cell's outcome is counted when the fix-up computes it
  count_wave(A.cpart, blk, C_FALLBACK, ro.fallback && !dep);
  count_wave(A.cpart, blk, C_FAIL, ro.fail && !dep);
  count_wave(A.cpart, blk, C_NEGRHO, ro.negrho && !dep);
  count_wave(A.cpart, blk, C_NEGP, ro.negp && !dep);
  count_wave(A.cpart, blk, C_VSQGT1, ro.vsqgt1 && !dep);
  if (ro.vsqgt1 > 1 && !dep) atomicAdd(&A.cnt[C_VSQGT1], (unsigned long long)(ro.vsqgt1 - 1));  // recov2: per trip
  if (ro.det0 && !dep) atomicAdd(&A.cnt[C_DET0], (unsigned long long)ro.det0);  // rare: one event per det=0 trip
  count_wave(A.cpart, blk, C_WGCARRY, carry);
  count_wave(A.cpart, blk, C_GUARD, ro.guard && !dep);
  if (dep) {
    // remember the cell for the in-order resolution
    A.cflag[c] = 1;
    const unsigned long long slot = atomicAdd(&A.carry[0], 1ULL);
    if (slot < (unsigned long long)A.carry_cap) A.carry[1 + slot] = (unsigned long long)c;
  }
}

// ------------------------------------------------------------------------
// Ghost fill + floors + conserved variables, every cell.
// bnd4 (bnd.f90:4-1059) is composed per cell: the value a ghost holds after
// the periodic pass, then the x1, x2, x3 passes.  Then infchk (bnd.f90:1636),
// atmosphere1a (correction.f90:226), pminmax2b (correction.f90:744) and
// caluu1 (caluu.f90:2).
// ------------------------------------------------------------------------
struct BndCtx {
  const Geo* G;
  const double* Wpre;   // post-recov interior values
  const double* Wold;   // values before this stage (ghosts not rewritten keep these)
  const double* Wo;     // primitives at the start of the step (fixed-value type 2)
  const int* bc;        // [axis*2+side][m] boundary codes
};

RH_DEV bool is_ghost(int p, int n, int nm1) { return p <= nm1 || p > n - nm1; }

// value after the periodic pass (bndpefb4)
RH_DEV double bnd_s0(const BndCtx& B, int m, int i, int j, int k) {
  const Geo& G = *B.G;
  const int nm1 = G.nm1;
  const bool gi = is_ghost(i, G.ni, nm1), gj = is_ghost(j, G.nj, nm1), gk = is_ghost(k, G.nk, nm1);
  const int ng = (int)gi + (int)gj + (int)gk;
  if (ng == 0) return B.Wpre[(long long)m * G.N + cidx(i, j, k, G)];
  if (ng == 1) {
    if (gi && (B.bc[0 * 9 + m] == 1 || B.bc[1 * 9 + m] == 1)) {
      const int Wd = G.ni - 2 * nm1;
      const int ii = nm1 + 1 + (((i - (nm1 + 1)) % Wd) + Wd) % Wd;
      return B.Wpre[(long long)m * G.N + cidx(ii, j, k, G)];
    }
    if (gj && (B.bc[2 * 9 + m] == 1 || B.bc[3 * 9 + m] == 1)) {
      const int Wd = G.nj - 2 * nm1;
      const int jj = nm1 + 1 + (((j - (nm1 + 1)) % Wd) + Wd) % Wd;
      return B.Wpre[(long long)m * G.N + cidx(i, jj, k, G)];
    }
    if (gk && (B.bc[4 * 9 + m] == 1 || B.bc[5 * 9 + m] == 1)) {
      const int Wd = G.nk - 2 * nm1;
      const int kk = nm1 + 1 + (((k - (nm1 + 1)) % Wd) + Wd) % Wd;
      return B.Wpre[(long long)m * G.N + cidx(i, j, kk, G)];
    }
  }
  return B.Wold[(long long)m * G.N + cidx(i, j, k, G)];
}

// Value of variable m at ghost (i,j,k) after the three physical passes.  Each
// pass maps the cell to its source in the previous pass (copy, mirror, with a
// sign for antisymmetric codes) or ends in a constant: walking x3 -> x2 -> x1
// and then reading the periodic-pass value composes the passes exactly.
RH_DEV double bnd_value(const BndCtx& B, int m, int i, int j, int k) {
  const Geo& G = *B.G;
  const int nm1 = G.nm1;
  double sgn = 1.0;
  int c[4] = {0, i, j, k};
  const int n[4] = {0, G.ni, G.nj, G.nk};
#pragma unroll
  for (int ax = 3; ax >= 1; ax--) {
    const int p = c[ax];
    if (!is_ghost(p, n[ax], nm1)) continue;
    const bool outer = p > n[ax] - nm1;
    const int type = B.bc[((ax - 1) * 2 + (outer ? 1 : 0)) * 9 + m];
    switch (type) {
      case 2: return sgn * B.Wo[(long long)m * G.N + cidx(c[1], c[2], c[3], G)];
      case 5: c[ax] = outer ? n[ax] - nm1 : nm1 + 1; break;
      case 12: sgn = -sgn; c[ax] = outer ? 2 * n[ax] - 2 * nm1 + 1 - p : 2 * nm1 - p + 1; break;
      case 13: return sgn * 0.0;
      case 14: c[ax] = outer ? 2 * n[ax] - 2 * nm1 + 1 - p : 2 * nm1 - p + 1; break;
      case 18: {
        // bnd.f90:598-610 (x1 outer only): no inflow through the outer boundary
        const double v2 = bnd_s0(B, 1, n[ax] - nm1, c[2], c[3]);
        if (v2 < 0.0) return sgn * 0.0;
        c[ax] = n[ax] - nm1;
        break;
      }
      default: break;  // 1: periodic, filled by the periodic pass
    }
  }
  return sgn * bnd_s0(B, m, c[1], c[2], c[3]);
}

// 1-based (i,j,k) of the t-th ghost cell: the two theta slabs (all i, j),
// then the two phi slabs (interior k), then the two r slabs (interior j, k).
RH_DEV bool ghost_index(long long t, const Geo& G, int& i, int& j, int& k) {
  const long long ni = G.ni, nj = G.nj, nk = G.nk, g = G.nm1;
  if (t < 0) return false;
  if (G.jfix) {
    // collapsed: the stored row's theta and r ghosts (no phi ghosts are stored)
    const long long nkS1 = 2 * g * ni, niS1 = 2 * g * (nk - 2 * g);
    j = G.jfix;
    if (t < nkS1) {
      i = (int)(t % ni) + 1;
      const long long kk = t / ni;
      k = (int)(kk < g ? kk : nk - 2 * g + kk) + 1;
      return true;
    }
    t -= nkS1;
    if (t < niS1) {
      const long long ii = t % (2 * g);
      i = (int)(ii < g ? ii : ni - 2 * g + ii) + 1;
      k = (int)(g + t / (2 * g)) + 1;
      return true;
    }
    return false;
  }
  const long long nkS = 2 * g * ni * nj, njS = 2 * g * ni * (nk - 2 * g), niS = 2 * g * (nj - 2 * g) * (nk - 2 * g);
  if (t < nkS) {
    i = (int)(t % ni) + 1;
    const long long r = t / ni;
    j = (int)(r % nj) + 1;
    const long long kk = r / nj;
    k = (int)(kk < g ? kk : nk - 2 * g + kk) + 1;
    return true;
  }
  t -= nkS;
  if (t < njS) {
    i = (int)(t % ni) + 1;
    const long long r = t / ni;
    const long long jj = r % (2 * g);
    j = (int)(jj < g ? jj : nj - 2 * g + jj) + 1;
    k = (int)(g + r / (2 * g)) + 1;
    return true;
  }
  t -= njS;
  if (t < niS) {
    const long long ii = t % (2 * g);
    i = (int)(ii < g ? ii : ni - 2 * g + ii) + 1;
    const long long r = t / (2 * g);
    j = (int)(g + r % (nj - 2 * g)) + 1;
    k = (int)(g + r / (nj - 2 * g)) + 1;
    return true;
  }
  return false;
}

// Ghost fill of the stage: bnd4 composed per ghost cell, written into the
// ghost slots of the recovered primitives.  bnd_value reads only interior
// cells of that array (and the stage-start values of other ghosts), so the
// writes cannot race with the reads.
RH_DEV void b_bnd(const Geo& G, const PostArgs& A, long long blk) {
  int i, j, k;
  if (!ghost_index(blk * blockDim.x + threadIdx.x, G, i, j, k)) return;
  const long long N = G.N, c = cidx(i, j, k, G);
  const BndCtx B{&G, A.Wpre, A.Wold, A.Wo, A.bc};
#pragma unroll 1
  for (int n = 0; n < 9; n++) A.Wghost[(long long)n * N + c] = bnd_value(B, n, i, j, k);
}

// infchk + atmosphere1a + pminmax2b + caluu1 on every cell, after k_bnd.
// The metric of cell (i,k) through a pointer the compiler cannot see
// through, so its values are loaded again (from cache) in the next phase
// instead of being held in registers across phases.  k_post's three phases
// each read the 32 metric values; holding them costs two waves per SIMD.
RH_DEV Met reload_met(const double* metC, const Geo& G, int i, int k) {
  asm volatile("" : "+s"(metC));
  return Met{metC, G.P2, q2(i, k, G)};
}

RH_DEV void b_post(const Geo& G, const PostArgs& A, long long blk) {
  const long long t = blk * blockDim.x + threadIdx.x;
  int i, j, k;
  if (!box_index(t, 1, G.ni, jlo(G, 1), jhi(G, G.nj), 1, G.nk, i, j, k)) return;
  const long long N = G.N, c = cidx(i, j, k, G);
  const double gam = A.gam;
  const Met M{A.metC, G.P2, q2(i, k, G)};
  double w[9];
#pragma unroll
  for (int n = 0; n < 9; n++) w[n] = A.Wpre[(long long)n * N + c];
  // infchk, outer radial ghosts (bnd.f90:1740-1819); model 7..20
  if (A.model >= 7 && A.model <= 20 && i > G.ni - G.nm1) {
    const double gfl = calgfl(w[1], w[2], w[3], M);
    double vel[4];
    vel[1] = w[1] * (1.0 / gfl);
    vel[2] = w[2] * (1.0 / gfl);
    vel[3] = w[3] * (1.0 / gfl);
    if (vel[1] < 0.0) {
      const double alpha1 = 1.0 / sqrt(-M.gi(0, 0));
      const double beta1 = alpha1 * alpha1 * M.gi(0, 1);
      vel[1] = beta1 * (1.0 / alpha1);
      double vsq = 0.0;
#pragma unroll
      for (int m = 1; m <= 3; m++)
#pragma unroll
        for (int nn = 1; nn <= 3; nn++) vsq = vsq + M.g(m, nn) * vel[m] * vel[nn];
      if (vsq >= 1.0) vsq = 1.0 - 1.0 / (10.0 * 10.0);
      const double g2 = 1.0 / sqrt(1.0 - vsq);
      w[1] = vel[1] * g2; w[2] = vel[2] * g2; w[3] = vel[3] * g2;
    }
  }
  const double demin = A.dmin * A.r15[i - 1];
  const double prmin = (gam - 1.0) * A.pmin * A.r25[i - 1];
  // atmosphere1a (model 7..15): correction.f90:276-393
  bool atm = false;
  if (A.model >= 7 && A.model <= 15) {
    const double dthe = 1.0 * (1.0 + 1.0e-2) * demin;
    const double dthe1 = 1.0 * (1.0 + 1.0e-2) * prmin;
    const double bsq = w[6] * w[6] + w[7] * w[7] + w[8] * w[8];
    if (w[0] < dthe || (w[4] < dthe1 && bsq <= 1.0e-20)) {
      atm = true;
      double vsq = 0.0;
      const double vel[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int m = 1; m <= 3; m++)
#pragma unroll
        for (int nn = 1; nn <= 3; nn++) vsq = vsq + M.g(m, nn) * vel[m] * vel[nn];
      if (vsq >= 1.0) vsq = 1.0 - 1.0 / (10.0 * 10.0);
      const double g2 = 1.0 / sqrt(1.0 - vsq);
      w[0] = demin; w[4] = prmin;
      w[1] = vel[1] * g2; w[2] = vel[2] * g2; w[3] = vel[3] * g2;
    }
    w[5] = 0.0;
  }
  // a collapsed model's cell stands for its nj phi copies, each counted by the reference
  const unsigned long long rows = G.jfix ? (unsigned long long)G.nj : 1ULL;
  count_wave(A.cpart, blk, C_ATM, atm, rows);
  // pminmax2b: correction.f90:808-961
  bool scaled = false, hit = false;
  {
    const Met M = reload_met(A.metC, G, i, k);
    const double gflmax = 20.0;
    double g3[9], gi3[9];
    cal3met(M, g3, gi3);
    double gfl = calgfl(w[1], w[2], w[3], M);
    double vel[4];
    vel[1] = w[1] * (1.0 / gfl);
    vel[2] = w[2] * (1.0 / gfl);
    vel[3] = w[3] * (1.0 / gfl);
    double vsq = 0.0;
#pragma unroll
    for (int m = 1; m <= 3; m++)
#pragma unroll
      for (int nn = 1; nn <= 3; nn++) vsq = vsq + g3[(m - 1) + 3 * (nn - 1)] * vel[m] * vel[nn];
    if (vsq >= 1.0 || gfl > gflmax) vsq = 1.0 - 1.0 / (gflmax * gflmax);
    gfl = 1.0 / sqrt(1.0 - vsq);
    Prel s;
    const double wl[9] = {w[0], vel[1] * gfl, vel[2] * gfl, vel[3] * gfl, w[4], w[5], w[6], w[7], w[8]};
    prelude(wl, gam, M, s);  // gfl, u^mu, b^mu of the limited velocity
    gfl = s.gfl;
    const double bcon1[4] = {0.0, w[6], w[7], w[8]};
    double bcov1[4];
    lower(bcon1, bcov1, M);
    const double bsq = bcon1[1] * bcov1[1] + bcon1[2] * bcov1[2] + bcon1[3] * bcov1[3];
    double de = w[0], pr = w[4];
    int guard = 0;
    for (; guard < 100000; guard++) {
      const double roh = de + (gam / (gam - 1.0)) * pr;
      const double t1a = roh * gfl * gfl - pr;
      const double t1b = fabs((double)0.05f * bsq);
      if (t1a <= t1b) { de = (double)1.05f * de; pr = (double)1.05f * pr; scaled = true; continue; }
      break;
    }
    if (de != w[0]) w[0] = de;
    else if (pr != w[4]) w[4] = pr;
    if (pr < prmin) { w[4] = prmin; hit = true; }
    if (pr > A.pmax) { w[4] = A.pmax; hit = true; }
    if (de < demin) { w[0] = demin; hit = true; }
    const double prro = w[4] * (1.0 / w[0]);
    if (prro > 1.0e6) { w[0] = w[4] / 1.0e6; hit = true; }
    const double bbro = s.bbsq * (1.0 / w[0]);
    const double bbpr = s.bbsq * (1.0 / w[4]);
    if (bbro > 50.0) { w[0] = s.bbsq / 50.0; hit = true; }
    if (bbpr > 2.0e3) { w[4] = s.bbsq / 2.0e3; hit = true; }
  }
  count_wave(A.cpart, blk, C_PMMSCALE, scaled, rows);
  count_wave(A.cpart, blk, C_PMMFLOOR, hit, rows);
#pragma unroll
  for (int n = 0; n < 9; n++) A.Wout[(long long)n * N + c] = w[n];
  // caluu1 (caluu.f90:2-126): conserved variables of the final primitives
  {
    const Met M = reload_met(A.metC, G, i, k);
    Prel s2;
    prelude(w, gam, M, s2);
    double u[9];
    tmunu_row(s2, 0, M.detg(), pow(s2.de, gam), u);
#pragma unroll
    for (int n = 0; n < 9; n++) A.Uout[(long long)n * N + c] = u[n];
  }
}

// ------------------------------------------------------------------------
// CFL: cdtcfl4 (cdtcfl.f90:2-129) with calcha4 centre speeds.
// The reference folds dtcfl = MIN(dtcfl, x) over the cells in (k,j,i) order,
// and MIN(a,b) = a<b?a:b: a NaN element sets the running value to NaN and
// the next element replaces it.  So the result is the minimum over the
// elements after the last NaN (NaN if a NaN is the last element; 1.0 and all
// elements if there is none).  Each block reduces its contiguous range to
// (position of its last NaN, minimum after it) with wavefront shuffles; one
// block then combines the block results in order.
// Accumulators: 0 dtcfl (three elements per cell: dtx1a, dtx2a, dtx3a),
// 1..3 dtx1..dtx3 (one element per cell).
// ------------------------------------------------------------------------
RH_DEV double wave_min(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const double o = __shfl_xor(v, off, 64);
    v = (o < v) ? o : v;  // operands are never NaN here
  }
  return v;
}
RH_DEV long long wave_max_ll(long long v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const long long o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}

template <int ICHA>
RH_DEV void b_cfl(const Geo& G, const CflArgs& A, long long blk) {
  __shared__ long long s_last[4][4];
  __shared__ double s_min[4][4];
  const long long t = blk * blockDim.x + threadIdx.x;
  int i, j, k;
  double e[4] = {__builtin_inf(), __builtin_inf(), __builtin_inf(), 0.0};  // dtx1a, dtx2a, dtx3a
  bool valid = box_index(t, 2, G.ni - 1, jlo(G, 2), jhi(G, G.nj - 1), 2, G.nk - 1, i, j, k);
  if (valid) {
    const long long N = G.N, c = cidx(i, j, k, G);
    const Met M{A.metC, G.P2, q2(i, k, G)};
    double w[9];
#pragma unroll
    for (int n = 0; n < 9; n++) w[n] = A.W[(long long)n * N + c];
    Prel s;
    prelude(w, A.gam, M, s);
    double vp[4], vm[4];
    centre_speeds<ICHA>(s, A.gam, M, vp, vm);
    const int nmax = G.nm1 * 2 + (G.nm1 - 1);
    const int nd[4] = {0, G.ni, G.nj, G.nk};
    const double dxv[4] = {0.0, A.dx1b[i - 1], A.dx2b[j - 1], A.dx3b[k - 1]};
#pragma unroll
    for (int q = 1; q <= 3; q++) {
      double vf = fmaxf_(fmaxf_(0.0, fabs(vp[q])), fabs(vm[q]));
      if (nd[q] > nmax && vf > 0.0) {
        if (vf > 1.0) vf = 1.0;
        e[q - 1] = fabs(dxv[q] / (vf + 1.0e-12));
      } else {
        e[q - 1] = 1.0e10;
      }
    }
  }
  // per accumulator: this thread's last NaN position and minimum after it
  long long last[4];
  double mn[4];
  {
    long long l0 = -1;
    double m0 = __builtin_inf();
    if (valid)
#pragma unroll
      for (int q = 0; q < 3; q++) {
        if (isnan(e[q])) { l0 = 3 * t + q; m0 = __builtin_inf(); }
        else m0 = (e[q] < m0) ? e[q] : m0;
      }
    last[0] = l0; mn[0] = m0;
#pragma unroll
    for (int q = 0; q < 3; q++) {
      const bool nan = valid && isnan(e[q]);
      last[q + 1] = nan ? t : -1;
      mn[q + 1] = (valid && !nan) ? e[q] : __builtin_inf();
    }
  }
  const int lane = threadIdx.x % 64, wid = threadIdx.x / 64;
  long long L[4];
#pragma unroll
  for (int a = 0; a < 4; a++) {
    const long long v = wave_max_ll(last[a]);
    if (lane == 0) s_last[wid][a] = v;
  }
  __syncthreads();
#pragma unroll
  for (int a = 0; a < 4; a++) {
    long long v = s_last[0][a];
    for (int w = 1; w < (int)(blockDim.x / 64); w++) v = s_last[w][a] > v ? s_last[w][a] : v;
    L[a] = v;
  }
  // contribute the minimum over own elements after the block's last NaN
#pragma unroll
  for (int a = 0; a < 4; a++) {
    double m = __builtin_inf();
    if (valid) {
      if (a == 0) {
#pragma unroll
        for (int q = 0; q < 3; q++)
          if (3 * t + q > L[0] && !isnan(e[q])) m = (e[q] < m) ? e[q] : m;
      } else if (t > L[a] && !isnan(e[a - 1])) {
        m = e[a - 1];
      }
    }
    const double v = wave_min(m);
    if (lane == 0) s_min[wid][a] = v;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    const int a = threadIdx.x;
    double v = s_min[0][a];
    for (int w = 1; w < (int)(blockDim.x / 64); w++) v = (s_min[w][a] < v) ? s_min[w][a] : v;
    A.blast[4 * blk + a] = L[a];
    A.bmin[4 * blk + a] = v;
  }
}

// Combine the per-block results in block order (one workgroup): the last NaN
// lies in the last block that has one; that block's minimum after it and the
// full minima of every later block make the result.
RH_DEV void b_cfl_finish(const CflArgs& A) {
  // the four accumulators together: two passes over the per-block results
  // (last NaN, then the minimum after it); max and min do not depend on the
  // order, so the result is the sequential fold's
  const int nblocks = A.nblocks;
  __shared__ long long s_l[16][4];
  __shared__ double s_m[16][4];
  __shared__ long long s_B[4];
  const int lane = threadIdx.x % 64, wid = threadIdx.x / 64, nw = blockDim.x / 64;
  long long bg[4] = {-1, -1, -1, -1};
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
#pragma unroll
    for (int a = 0; a < 4; a++)
      if (A.blast[4 * b + a] >= 0) bg[a] = b > bg[a] ? b : bg[a];
#pragma unroll
  for (int a = 0; a < 4; a++) {
    const long long v = wave_max_ll(bg[a]);
    if (lane == 0) s_l[wid][a] = v;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    long long B = s_l[0][threadIdx.x];
    for (int w = 1; w < nw; w++) B = s_l[w][threadIdx.x] > B ? s_l[w][threadIdx.x] : B;
    s_B[threadIdx.x] = B;
  }
  __syncthreads();
  long long B[4];
  double m[4];
#pragma unroll
  for (int a = 0; a < 4; a++) { B[a] = s_B[a]; m[a] = __builtin_inf(); }
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
#pragma unroll
    for (int a = 0; a < 4; a++)
      if (b >= B[a]) { const double v = A.bmin[4 * b + a]; m[a] = (v < m[a]) ? v : m[a]; }
#pragma unroll
  for (int a = 0; a < 4; a++) {
    const double v = wave_min(m[a]);
    if (lane == 0) s_m[wid][a] = v;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    const int a = threadIdx.x;
    double M = s_m[0][a];
    for (int w = 1; w < nw; w++) M = (s_m[w][a] < M) ? s_m[w][a] : M;
    double r;
    if (B[a] < 0) r = fminf_(1.0, M);                              // no NaN: MIN over 1.0 and all
    else r = (M == __builtin_inf()) ? __builtin_nan("") : M;  // a trailing NaN stays
    A.dtbits[a] = (unsigned long long)__double_as_longlong(r);
  }
}

// time <- time + cfl*dtcfl ; last dt kept (main.f90:760-781)
RH_DEV void b_advance_time(const AdvArgs& A) {
  if (threadIdx.x == 0) {
    const double dt = A.cfl * __longlong_as_double((long long)A.dtbits[0]);
    A.tdt[0] = A.tdt[0] + dt;
    A.tdt[1] = dt;
  }
}

// Resolution of cells whose initial W is inherited from the previous cell in
// the reference's (k,j,i) loop order (recov.f90:815-817 leaves wg unchanged,
// recov.f90:632 starts it at 0).  A flagged cell is computed once its
// predecessor's final W is known.
//
// Chain heads (flagged cells whose predecessor is not flagged, or that open
// the loop) are independent: every workgroup of the launch takes a share of
// them.  The last workgroup to finish (a device-scope counter after a
// device-scope fence) then resolves the remaining links in passes over the
// flagged list, each pass taking the cells whose predecessor is resolved,
// until none is left: a chain of consecutive flagged cells takes one pass
// per link.  A batch runs one workgroup per member, which does both parts.

// Predecessor of interior cell c in (k,j,i) order, -1 for the first cell.
RH_DEV long long carry_pred(const Geo& G, long long c, int& i, int& k) {
  const long long sk = kstride(G);
  i = (int)(c % G.ni) + 1;
  k = (int)(c / sk) + 1;
  const int j = G.jfix ? G.jfix : (int)((c / G.ni) % G.nj) + 1;
  int pi = i - 1, pj = j, pk = k;
  if (pi < G.nm1 + 1) { pi = G.ni - G.nm1; pj = j - 1; if (pj < G.nm1 + 1) { pj = G.nj - G.nm1; pk = k - 1; } }
  return pk >= G.nm1 + 1 ? cidx(pi, pj, pk, G) : -1;
}

template <int IW>
RH_DEV void carry_cell(const Geo& G, const FixArgs& A, long long c, int i, int k, double wg_prev) {
  const long long N = G.N;
  const Met M{A.metC, G.P2, q2(i, k, G)};
  double u[9], w[9];
  for (int m = 0; m < 9; m++) { u[m] = A.Ustage[(long long)m * N + c]; w[m] = A.W[(long long)m * N + c]; }
  RecovOut ro;
  bool carry;
  double wgout;
  recov_any<IW>(u, w, M, A.gam, A.dmin, A.pmin, A.iter, A.r15[i - 1], A.r25[i - 1], wg_prev, carry, ro, wgout);
  for (int m = 0; m < 9; m++) A.Wpre[(long long)m * N + c] = ro.w[m];
  A.wgout[c] = wgout;
  if (ro.fallback) atomicAdd(&A.cnt[C_FALLBACK], 1ULL);
  if (ro.fail) atomicAdd(&A.cnt[C_FAIL], 1ULL);
  if (ro.negrho) atomicAdd(&A.cnt[C_NEGRHO], 1ULL);
  if (ro.negp) atomicAdd(&A.cnt[C_NEGP], 1ULL);
  if (ro.vsqgt1) atomicAdd(&A.cnt[C_VSQGT1], (unsigned long long)ro.vsqgt1);
  if (ro.det0) atomicAdd(&A.cnt[C_DET0], (unsigned long long)ro.det0);
  if (ro.guard) atomicAdd(&A.cnt[C_GUARD], 1ULL);
}

// grp/ngrp: this workgroup's place among those sharing the chain heads
// (ngrp == 1: the workgroup does everything and needs no counter).
// Pass -1 takes the chain heads (pend 2 = resolved, 1 = waits for its
// predecessor); passes 0, 1, ... (last workgroup) take the links.  One loop
// body, so the recovery is compiled once.
template <int IW>
RH_DEV void b_recov_carry(const Geo& G, const FixArgs& A, int grp, int ngrp) {
  __shared__ int s_progress, s_left, s_last;
  unsigned long long n = A.carry[0];
  if (n == 0) return;
  if (n > (unsigned long long)A.carry_cap) n = (unsigned long long)A.carry_cap;
  const unsigned long long* list = A.carry + 1;
  unsigned char* pend = A.pend;
  for (int pass = -1; pass < (1 << 22); pass++) {
    const bool heads = pass < 0;
    if (!heads) {
      if (threadIdx.x == 0) { s_progress = 0; s_left = 0; }
      __syncthreads();
    }
    const unsigned long long e0 = heads ? (unsigned long long)grp * blockDim.x + threadIdx.x : threadIdx.x;
    const unsigned long long es = heads ? (unsigned long long)ngrp * blockDim.x : blockDim.x;
    for (unsigned long long e = e0; e < n; e += es) {
      if (!heads && pend[e] != 1) continue;
      const long long c = (long long)list[e];
      int i, k;
      const long long pc = carry_pred(G, c, i, k);  // -1 only for the first cell, a head
      if (pc >= 0 && A.cflag[pc]) {  // predecessor not resolved yet
        if (heads) pend[e] = 1;
        else s_left = 1;
        continue;
      }
      carry_cell<IW>(G, A, c, i, k, pc >= 0 ? A.wgout[pc] : 0.0);
      pend[e] = 2;
      if (!heads) s_progress = 1;
    }
    if (heads && ngrp > 1) {
      unsigned long long* done = A.carry + 1 + A.carry_cap;
      __threadfence();  // each thread's results visible on the device before the count
      __syncthreads();
      if (threadIdx.x == 0) s_last = atomicAdd(done, 1ULL) == (unsigned long long)(ngrp - 1);
      __syncthreads();
      if (!s_last) return;
      __threadfence();  // every workgroup's results before the links are read
    }
    __syncthreads();
    if (heads && threadIdx.x == 0 && A.carry[0] > (unsigned long long)A.carry_cap) atomicAdd(&A.cnt[C_GUARD], 1ULL);
    for (unsigned long long e = threadIdx.x; e < n; e += blockDim.x)
      if (pend[e] == 2) { pend[e] = 0; A.cflag[list[e]] = 0; }
    __syncthreads();
    if (heads) continue;
    const bool stop = !s_left || !s_progress;  // no progress: nothing can resolve further
    __syncthreads();                            // all have read before the next reset
    if (stop) break;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    A.carry[0] = 0;
    if (ngrp > 1) A.carry[1 + A.carry_cap] = 0;
  }
}

// Set-up tables of an axisymmetric metric (metric 303 or 1) on the device,
// one thread per (i,k): the metric at centres, i-, j- and k-faces (j-face =
// centre, coord.f90:116,121) with sqrt(-g), and the Christoffel symbols.
// tr holds per k the host's sin/cos of the folded theta (clamped): at
// x3(k), x3(k)+1e-5, x3(k)-1e-5 and x3a(k); everything else is arithmetic,
// so the tables equal the host set-up's bit for bit.  *sym is cleared where
// Gamma^l_{ab} != Gamma^l_{ba} (the caller sets it to 1 first).
__global__ void __launch_bounds__(256) k_metric_tables(int ni, int nk, int metric, double akm,
                                                       const double* __restrict__ x1, const double* __restrict__ x1a,
                                                       const double* __restrict__ x3, double ph,
                                                       const double* __restrict__ tr, double* met0, double* met1,
                                                       double* met2, double* met3, double* chr, int* sym) {
  const long long P2 = (long long)ni * nk;
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= P2) return;
  const int i = (int)(q % ni), k = (int)(q / ni);
  const double* T = tr + 8LL * k;
  const double r = x1[i];
  double g[16], h[16];
  auto put = [&](double* met) {
    for (int e = 0; e < 16; e++) { met[e * P2 + q] = g[e]; met[(16 + e) * P2 + q] = h[e]; }
    met[32 * P2 + q] = rhmet::detg(g);
  };
  rhmet::cov(metric, akm, g, r, T[0], T[1]);
  rhmet::con(metric, akm, h, r, T[0], T[1]);
  put(met0);
  put(met2);
  rhmet::cov(metric, akm, g, x1a[i], T[0], T[1]);
  rhmet::con(metric, akm, h, x1a[i], T[0], T[1]);
  put(met1);
  rhmet::cov(metric, akm, g, r, T[6], T[7]);
  rhmet::con(metric, akm, h, r, T[6], T[7]);
  put(met3);
  const rhmet::ThetaTrig tt{T[0], T[1], T[2], T[3], T[4], T[5]};
  double ch[64];
  rhmet::christoffel(metric, akm, r, x3[k], ph, tt, ch);
  bool s = true;
  for (int e = 0; e < 64; e++) chr[e * P2 + q] = ch[e];
  for (int l = 0; l < 4; l++)
    for (int a = 0; a < 4; a++)
      for (int b = a + 1; b < 4; b++)
        s = s && __double_as_longlong(ch[l + 4 * a + 16 * b]) == __double_as_longlong(ch[l + 4 * b + 16 * a]);
  if (!s) atomicAnd(sym, 0);
}

hipError_t launch_metric_tables(int ni, int nk, int metric, double akm, const double* x1, const double* x1a,
                                const double* x3, double ph, const double* tr, double* const met[4], double* chr,
                                int* sym, hipStream_t st) {
  const long long P2 = (long long)ni * nk;
  if (P2 <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_metric_tables, dim3((unsigned)((P2 + 255) / 256)), dim3(256), 0, st, ni, nk, metric, akm, x1,
                     x1a, x3, ph, tr, met[0], met[1], met[2], met[3], chr, sym);
  return hipGetLastError();
}

// caluu1 (caluu.f90:2-126) on every cell: conserved variables from primitives.
__global__ void __launch_bounds__(256) k_caluu1(Geo G, const double* __restrict__ W, const double* __restrict__ metC,
                                                double gam, double* __restrict__ U) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int i, j, k;
  if (!box_index(t, 1, G.ni, jlo(G, 1), jhi(G, G.nj), 1, G.nk, i, j, k)) return;
  const long long N = G.N, c = cidx(i, j, k, G);
  const Met M{metC, G.P2, q2(i, k, G)};
  double w[9], u[9];
#pragma unroll
  for (int n = 0; n < 9; n++) w[n] = W[(long long)n * N + c];
  Prel s;
  prelude(w, gam, M, s);
  tmunu_row(s, 0, M.detg(), pow(s.de, gam), u);
#pragma unroll
  for (int n = 0; n < 9; n++) U[(long long)n * N + c] = u[n];
}

// ------------------------------------------------------------------------
// Layout conversion between the reference's (n,i,j,k) arrays and planes.
// ------------------------------------------------------------------------
__global__ void k_aos_to_soa(const double* __restrict__ aos, double* __restrict__ soa, long long N) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N * 9) return;
  const long long c = t / 9, n = t % 9;
  soa[n * N + c] = aos[t];
}
__global__ void k_soa_to_aos(const double* __restrict__ soa, double* __restrict__ aos, long long N) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N * 9) return;
  const long long c = t / 9, n = t % 9;
  aos[t] = soa[n * N + c];
}

// ww(3,9,i,j,k) of the reference (d fastest, then n) <-> planes d*9+n.
__global__ void k_ww_to_planes(const double* __restrict__ ww, double* __restrict__ F, long long N) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
#pragma unroll
  for (int d = 0; d < 3; d++)
#pragma unroll
    for (int n = 0; n < 9; n++) F[(long long)(d * 9 + n) * N + t] = ww[27 * t + d + 3 * n];
}
// the stored flux planes between the two precision forms
template <class S, class T>
__global__ void k_retype(const S* __restrict__ src, T* __restrict__ dst, long long n) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[t] = T(src[t]);
}
__global__ void k_planes_to_ww(const double* __restrict__ F, double* __restrict__ ww, long long N) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
#pragma unroll
  for (int d = 0; d < 3; d++)
#pragma unroll
    for (int n = 0; n < 9; n++) ww[27 * t + d + 3 * n] = F[(long long)(d * 9 + n) * N + t];
}

// ------------------------------------------------------------------------
// Kernels: one model per launch (the workgroups renumbered per XCD where the
// stencil benefits), and the batched twins, where one launch covers the same
// step of many models.  A batched launch reads a table of BatchItem (the
// model's Geo, its arguments, and the first workgroup of its part of the
// grid); each workgroup finds its model by binary search and runs the same
// body with its index inside that model's part.
// ------------------------------------------------------------------------
RH_DEV int batch_find(const BatchItem* __restrict__ it, int n, long long b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (it[mid].off <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

RH_DEV void b_copy(const CopyArgs& A, long long blk) {
  const long long t = blk * blockDim.x + threadIdx.x;
  if (t < A.n) A.dst[t] = A.src[t];
}

template <int D, class R>
__global__ void __launch_bounds__(256) k_ppm_edges(Geo G, EdgeArgs A) { b_ppm_edges<D, R>(G, A, xcd_block()); }
template <int D, class R>
__global__ void __launch_bounds__(256) k_mc_edges(Geo G, EdgeArgs A) { b_mc_edges<D, R>(G, A, xcd_block()); }
// the sweeps over precomputed face values (PPM, MC) fit 3 waves per SIMD in
// FP64 (168 VGPRs, against 170 and 2 waves unconstrained) and 5 in float (96);
// MP5 and WENO5 are left to the compiler
template <class R, int IREC>
constexpr int sweep_waves = (IREC == 5 || IREC == 1) ? (sizeof(R) == 4 ? 5 : 3) : 1;
template <int D, int IREC, class R, int ICHA>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sweep_waves<R, IREC>)))
k_sweep(Geo G, SweepArgs A) { b_sweep<D, IREC, R, ICHA>(G, A, xcd_block()); }
template <int D, class R, int ICHA>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sweep_waves<R, 5>)))
k_psweep(Geo G, PSweepArgs A) { b_psweep<D, R, ICHA>(G, A, xcd_block()); }
template <class FT>
__global__ void __launch_bounds__(256) k_fluxct(Geo G, CtArgs A) { b_fluxct<FT>(G, A, xcd_block()); }
template <int STAGE, bool SYM, class FT>
__global__ void __launch_bounds__(256) k_rk(Geo G, StageArgs A) { b_rk<STAGE, SYM, FT>(G, A, xcd_block()); }
template <int IW>
__global__ void __launch_bounds__(256) k_recov(Geo G, StageArgs A) { b_recov<IW>(G, A, xcd_block()); }
template <int IW>
__global__ void __launch_bounds__(1024) k_recov_carry(Geo G, FixArgs A) {
  b_recov_carry<IW>(G, A, blockIdx.x, gridDim.x);
}
__global__ void __launch_bounds__(256) k_bnd(Geo G, PostArgs A) { b_bnd(G, A, blockIdx.x); }
__global__ void __launch_bounds__(256) k_post(Geo G, PostArgs A) { b_post(G, A, xcd_block()); }
template <int ICHA>
__global__ void __launch_bounds__(256) k_cfl(Geo G, CflArgs A) { b_cfl<ICHA>(G, A, blockIdx.x); }
__global__ void __launch_bounds__(1024) k_cfl_finish(CflArgs A) { b_cfl_finish(A); }
__global__ void __launch_bounds__(64) k_advance_time(AdvArgs A) { b_advance_time(A); }

// batched twins: BODY(G, A, blk) for the model that owns workgroup blockIdx.x
#define RH_BATCH_PROLOGUE                                  \
  const int q = batch_find(it, n, (long long)blockIdx.x);  \
  const Geo G = it[q].G;                                   \
  const long long blk = (long long)blockIdx.x - it[q].off;
template <int D, class R>
__global__ void __launch_bounds__(256) kb_ppm_edges(const BatchItem* __restrict__ it, int n) {
  RH_BATCH_PROLOGUE
  const EdgeArgs A = it[q].A.e;
  b_ppm_edges<D, R>(G, A, blk);
}
template <int D, class R>
__global__ void __launch_bounds__(256) kb_mc_edges(const BatchItem* __restrict__ it, int n) {
  RH_BATCH_PROLOGUE
  const EdgeArgs A = it[q].A.e;
  b_mc_edges<D, R>(G, A, blk);
}
template <int D, int IREC, class R, int ICHA>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sweep_waves<R, IREC>)))
kb_sweep(const BatchItem* __restrict__ it, int n) {
  RH_BATCH_PROLOGUE
  const SweepArgs A = it[q].A.s;
  b_sweep<D, IREC, R, ICHA>(G, A, blk);
}
template <int D, class R, int ICHA>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sweep_waves<R, 5>)))
kb_psweep(const BatchItem* __restrict__ it, int n) {
  RH_BATCH_PROLOGUE
  const PSweepArgs A = it[q].A.ps;
  b_psweep<D, R, ICHA>(G, A, blk);
}
template <class FT>
__global__ void __launch_bounds__(256) kb_fluxct(const BatchItem* __restrict__ it, int n) {
  RH_BATCH_PROLOGUE
  const CtArgs A = it[q].A.c;
  b_fluxct<FT>(G, A, blk);
}
template <int STAGE, bool SYM, class FT>
__global__ void __launch_bounds__(256) kb_rk(const BatchItem* __restrict__ it, int n) {
  RH_BATCH_PROLOGUE
  const StageArgs A = it[q].A.st;
  b_rk<STAGE, SYM, FT>(G, A, blk);
}
template <int IW>
__global__ void __launch_bounds__(256) kb_recov(const BatchItem* __restrict__ it, int n) {
  RH_BATCH_PROLOGUE
  const StageArgs A = it[q].A.st;
  b_recov<IW>(G, A, blk);
}
__global__ void __launch_bounds__(256) kb_bnd(const BatchItem* __restrict__ it, int n) {
  RH_BATCH_PROLOGUE
  const PostArgs A = it[q].A.p;
  b_bnd(G, A, blk);
}
__global__ void __launch_bounds__(256) kb_post(const BatchItem* __restrict__ it, int n) {
  RH_BATCH_PROLOGUE
  const PostArgs A = it[q].A.p;
  b_post(G, A, blk);
}
template <int ICHA>
__global__ void __launch_bounds__(256) kb_cfl(const BatchItem* __restrict__ it, int n) {
  RH_BATCH_PROLOGUE
  const CflArgs A = it[q].A.cf;
  b_cfl<ICHA>(G, A, blk);
}
__global__ void __launch_bounds__(256) kb_copy(const BatchItem* __restrict__ it, int n) {
  RH_BATCH_PROLOGUE
  (void)G;
  const CopyArgs A = it[q].A.cp;
  b_copy(A, blk);
}
#undef RH_BATCH_PROLOGUE
// one workgroup per model
template <int IW>
__global__ void __launch_bounds__(1024) kb_recov_carry(const BatchItem* __restrict__ it, int n) {
  if ((int)blockIdx.x >= n) return;
  const Geo G = it[blockIdx.x].G;
  const FixArgs A = it[blockIdx.x].A.fx;
  b_recov_carry<IW>(G, A, 0, 1);
}
__global__ void __launch_bounds__(1024) kb_cfl_finish(const BatchItem* __restrict__ it, int n) {
  if ((int)blockIdx.x >= n) return;
  const CflArgs A = it[blockIdx.x].A.cf;
  b_cfl_finish(A);
}
__global__ void __launch_bounds__(64) kb_advance_time(const BatchItem* __restrict__ it, int n) {
  if ((int)blockIdx.x >= n) return;
  const AdvArgs A = it[blockIdx.x].A.a;
  b_advance_time(A);
}

// ------------------------------------------------------------------------
// Launch helpers (host side).  While a recorder is set, each helper appends
// the launch it would make (kind, template variant, workgroups, arguments)
// instead of launching; the batch executor merges the records of many models.
// ------------------------------------------------------------------------
static inline unsigned nblk(long long n, int b) { return (unsigned)((n + b - 1) / b); }
static inline long long box(int lo, int hi) { return hi >= lo ? (long long)(hi - lo + 1) : 0; }
// rows of a box along j: one for a collapsed model (its stored row lies in every box)
static inline long long jbox(const Geo& G, int lo, int hi) { return G.jfix ? 1 : box(lo, hi); }

static thread_local std::vector<LaunchRec>* g_rec = nullptr;

void set_recorder(std::vector<LaunchRec>* r) { g_rec = r; }

static void record(int kind, int variant, int threads, long long nb, const Geo& G, const AnyArgs& A) {
  LaunchRec r;
  r.kind = kind; r.variant = variant; r.threads = threads; r.nb = nb; r.G = G; r.A = A;
  g_rec->push_back(r);
}

int block_threads(int kind) {
  switch (kind) {
    case LK_CARRY: case LK_CFL_FINISH: return 1024;
    case LK_ADVANCE: return 64;
    default: return 256;
  }
}

bool one_group_per_model(int kind) { return kind == LK_CARRY || kind == LK_CFL_FINISH || kind == LK_ADVANCE; }

hipError_t launch_edges(const Geo& G, int dir, int irec, const double* W, const double* dxs, const double* xs,
                        const double* tab, double gam, void* E, bool mixed, hipStream_t st) {
  const long long ci = box(G.nm0, G.ni - G.nm0 + (dir == 1));
  // PPM along j and k: whole 64-lane rows (box_index_wave)
  const long long ciw = (irec == 5 && dir != 1) ? 64 * ((ci + 63) / 64) : ci;
  const long long n = ciw * jbox(G, G.nm0, G.nj - G.nm0 + (dir == 2)) * box(G.nm0, G.nk - G.nm0 + (dir == 3));
  EdgeArgs A{W, dxs, xs, tab, gam, E};
  const int bt = slot_threads(G, TS_EDGES);
  const unsigned nb = nblk(n, bt);
  const bool mc = irec == 1;
  if (g_rec) {
    AnyArgs a{};
    a.e = A;
    record(LK_EDGES, dir + 4 * (int)mixed + 16 * (int)mc, bt, nb, G, a);
    return hipSuccess;
  }
  if (n == 0) return hipSuccess;
#define L(K, Dd, T) hipLaunchKernelGGL((K<Dd, T>), dim3(nb), dim3(bt), 0, st, G, A)
#define LD(K, T) do { if (dir == 1) L(K, 1, T); else if (dir == 2) L(K, 2, T); else L(K, 3, T); } while (0)
  if (mc) { if (mixed) LD(k_mc_edges, float); else LD(k_mc_edges, double); }
  else { if (mixed) LD(k_ppm_edges, float); else LD(k_ppm_edges, double); }
#undef LD
#undef L
  return hipGetLastError();
}

// Sweep kernels are instantiated per (direction, reconstruction, arithmetic
// type, wave-speed formula); the variant number packs the four so that batch
// plans group launches of the same kernel:
//   dir + 4*mixed + 8*irec + 64*icha.
template <int V> using IC = std::integral_constant<int, V>;
template <class T> struct TypeTag { using type = T; };

static int sweep_variant(int dir, bool mixed, int irec, int icha) { return dir + 4 * (int)mixed + 8 * irec + 64 * icha; }

template <class F>
static void with_icha(int icha, F&& f) {
  if (icha == 2) f(IC<2>{});
  else if (icha == 1) f(IC<1>{});
  else f(IC<0>{});
}

template <class F>
static void with_sweep(int variant, F&& f) {
  const int d = variant & 3, irec = (variant >> 3) & 7;
  const bool mx = (variant & 4) != 0;
  with_icha(variant >> 6, [&](auto C) {
    auto t = [&](auto D, auto I) {
      if (mx) f(D, I, TypeTag<float>{}, C);
      else f(D, I, TypeTag<double>{}, C);
    };
    auto r = [&](auto I) {
      if (d == 1) t(IC<1>{}, I);
      else if (d == 2) t(IC<2>{}, I);
      else t(IC<3>{}, I);
    };
    if (irec == 5) r(IC<5>{});
    else if (irec == 1) r(IC<1>{});
    else if (irec == 6) r(IC<6>{});
    else r(IC<7>{});
  });
}

hipError_t launch_sweep(const Geo& G, int dir, int irec, const double* W, const double* metF, const void* E,
                        double gam, void* F, void* CTsrc, bool mixed, hipStream_t st) {
  const long long n = box(G.nm0, G.ni - G.nm0) * jbox(G, G.nm0, G.nj - G.nm0) * box(G.nm0, G.nk - G.nm0);
  SweepArgs A{W, metF, E, gam, F, CTsrc};
  const int bt = slot_threads(G, TS_SWEEP1 + dir - 1);
  const unsigned nb = nblk(n, bt);
  if (g_rec) {
    AnyArgs a{};
    a.s = A;
    record(LK_SWEEP, sweep_variant(dir, mixed, irec, G.icha), bt, nb, G, a);
    return hipSuccess;
  }
  if (n == 0) return hipSuccess;
  with_sweep(sweep_variant(dir, mixed, irec, G.icha), [&](auto D, auto I, auto T, auto C) {
    hipLaunchKernelGGL((k_sweep<D.value, I.value, typename decltype(T)::type, C.value>), dim3(nb), dim3(bt), 0, st,
                       G, A);
  });
  return hipGetLastError();
}

// Fused PPM sweeps: variant = dir + 4*mixed + 64*icha (as the sweeps').
static int psweep_variant(int dir, bool mixed, int icha) { return dir + 4 * (int)mixed + 64 * icha; }

template <class F>
static void with_psweep(int variant, F&& f) {
  const int d = variant & 3;
  const bool mx = (variant & 4) != 0;
  with_icha(variant >> 6, [&](auto C) {
    auto t = [&](auto D) {
      if (mx) f(D, TypeTag<float>{}, C);
      else f(D, TypeTag<double>{}, C);
    };
    if (d == 1) t(IC<1>{});
    else if (d == 2) t(IC<2>{});
    else t(IC<3>{});
  });
}

// Waves of one fused PPM sweep (per model): D = 1, 64-cell chunks of the i
// rows (63 interfaces each) over (j, k); D = 2, 3, tiles of 64 / seg i by
// seg cells along D (seg - 1 interfaces) over the transverse axis.
static long long psweep_waves(const Geo& G, int D, int seg) {
  const long long nj = G.jfix ? 1 : (long long)(G.nj - 2 * G.nm0 + 1);
  const long long ni = G.ni - 2 * G.nm0 + 1, nk = G.nk - 2 * G.nm0 + 1;
  if (ni <= 0 || nj <= 0 || nk <= 0) return 0;
  if (D == 1) return ((ni + 62) / 63) * nj * nk;
  if (D == 2 && G.jfix) return ((ni + 63) / 64) * nk;  // one stored row: 64 i per wave
  const long long along = D == 2 ? nj : nk, across = D == 2 ? nk : nj, ti = 64 / seg;
  return ((along + seg - 2) / (seg - 1)) * ((ni + ti - 1) / ti) * across;
}

// The tile height in use along dir: the chosen one (or the default), but no
// taller than the axis needs (the smallest power of two above its interface
// count), so short axes (Riemann tubes' transverse axes) waste no lanes.
int psweep_segment(const Geo& G, int dir) {
  const int want = (dir == 2 ? G.seg[0] : dir == 3 ? G.seg[1] : 0);
  int s = want > 0 ? want : PSWEEP_SEG_DEFAULT;
  const int n = dir == 2 ? (G.jfix ? 1 : G.nj - 2 * G.nm0 + 1) : dir == 3 ? G.nk - 2 * G.nm0 + 1 : 0;
  int cap = 2;
  while (cap < 64 && cap - 1 < n) cap *= 2;
  return s < cap ? s : cap;
}

size_t psweep_lds_bytes(int threads, bool mixed) { return (size_t)18 * threads * (mixed ? sizeof(float) : sizeof(double)); }

hipError_t launch_psweep(const Geo& G, int dir, const double* W, const double* metF, const double* dxs,
                         const double* xs, const double* pc, double gam, void* F, void* CTsrc, bool mixed,
                         hipStream_t st) {
  const int seg = psweep_segment(G, dir);
  const long long waves = psweep_waves(G, dir, seg);
  PSweepArgs A{W, metF, dxs, xs, pc, gam, F, CTsrc, seg};
  const int bt = slot_threads(G, TS_SWEEP1 + dir - 1);
  const unsigned nb = nblk(waves * 64, bt);
  const int variant = psweep_variant(dir, mixed, G.icha);
  if (g_rec) {
    AnyArgs a{};
    a.ps = A;
    record(LK_PSWEEP, variant, bt, nb, G, a);
    return hipSuccess;
  }
  if (waves == 0) return hipSuccess;
  const size_t lds = psweep_lds_bytes(bt, mixed);
  with_psweep(variant, [&](auto D, auto T, auto C) {
    hipLaunchKernelGGL((k_psweep<D.value, typename decltype(T)::type, C.value>), dim3(nb), dim3(bt), lds, st, G, A);
  });
  return hipGetLastError();
}

hipError_t launch_fluxct(const Geo& G, const void* CTsrc, void* F1, void* F2, void* F3, bool f32, double* CTd,
                         hipStream_t st) {
  const long long n = box(G.nm1, G.ni - G.nm1) * jbox(G, G.nm1, G.nj - G.nm1) * box(G.nm1, G.nk - G.nm1);
  CtArgs A{CTsrc, F1, F2, F3, f32 ? CTd : nullptr};
  const int bt = slot_threads(G, TS_FLUXCT);
  if (g_rec) { AnyArgs a{}; a.c = A; record(LK_FLUXCT, (int)f32, bt, nblk(n, bt), G, a); return hipSuccess; }
  if (n == 0) return hipSuccess;
  if (f32) hipLaunchKernelGGL(k_fluxct<float>, dim3(nblk(n, bt)), dim3(bt), 0, st, G, A);
  else hipLaunchKernelGGL(k_fluxct<double>, dim3(nblk(n, bt)), dim3(bt), 0, st, G, A);
  return hipGetLastError();
}

hipError_t launch_update(const Geo& G, int stage, const StageArgs& A, hipStream_t st) {
  const long long n = box(G.nm1 + 1, G.ni - G.nm1) * jbox(G, G.nm1 + 1, G.nj - G.nm1) * box(G.nm1 + 1, G.nk - G.nm1);
  const int bt = slot_threads(G, TS_RK);
  const unsigned nb = nblk(n, bt);
  if (g_rec) {
    AnyArgs a{};
    a.st = A;
    record(LK_RK, stage + 8 * (int)A.chr_sym + 16 * (int)A.f32flux, bt, nb, G, a);
    return hipSuccess;
  }
  if (n == 0) return hipSuccess;
#define LT(S, Y) \
  do { if (A.f32flux) hipLaunchKernelGGL((k_rk<S, Y, float>), dim3(nb), dim3(bt), 0, st, G, A); \
       else hipLaunchKernelGGL((k_rk<S, Y, double>), dim3(nb), dim3(bt), 0, st, G, A); } while (0)
#define L(S) do { if (A.chr_sym) LT(S, true); else LT(S, false); } while (0)
  if (stage == 1) L(1);
  else if (stage == 2) L(2);
  else if (stage == 3) L(3);
  else L(4);
#undef L
#undef LT
  return hipGetLastError();
}

hipError_t launch_recov(const Geo& G, const StageArgs& A, hipStream_t st) {
  const long long n = box(G.nm1 + 1, G.ni - G.nm1) * jbox(G, G.nm1 + 1, G.nj - G.nm1) * box(G.nm1 + 1, G.nk - G.nm1);
  const int bt = slot_threads(G, TS_RECOV);
  if (g_rec) { AnyArgs a{}; a.st = A; record(LK_RECOV, G.iwvec, bt, nblk(n, bt), G, a); return hipSuccess; }
  if (n == 0) return hipSuccess;
  if (G.iwvec == 2) hipLaunchKernelGGL(k_recov<2>, dim3(nblk(n, bt)), dim3(bt), 0, st, G, A);
  else hipLaunchKernelGGL(k_recov<1>, dim3(nblk(n, bt)), dim3(bt), 0, st, G, A);
  return hipGetLastError();
}

hipError_t launch_recov_carry(const Geo& G, const FixArgs& A, hipStream_t st) {
  if (g_rec) { AnyArgs a{}; a.fx = A; record(LK_CARRY, G.iwvec, 1024, 1, G, a); return hipSuccess; }
  // workgroups sharing the chain heads: one per 1024 flaggable cells, at most 64
  const int ng = (int)std::min<long long>(64, std::max<long long>(1, (A.carry_cap + 1023) / 1024));
  if (G.iwvec == 2) hipLaunchKernelGGL(k_recov_carry<2>, dim3(ng), dim3(1024), 0, st, G, A);
  else hipLaunchKernelGGL(k_recov_carry<1>, dim3(ng), dim3(1024), 0, st, G, A);
  return hipGetLastError();
}

hipError_t launch_post(const Geo& G, const PostArgs& A, hipStream_t st) {
  const long long g = G.nm1;
  const long long ng = G.N - (long long)(G.ni - 2 * g) * (G.jfix ? 1 : G.nj - 2 * g) * (G.nk - 2 * g);
  const int bt = slot_threads(G, TS_POST);
  if (g_rec) {
    AnyArgs a{};
    a.p = A;
    record(LK_BND, 0, bt, ng > 0 ? nblk(ng, bt) : 0, G, a);
    record(LK_POST, 0, bt, nblk(G.N, bt), G, a);
    return hipSuccess;
  }
  if (ng > 0) hipLaunchKernelGGL(k_bnd, dim3(nblk(ng, bt)), dim3(bt), 0, st, G, A);
  hipLaunchKernelGGL(k_post, dim3(nblk(G.N, bt)), dim3(bt), 0, st, G, A);
  return hipGetLastError();
}

hipError_t launch_sum_counters(unsigned long long* cpart, long long rows, unsigned long long* cnt, hipStream_t st) {
  hipLaunchKernelGGL(k_sum_counters, dim3(1), dim3(1024), 0, st, cpart, rows, cnt);
  return hipGetLastError();
}

long long cfl_blocks(const Geo& G) {
  const long long n = box(2, G.ni - 1) * jbox(G, 2, G.nj - 1) * box(2, G.nk - 1);
  return n > 0 ? (long long)nblk(n, 256) : 1;
}

hipError_t launch_cfl(const Geo& G, const CflArgs& A0, hipStream_t st) {
  CflArgs A = A0;
  const unsigned nb = (unsigned)cfl_blocks(G);
  A.nblocks = (int)nb;
  if (g_rec) {
    AnyArgs a{};
    a.cf = A;
    record(LK_CFL, G.icha, 256, nb, G, a);
    record(LK_CFL_FINISH, 0, 1024, 1, G, a);
    return hipSuccess;
  }
  with_icha(G.icha, [&](auto C) { hipLaunchKernelGGL((k_cfl<C.value>), dim3(nb), dim3(256), 0, st, G, A); });
  hipLaunchKernelGGL(k_cfl_finish, dim3(1), dim3(1024), 0, st, A);
  return hipGetLastError();
}

hipError_t launch_advance_time(const Geo& G, const unsigned long long* dtbits, double cfl, double* tdt,
                               hipStream_t st) {
  AdvArgs A{dtbits, cfl, tdt};
  if (g_rec) { AnyArgs a{}; a.a = A; record(LK_ADVANCE, 0, 64, 1, G, a); return hipSuccess; }
  hipLaunchKernelGGL(k_advance_time, dim3(1), dim3(64), 0, st, A);
  return hipGetLastError();
}

hipError_t launch_copy(const Geo& G, double* dst, const double* src, long long n, hipStream_t st) {
  if (g_rec) {
    AnyArgs a{};
    a.cp = CopyArgs{src, dst, n};
    record(LK_COPY, 0, 256, nblk(n, 256), G, a);
    return hipSuccess;
  }
  return hipMemcpyAsync(dst, src, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, st);
}

hipError_t launch_batched(int kind, int variant, int threads, const BatchItem* it, int n, long long total_blocks,
                          hipStream_t st) {
  if (n <= 0 || total_blocks <= 0) return hipSuccess;
  if (total_blocks > 0x7fffffffLL) return hipErrorInvalidValue;
  if (!one_group_per_model(kind) && (threads < 64 || threads > 256 || threads % 64)) return hipErrorInvalidValue;
  const dim3 g((unsigned)total_blocks), b((unsigned)threads);
  switch (kind) {
    case LK_EDGES: {
      const int d = variant & 3;
      const bool mx = (variant & 4) != 0, mc = (variant & 16) != 0;
#define L(K, Dd, T) hipLaunchKernelGGL((K<Dd, T>), g, b, 0, st, it, n)
#define LD(K, T) do { if (d == 1) L(K, 1, T); else if (d == 2) L(K, 2, T); else L(K, 3, T); } while (0)
      if (mc) { if (mx) LD(kb_mc_edges, float); else LD(kb_mc_edges, double); }
      else { if (mx) LD(kb_ppm_edges, float); else LD(kb_ppm_edges, double); }
#undef LD
#undef L
      break;
    }
    case LK_SWEEP:
      with_sweep(variant, [&](auto D, auto I, auto T, auto C) {
        hipLaunchKernelGGL((kb_sweep<D.value, I.value, typename decltype(T)::type, C.value>), g, b, 0, st, it, n);
      });
      break;
    case LK_PSWEEP: {
      const size_t lds = psweep_lds_bytes(threads, (variant & 4) != 0);
      with_psweep(variant, [&](auto D, auto T, auto C) {
        hipLaunchKernelGGL((kb_psweep<D.value, typename decltype(T)::type, C.value>), g, b, lds, st, it, n);
      });
      break;
    }
    case LK_FLUXCT:
      if (variant == 1) hipLaunchKernelGGL(kb_fluxct<float>, g, b, 0, st, it, n);
      else hipLaunchKernelGGL(kb_fluxct<double>, g, b, 0, st, it, n);
      break;
    case LK_RK: {
      const int stg = variant & 7;
      const bool sym = (variant & 8) != 0, f32 = (variant & 16) != 0;
#define LT(S, Y) \
  do { if (f32) hipLaunchKernelGGL((kb_rk<S, Y, float>), g, b, 0, st, it, n); \
       else hipLaunchKernelGGL((kb_rk<S, Y, double>), g, b, 0, st, it, n); } while (0)
#define L(S) do { if (sym) LT(S, true); else LT(S, false); } while (0)
      if (stg == 1) L(1);
      else if (stg == 2) L(2);
      else if (stg == 3) L(3);
      else L(4);
#undef L
#undef LT
      break;
    }
    case LK_RECOV:
      if (variant == 2) hipLaunchKernelGGL(kb_recov<2>, g, b, 0, st, it, n);
      else hipLaunchKernelGGL(kb_recov<1>, g, b, 0, st, it, n);
      break;
    case LK_BND: hipLaunchKernelGGL(kb_bnd, g, b, 0, st, it, n); break;
    case LK_POST: hipLaunchKernelGGL(kb_post, g, b, 0, st, it, n); break;
    case LK_CFL:
      with_icha(variant, [&](auto C) { hipLaunchKernelGGL((kb_cfl<C.value>), g, b, 0, st, it, n); });
      break;
    case LK_COPY: hipLaunchKernelGGL(kb_copy, g, b, 0, st, it, n); break;
    case LK_CARRY:
      if (variant == 2) hipLaunchKernelGGL(kb_recov_carry<2>, dim3(n), dim3(1024), 0, st, it, n);
      else hipLaunchKernelGGL(kb_recov_carry<1>, dim3(n), dim3(1024), 0, st, it, n);
      break;
    case LK_CFL_FINISH: hipLaunchKernelGGL(kb_cfl_finish, dim3(n), dim3(1024), 0, st, it, n); break;
    case LK_ADVANCE: hipLaunchKernelGGL(kb_advance_time, dim3(n), dim3(64), 0, st, it, n); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

__global__ void k_gather_first(const double* const* __restrict__ src, double* __restrict__ dst, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[t] = src[t][0];
}

hipError_t launch_gather_first(const double* const* src, double* dst, int n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_first, dim3(nblk(n, 256)), dim3(256), 0, st, src, dst, n);
  return hipGetLastError();
}

hipError_t launch_caluu1(const Geo& G, const double* W, const double* metC, double gam, double* U, hipStream_t st) {
  hipLaunchKernelGGL(k_caluu1, dim3(nblk(G.N, 256)), dim3(256), 0, st, G, W, metC, gam, U);
  return hipGetLastError();
}

hipError_t launch_aos_to_soa(const double* aos, double* soa, long long N, hipStream_t st) {
  hipLaunchKernelGGL(k_aos_to_soa, dim3(nblk(N * 9, 256)), dim3(256), 0, st, aos, soa, N);
  return hipGetLastError();
}
hipError_t launch_soa_to_aos(const double* soa, double* aos, long long N, hipStream_t st) {
  hipLaunchKernelGGL(k_soa_to_aos, dim3(nblk(N * 9, 256)), dim3(256), 0, st, soa, aos, N);
  return hipGetLastError();
}

hipError_t launch_ww_to_planes(const double* ww, double* F, long long N, hipStream_t st) {
  if (N == 0) return hipSuccess;
  hipLaunchKernelGGL(k_ww_to_planes, dim3(nblk(N, 256)), dim3(256), 0, st, ww, F, N);
  return hipGetLastError();
}
hipError_t launch_planes_to_ww(const double* F, double* ww, long long N, hipStream_t st) {
  if (N == 0) return hipSuccess;
  hipLaunchKernelGGL(k_planes_to_ww, dim3(nblk(N, 256)), dim3(256), 0, st, F, ww, N);
  return hipGetLastError();
}

hipError_t launch_retype(const void* src, void* dst, long long n, bool to_float, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (to_float)
    hipLaunchKernelGGL((k_retype<double, float>), dim3(nblk(n, 256)), dim3(256), 0, st, (const double*)src, (float*)dst, n);
  else
    hipLaunchKernelGGL((k_retype<float, double>), dim3(nblk(n, 256)), dim3(256), 0, st, (const float*)src, (double*)dst, n);
  return hipGetLastError();
}

}  // namespace rh
