// The magnetised Fishbone-Moncrief torus built on the device, straight into
// a model's state (raishin_model_set_torus): the device-side counterpart of
// raishin_mdtorus1 (mdgrmhd.f90:879-1654), for models whose metric tables the
// device built, so that set-up needs no host arrays at all.
//
// The column arithmetic is rhtorus::column, shared with the host set-up; the
// host supplies what needs glibc (the torus constants, sin and cos of every
// theta), so the device torus differs from the host's only where the device's
// log, exp and pow round differently.  The steps follow the host set-up:
//   columns (r, theta) -> rho_max -> normalise, noise and floor every cell ->
//   A_phi at the corners -> B = curl A / sqrt(-g) -> min 2p/b^2 -> scale B.
// The pressure noise keeps the host's stream: draw n of the seeded splitmix64
// stream goes to the n-th torus cell in (k, j, i) order (one draw per (i,k)
// column when one phi row stands for all), n found from per-row counts of
// torus columns.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/raishin_hip.h"
#include "../../include/raishin_setup.h"
#include "rh_model.hpp"
#include "rh_torus.hpp"

namespace rhm {
rhtorus::Consts torus_consts(const raishin_pram& p, const raishin_torus& t);
}

namespace {

using rh::Geo;

struct TorusArgs {
  rhtorus::Consts c;
  const double *x1, *x3, *trig;  // trig per k: sin(theta) clamped, cos(theta), folded pair (s, c)
  const double *r15, *r25, *metC;
  double* W;     // the model's primitives (9 planes)
  double* col;   // 7 planes of (i,k): de, pr, u1, u2, u3, torus flag, torus columns before i in row k
  double* pom;   // A_phi at cell corners, one per stored cell
  double* red;   // [0] rho_max [1] min 2p/b^2 [2, 2+nk) columns in rows before k [2+nk, 2+2nk) in row k,
                 // then per-workgroup partial minima
  double dmin, pmin, gam, noise, acut, beta;
  unsigned long long seed;
  int nred;      // workgroups of the partial minima
  int axisym;    // one noise draw per (i,k) column (mdtorus1's rule: at most one phi cell inside nm0 + 1)
};

__device__ long long cell_of(const Geo& G, int i, int j, int k) {
  const long long jr = G.jfix ? 0 : (long long)(j - 1);
  return (long long)(i - 1) + (long long)G.ni * (jr + (long long)G.njs * (long long)(k - 1));
}
__device__ long long col_of(const Geo& G, int i, int k) { return (long long)(i - 1) + (long long)G.ni * (k - 1); }
// stored cell t -> (i, j, k), j the row it stands for
__device__ bool stored_cell(const Geo& G, long long t, int& i, int& j, int& k) {
  if (t >= G.N) return false;
  i = (int)(t % G.ni) + 1;
  const long long r = t / G.ni;
  j = G.jfix ? G.jfix : (int)(r % G.njs) + 1;
  k = (int)(r / G.njs) + 1;
  return true;
}

// one thread per (i,k) column
__global__ void __launch_bounds__(256) k_torus_columns(Geo G, TorusArgs A) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long P2 = (long long)G.ni * G.nk;
  if (t >= P2) return;
  const int i = (int)(t % G.ni) + 1, k = (int)(t / G.ni) + 1;
  const double* tr = A.trig + 4 * (k - 1);
  const rhtorus::Column c = rhtorus::column(A.c, A.x1[i - 1], A.x3[k - 1], tr[0], tr[1], tr[2], tr[3]);
  A.col[0 * P2 + t] = c.de;
  A.col[1 * P2 + t] = c.pr;
  A.col[2 * P2 + t] = c.u[1];
  A.col[3 * P2 + t] = c.u[2];
  A.col[4 * P2 + t] = c.u[3];
  A.col[5 * P2 + t] = c.in ? 1.0 : 0.0;
}

// One workgroup of 1024: rho_max over the torus columns, and per row k the
// torus columns before each i and in the whole row, then the rows before k.
__global__ void __launch_bounds__(1024) k_torus_scan(Geo G, TorusArgs A) {
  __shared__ double s_max[1024];
  const long long P2 = (long long)G.ni * G.nk;
  double mx = 0.0;
  for (long long t = threadIdx.x; t < P2; t += blockDim.x)
    if (A.col[5 * P2 + t] != 0.0 && A.col[t] > mx) mx = A.col[t];
  s_max[threadIdx.x] = mx;
  double* C = A.red + 2;
  double* R = A.red + 2 + G.nk;
  for (int k = threadIdx.x + 1; k <= G.nk; k += blockDim.x) {
    double n = 0.0;
    for (int i = 1; i <= G.ni; i++) {
      A.col[6 * P2 + col_of(G, i, k)] = n;
      if (A.col[5 * P2 + col_of(G, i, k)] != 0.0) n += 1.0;
    }
    R[k - 1] = n;
  }
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s && s_max[threadIdx.x + s] > s_max[threadIdx.x]) s_max[threadIdx.x] = s_max[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    A.red[0] = s_max[0];
    double acc = 0.0;
    for (int k = 1; k <= G.nk; k++) {
      C[k - 1] = acc;
      acc += R[k - 1];
    }
  }
}

// every stored cell: the atmosphere, or the normalised torus with its noise
// and the floor check (mdtorus1's second pass)
__global__ void __launch_bounds__(256) k_torus_fill(Geo G, TorusArgs A) {
  int i, j, k;
  if (!stored_cell(G, (long long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
  const long long N = G.N, c = cell_of(G, i, j, k), P2 = (long long)G.ni * G.nk, q = col_of(G, i, k);
  const double gam = A.gam;
  const double de0 = A.dmin * A.r15[i - 1], pr0 = A.pmin * (gam - 1.0) * A.r25[i - 1];
  double w[9] = {de0, 0.0, 0.0, 0.0, pr0, pr0 * (1.0 / pow(de0, gam - 1.0)), 0.0, 0.0, 0.0};
  if (A.col[5 * P2 + q] != 0.0) {
    const double rhomax = A.red[0];
    double de = A.col[q] * (1.0 / rhomax), pr = A.col[P2 + q] * (1.0 / rhomax);
    const double Ck = A.red[2 + (k - 1)], Rk = A.red[2 + G.nk + (k - 1)], before = A.col[6 * P2 + q];
    // the draw of this cell: one per column when one row stands for all
    const double n = A.axisym ? Ck + before + 1.0 : (double)G.nj * Ck + (double)(j - 1) * Rk + before + 1.0;
    const double u = rhtorus::splitmix_nth(A.seed, (unsigned long long)n);
    pr = pr * (1.0 + A.noise * (u - 0.5));
    if (de <= de0 || pr <= pr0) {
      de = de0;
      pr = pr0;
    } else {
      w[1] = A.col[2 * P2 + q]; w[2] = A.col[3 * P2 + q]; w[3] = A.col[4 * P2 + q];
    }
    w[0] = de; w[4] = pr; w[5] = pr * (1.0 / pow(de, gam - 1.0));
  }
#pragma unroll
  for (int m = 0; m < 9; m++) A.W[(long long)m * N + c] = w[m];
}

// A_phi at the corner below (i,k) from the corner-averaged density
// (mdgrmhd.f90:1350-1460); the first row and column copy the second
__global__ void __launch_bounds__(256) k_torus_pom(Geo G, TorusArgs A) {
  int i, j, k;
  if (!stored_cell(G, (long long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
  const int ii = i < 2 ? 2 : i, kk = k < 2 ? 2 : k;
  const double* D = A.W;
  const double rhoav = 0.25 * (D[cell_of(G, ii, j, kk)] + D[cell_of(G, ii - 1, j, kk)] + D[cell_of(G, ii, j, kk - 1)] +
                               D[cell_of(G, ii - 1, j, kk - 1)]);
  const double pm = (rhoav != 0.0) ? rhoav - A.acut : 0.0;
  A.pom[cell_of(G, i, j, k)] = pm > 0.0 ? pm : 0.0;
}

// B^r and B^theta = curl A / sqrt(-g) at the cell, the edge cells copying
// the nearest interior one (the host's copies of k = 1, nk, then i = 1, ni)
__global__ void __launch_bounds__(256) k_torus_field(Geo G, TorusArgs A) {
  int i, j, k;
  if (!stored_cell(G, (long long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
  const int ii = i < 2 ? 2 : (i > G.ni - 1 ? G.ni - 1 : i);
  const int kk = k < 2 ? 2 : (k > G.nk - 1 ? G.nk - 1 : k);
  const long long N = G.N, c = cell_of(G, i, j, k);
  const double* P = A.pom;
  const double p11 = P[cell_of(G, ii + 1, j, kk + 1)], p01 = P[cell_of(G, ii, j, kk + 1)];
  const double p10 = P[cell_of(G, ii + 1, j, kk)], p00 = P[cell_of(G, ii, j, kk)];
  const double ta = 0.5 * (p11 + p01 - p10 - p00);
  const double dx3a = 0.5 * (A.x3[kk] + A.x3[kk - 1]) - 0.5 * (A.x3[kk - 1] + A.x3[kk - 2]);
  const double tb = 0.5 * (p11 + p10 - p01 - p00);
  const double dx1a = 0.5 * (A.x1[ii] + A.x1[ii - 1]) - 0.5 * (A.x1[ii - 1] + A.x1[ii - 2]);
  const double detg = A.metC[32LL * ((long long)G.ni * G.nk) + col_of(G, ii, kk)];
  A.W[6 * N + c] = (ta != 0.0) ? -ta * (1.0 / (dx3a * detg)) : 0.0;
  A.W[7 * N + c] = 0.0;
  A.W[8 * N + c] = (tb != 0.0) ? tb * (1.0 / (dx1a * detg)) : 0.0;
}

// min over cells of 2p/b^2 (1e5 where b^2 = 0): partial minima per workgroup
__global__ void __launch_bounds__(256) k_torus_beta(Geo G, TorusArgs A) {
  __shared__ double s_min[256];
  double mn = 1.0e5;
  const long long P2 = (long long)G.ni * G.nk;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < G.N; t += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    stored_cell(G, t, i, j, k);
    const long long c = cell_of(G, i, j, k), q = col_of(G, i, k);
    double w[9], g[16], gi[16];
    for (int m = 0; m < 9; m++) w[m] = A.W[(long long)m * G.N + c];
    for (int e = 0; e < 16; e++) { g[e] = A.metC[(long long)e * P2 + q]; gi[e] = A.metC[(long long)(16 + e) * P2 + q]; }
    const double b2 = rhtorus::bsq(w, g, gi);
    const double pb = (b2 != 0.0) ? 2.0 * w[4] / b2 : 1.0e5;
    if (pb < mn) mn = pb;
  }
  s_min[threadIdx.x] = mn;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s && s_min[threadIdx.x + s] < s_min[threadIdx.x]) s_min[threadIdx.x] = s_min[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) A.red[2 + 2 * G.nk + blockIdx.x] = s_min[0];
}

// beta_min: the smallest partial (one workgroup)
__global__ void __launch_bounds__(256) k_torus_betamin(Geo G, TorusArgs A) {
  __shared__ double s_min[256];
  double mn = 1.0e5;
  for (int b = threadIdx.x; b < A.nred; b += blockDim.x) {
    const double v = A.red[2 + 2 * G.nk + b];
    if (v < mn) mn = v;
  }
  s_min[threadIdx.x] = mn;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s && s_min[threadIdx.x + s] < s_min[threadIdx.x]) s_min[threadIdx.x] = s_min[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) A.red[1] = s_min[0];
}

// B *= sqrt(beta_min / beta)
__global__ void __launch_bounds__(256) k_torus_scale(Geo G, TorusArgs A) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= G.N) return;
  const double bnorm = (A.beta == 0.0) ? 0.0 : sqrt(A.red[1] / A.beta);
  for (int m = 6; m <= 8; m++) A.W[(long long)m * G.N + t] = A.W[(long long)m * G.N + t] * bnorm;
}

unsigned blocks(long long n, int bt) { return (unsigned)((n + bt - 1) / bt); }

}  // namespace

extern "C" int raishin_model_set_torus(raishin_model* m, const raishin_torus* t) {
  using namespace rhm;
  if (!m || !t) { set_error("raishin_model_set_torus: null argument"); return -1; }
  if (m->p.metric != 303) { set_error("raishin_model_set_torus: the torus needs the Kerr-Schild metric (303)"); return -1; }
  if (m->need_wo) { set_error("raishin_model_set_torus: fixed-value boundaries are not supported"); return -1; }
  if (m->G.ni < 4 || m->G.nk < 4) { set_error("raishin_model_set_torus: needs at least 4 cells in r and theta"); return -1; }
  auto go = [&]() -> bool {
    HIPCHK(hipStreamSynchronize(m->st));
    // the torus is the same in every phi row; one row stands for all when it can
    set_rows(m, collapsible(m));
    const Geo& G = m->G;
    const int ni = G.ni, nk = G.nk;
    const long long N = G.N, P2 = (long long)ni * nk;
    // host: the constants and per-theta sin / cos (glibc, as the host set-up)
    std::vector<double> x3(nk);
    HIPCHK(hipMemcpyAsync(x3.data(), m->xs[2], nk * sizeof(double), hipMemcpyDeviceToHost, m->st));
    HIPCHK(hipStreamSynchronize(m->st));
    const rhtorus::Consts c = torus_consts(m->p, *t);
    std::vector<double> trig(4 * (size_t)nk), tt(8 * (size_t)nk);
    theta_trig_table(m->p, x3.data(), x3.data(), tt.data());  // entries 0, 1: the folded, clamped pair
    for (int k = 0; k < nk; k++) {
      double sth = std::sin(x3[k]);
      if (std::fabs(sth) < c.small) sth = (sth >= 0.0) ? c.small : -c.small;
      trig[4 * k] = sth;
      trig[4 * k + 1] = std::cos(x3[k]);
      trig[4 * k + 2] = tt[8 * k];
      trig[4 * k + 3] = tt[8 * k + 1];
    }
    // scratch: the step's own work buffers (they hold nothing between steps)
    const int bt = 256;
    const int nred = (int)std::min<long long>(1024, blocks(N, bt));
    if (7 * P2 + N > 9 * N || 4 * (long long)nk + 2 + 2 * nk + nred > 9 * N) {
      set_error("raishin_model_set_torus: grid too small for the scratch layout");
      return false;
    }
    TorusArgs A{};
    A.c = c;
    A.x1 = m->xs[0];
    A.x3 = m->xs[2];
    double* trig_d = m->Wpre;
    A.trig = trig_d;
    A.red = m->Wpre + 4 * (long long)nk;
    A.col = m->Wb;
    A.pom = m->Wb + 7 * P2;
    A.r15 = m->r15;
    A.r25 = m->r25;
    A.metC = m->met[0];
    A.W = m->W;
    A.dmin = m->p.dmin;
    A.pmin = m->p.pmin;
    A.gam = m->p.gam;
    A.noise = t->noise;
    A.acut = t->acut;
    A.beta = t->beta;
    A.seed = t->seed;
    A.nred = nred;
    A.axisym = (G.nj - 2 * (G.nm0 + 1)) <= 1;
    HIPCHK(hipMemcpyAsync(trig_d, trig.data(), trig.size() * sizeof(double), hipMemcpyHostToDevice, m->st));
    hipLaunchKernelGGL(k_torus_columns, dim3(blocks(P2, bt)), dim3(bt), 0, m->st, G, A);
    hipLaunchKernelGGL(k_torus_scan, dim3(1), dim3(1024), 0, m->st, G, A);
    double rhomax = 0.0;
    HIPCHK(hipMemcpyAsync(&rhomax, A.red, sizeof(double), hipMemcpyDeviceToHost, m->st));
    HIPCHK(hipStreamSynchronize(m->st));
    if (!(rhomax > 0.0)) {
      set_error("raishin_model_set_torus: no torus cells on this grid");
      return false;
    }
    hipLaunchKernelGGL(k_torus_fill, dim3(blocks(N, bt)), dim3(bt), 0, m->st, G, A);
    hipLaunchKernelGGL(k_torus_pom, dim3(blocks(N, bt)), dim3(bt), 0, m->st, G, A);
    hipLaunchKernelGGL(k_torus_field, dim3(blocks(N, bt)), dim3(bt), 0, m->st, G, A);
    hipLaunchKernelGGL(k_torus_beta, dim3(nred), dim3(bt), 0, m->st, G, A);
    hipLaunchKernelGGL(k_torus_betamin, dim3(1), dim3(256), 0, m->st, G, A);
    hipLaunchKernelGGL(k_torus_scale, dim3(blocks(N, bt)), dim3(bt), 0, m->st, G, A);
    HIPCHK(hipGetLastError());
    // the rest as a set_state with uu computed: conserved variables (main.f90:530),
    // time 0, ww from zero (main.f90:374-376)
    HIPCHK(rh::launch_caluu1(G, m->W, m->met[0], m->p.gam, m->Ucur, m->st));
    const double tdt[2] = {0.0, 0.0};
    HIPCHK(hipMemcpyAsync(m->tdt, tdt, 16, hipMemcpyHostToDevice, m->st));
    HIPCHK(hipMemsetAsync(m->F, 0, (size_t)27 * N * 8, m->st));
    HIPCHK(hipStreamSynchronize(m->st));
    return true;
  };
  return go() ? 0 : -1;
}
