// C-ABI of the MI355X-native RAISHIN GRMHD step: resident models and the
// reference-named entry points.  Host code; kernels live in rh_kernels.hip.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/raishin_hip.h"
#include "rh_kernels.hpp"

using namespace rh;

namespace {

std::mutex g_err_mu;
std::string g_err;

void set_error(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  std::lock_guard<std::mutex> lk(g_err_mu);
  if (g_err.empty()) g_err = buf;  // sticky: keep the first error
}

#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) {                                                        \
      set_error("%s:%d %s: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_));    \
      return false;                                                                \
    }                                                                              \
  } while (0)

constexpr int NCOUNT = 11;

int ghost_nm0(int irec) {
  // main.f90:179-211
  if (irec == 1 || irec == 2 || irec == 3 || irec == 11 || irec == 12 || irec == 13) return 2;
  if (irec == 4 || irec == 6 || irec == 7 || irec == 8 || irec == 9 || irec == 10) return 3;
  if (irec == 5) return 4;
  return 1;
}

bool check_pram(const raishin_pram& p) {
  if (p.imax < 1 || p.jmax < 1 || p.kmax < 1) { set_error("bad grid %d %d %d", p.imax, p.jmax, p.kmax); return false; }
  if (p.irec != 5 && p.irec != 6 && p.irec != 7) { set_error("irec=%d not built (5 PPM, 6 MP5, 7 WENO5)", p.irec); return false; }
  if (p.irkt != 2 && p.irkt != 3) { set_error("irkt=%d not supported", p.irkt); return false; }
  if (p.ict != 1) { set_error("ict=%d not built (1 = flux-CT)", p.ict); return false; }
  if (p.ihll != 1) { set_error("ihll=%d not built (1 = HLLE)", p.ihll); return false; }
  if (p.icha != 2) { set_error("icha=%d not built (2 = Del Zanna speeds)", p.icha); return false; }
  if (p.iwvec != 1) { set_error("iwvec=%d not built (1 = recov1a)", p.iwvec); return false; }
  if (p.ieos != 0) { set_error("ieos=%d not built (0 = gamma law)", p.ieos); return false; }
  const int nm1 = ghost_nm0(p.irec) + 1;
  const int n[3] = {p.imax, p.jmax, p.kmax};
  const int* bcs[6] = {p.iboux1in, p.iboux1ot, p.iboux2in, p.iboux2ot, p.iboux3in, p.iboux3ot};
  for (int a = 0; a < 3; a++) {
    if (n[a] - 2 * nm1 < 1) { set_error("axis %d: %d cells leave no interior with %d ghosts", a + 1, n[a], nm1); return false; }
    for (int side = 0; side < 2; side++)
      for (int m = 0; m < 9; m++) {
        const int t = bcs[a * 2 + side][m];
        const bool ok = t == 1 || t == 2 || t == 5 || t == 12 || t == 13 || t == 14 || (t == 18 && a == 0 && side == 1);
        if (!ok) { set_error("boundary code %d (axis %d side %d var %d) not built", t, a + 1, side, m + 1); return false; }
        if ((t == 12 || t == 14) && n[a] - 2 * nm1 < nm1) {
          set_error("mirror boundary on axis %d needs at least %d interior cells", a + 1, nm1);
          return false;
        }
      }
  }
  if (p.irec == 5 && ghost_nm0(p.irec) != 4) return false;
  return true;
}

}  // namespace

// Kernel kinds timed with HIP events when timing is on (raishin_model_set_timing).
enum { K_SWEEP1 = 0, K_SWEEP2, K_SWEEP3, K_FLUXCT, K_UPDATE, K_CARRY, K_POST, K_CFL, K_NKINDS };

struct raishin_model {
  raishin_pram p;
  // per-kind HIP event pairs recorded around launches; resolved on read
  bool timing = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[K_NKINDS];
  size_t ev_used[K_NKINDS] = {0};
  double ev_ms[K_NKINDS] = {0};
  long long ev_n[K_NKINDS] = {0};
  Geo G;
  hipStream_t st = nullptr;
  bool need_wo = false;
  // state (9 planes each)
  double *W = nullptr, *Wb = nullptr, *Wpre = nullptr, *Wo = nullptr;
  double *Ucur = nullptr, *Uh = nullptr, *Us = nullptr, *Unext = nullptr;
  // F: ww(3,9,...) (27 planes); CT: the six transverse B flux copies ct.f90 averages
  double *F = nullptr, *CT = nullptr, *wgout = nullptr, *stage_aos = nullptr;
  // geometry
  double *met[4] = {nullptr, nullptr, nullptr, nullptr};  // centre, i-face, j-face, k-face
  double *chr = nullptr;
  double *xs[3] = {nullptr, nullptr, nullptr}, *dxb[3] = {nullptr, nullptr, nullptr}, *dxs[3] = {nullptr, nullptr, nullptr};
  double *r15 = nullptr, *r25 = nullptr;
  int* bc = nullptr;
  unsigned long long *dtbits = nullptr, *cnt = nullptr, *carry = nullptr;
  unsigned char *cflag = nullptr, *pend = nullptr;
  long long carry_cap = 0;  // room for every interior cell: the flagged list cannot overflow
  long long* cfl_last = nullptr;
  double* cfl_min = nullptr;
  double* tdt = nullptr;  // [0] time [1] last dt
};

namespace {

template <typename T>
bool dalloc(T** p, size_t n) {
  HIPCHK(hipMalloc((void**)p, sizeof(T) * (n ? n : 1)));
  return true;
}

void dfree_all(raishin_model* m) {
  double** bufs[] = {&m->W, &m->Wb, &m->Wpre, &m->Wo, &m->Ucur, &m->Uh, &m->Us, &m->Unext, &m->F, &m->CT,
                     &m->wgout, &m->stage_aos, &m->chr, &m->r15, &m->r25, &m->tdt};
  for (auto b : bufs) if (*b) { (void)hipFree(*b); *b = nullptr; }
  for (int a = 0; a < 4; a++) if (m->met[a]) { (void)hipFree(m->met[a]); m->met[a] = nullptr; }
  for (int a = 0; a < 3; a++) {
    if (m->xs[a]) (void)hipFree(m->xs[a]);
    if (m->dxb[a]) (void)hipFree(m->dxb[a]);
    if (m->dxs[a]) (void)hipFree(m->dxs[a]);
    m->xs[a] = m->dxb[a] = m->dxs[a] = nullptr;
  }
  if (m->bc) (void)hipFree(m->bc);
  if (m->dtbits) (void)hipFree(m->dtbits);
  if (m->cnt) (void)hipFree(m->cnt);
  if (m->carry) (void)hipFree(m->carry);
  if (m->cflag) (void)hipFree(m->cflag);
  if (m->pend) (void)hipFree(m->pend);
  if (m->cfl_last) (void)hipFree(m->cfl_last);
  if (m->cfl_min) (void)hipFree(m->cfl_min);
  m->cfl_last = nullptr; m->cfl_min = nullptr;
  m->bc = nullptr; m->dtbits = m->cnt = m->carry = nullptr;
  m->cflag = m->pend = nullptr;
}

// (i,k) tables of one metric location: planes 0..15 gcov, 16..31 gcon, 32 detg.
// The reference arrays are checked to be independent of j, bit for bit.
bool metric_tables(const raishin_pram& p, const double* gcov, const double* gcon, const double* detg,
                   std::vector<double>& out, const char* what) {
  const long long ni = p.imax, nj = p.jmax, nk = p.kmax, P2 = ni * nk;
  out.assign(33 * P2, 0.0);
  for (long long k = 0; k < nk; k++)
    for (long long j = 0; j < nj; j++)
      for (long long i = 0; i < ni; i++) {
        const long long c = i + ni * (j + nj * k), q = i + ni * k;
        for (int e = 0; e < 16; e++) {
          const double a = gcov[16 * c + e], b = gcon[16 * c + e];
          if (j == 0) { out[e * P2 + q] = a; out[(16 + e) * P2 + q] = b; }
          else if (std::memcmp(&a, &out[e * P2 + q], 8) || std::memcmp(&b, &out[(16 + e) * P2 + q], 8)) {
            set_error("%s metric depends on j at (%lld,%lld,%lld); only axisymmetric metrics are built", what, i + 1, j + 1, k + 1);
            return false;
          }
        }
        const double d = detg[c];
        if (j == 0) out[32 * P2 + q] = d;
        else if (std::memcmp(&d, &out[32 * P2 + q], 8)) { set_error("%s sqrt(-g) depends on j", what); return false; }
      }
  return true;
}

// Record an event pair around one launch of kind `kind` when timing is on.
struct TimedLaunch {
  raishin_model* m;
  int kind;
  hipEvent_t b = nullptr, e = nullptr;
  TimedLaunch(raishin_model* m_, int k) : m(m_), kind(k) {
    if (!m->timing) return;
    auto& v = m->ev[kind];
    if (m->ev_used[kind] == v.size()) {
      hipEvent_t x, y;
      if (hipEventCreate(&x) != hipSuccess || hipEventCreate(&y) != hipSuccess) return;
      v.emplace_back(x, y);
    }
    b = v[m->ev_used[kind]].first;
    e = v[m->ev_used[kind]].second;
    m->ev_used[kind]++;
    (void)hipEventRecord(b, m->st);
  }
  ~TimedLaunch() {
    if (b) (void)hipEventRecord(e, m->st);
  }
};

// Fold the recorded event pairs into per-kind totals (synchronises).
void resolve_timing(raishin_model* m) {
  (void)hipStreamSynchronize(m->st);
  for (int k = 0; k < K_NKINDS; k++) {
    for (size_t q = 0; q < m->ev_used[k]; q++) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, m->ev[k][q].first, m->ev[k][q].second) == hipSuccess) {
        m->ev_ms[k] += ms;
        m->ev_n[k] += 1;
      }
    }
    m->ev_used[k] = 0;
  }
}

bool upload(double* dst, const double* src, size_t n, hipStream_t st) {
  HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, st));
  return true;
}

bool build_model(raishin_model* m, const raishin_pram* p, const double* x1, const double* x2, const double* x3,
                 const double* dx1b, const double* dx2b, const double* dx3b, const double* const gcv[4],
                 const double* const gcn[4], const double* const dtg[4], const double* christ) {
  m->p = *p;
  if (!check_pram(*p)) return false;
  Geo& G = m->G;
  G.ni = p->imax; G.nj = p->jmax; G.nk = p->kmax;
  G.nm0 = ghost_nm0(p->irec); G.nm1 = G.nm0 + 1;
  G.N = (long long)G.ni * G.nj * G.nk;
  G.P2 = (long long)G.ni * G.nk;
  const long long N = G.N;
  HIPCHK(hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking));
  m->carry_cap = (long long)(G.ni - 2 * G.nm1) * (G.nj - 2 * G.nm1) * (G.nk - 2 * G.nm1);
  for (int a = 0; a < 6; a++)
    for (int q = 0; q < 9; q++) {
      const int* b = a == 0 ? p->iboux1in : a == 1 ? p->iboux1ot : a == 2 ? p->iboux2in : a == 3 ? p->iboux2ot
                                                                   : a == 4 ? p->iboux3in : p->iboux3ot;
      if (b[q] == 2) m->need_wo = true;
    }
  if (!dalloc(&m->W, 9 * N) || !dalloc(&m->Wb, 9 * N) || !dalloc(&m->Wpre, 9 * N) || !dalloc(&m->Ucur, 9 * N) ||
      !dalloc(&m->Uh, 9 * N) || !dalloc(&m->Us, 9 * N) || !dalloc(&m->Unext, 9 * N) || !dalloc(&m->F, 27 * N) ||
      !dalloc(&m->CT, 6 * N) || !dalloc(&m->wgout, N) || !dalloc(&m->chr, 64 * G.P2) || !dalloc(&m->r15, G.ni) ||
      !dalloc(&m->r25, G.ni) || !dalloc(&m->bc, 54) || !dalloc(&m->dtbits, 4) || !dalloc(&m->cnt, NCOUNT) ||
      !dalloc(&m->carry, 1 + m->carry_cap) || !dalloc(&m->tdt, 2) || !dalloc(&m->cflag, N) ||
      !dalloc(&m->pend, m->carry_cap) || !dalloc(&m->cfl_last, 4 * cfl_blocks(G)) ||
      !dalloc(&m->cfl_min, 4 * cfl_blocks(G)))
    return false;
  if (m->need_wo && !dalloc(&m->Wo, 9 * N)) return false;
  HIPCHK(hipMemsetAsync(m->cnt, 0, NCOUNT * 8, m->st));
  HIPCHK(hipMemsetAsync(m->carry, 0, (1 + m->carry_cap) * 8, m->st));
  HIPCHK(hipMemsetAsync(m->tdt, 0, 16, m->st));
  HIPCHK(hipMemsetAsync(m->W, 0, 9 * N * 8, m->st));
  // the reference zeroes its flux array once (main.f90:372-381); an interface
  // whose speeds are unordered (NaN) keeps its previous flux (hll.f90:86-115)
  HIPCHK(hipMemsetAsync(m->F, 0, 27 * N * 8, m->st));
  HIPCHK(hipMemsetAsync(m->CT, 0, 6 * N * 8, m->st));
  HIPCHK(hipMemsetAsync(m->Wpre, 0, 9 * N * 8, m->st));
  HIPCHK(hipMemsetAsync(m->cflag, 0, N, m->st));
  // metric tables
  const char* names[4] = {"centre", "i-face", "j-face", "k-face"};
  std::vector<double> tab;
  for (int a = 0; a < 4; a++) {
    if (!metric_tables(*p, gcv[a], gcn[a], dtg[a], tab, names[a])) return false;
    if (!dalloc(&m->met[a], tab.size())) return false;
    HIPCHK(hipMemcpy(m->met[a], tab.data(), tab.size() * 8, hipMemcpyHostToDevice));
  }
  {
    std::vector<double> ct(64 * G.P2);
    for (long long k = 0; k < G.nk; k++)
      for (long long j = 0; j < G.nj; j++)
        for (long long i = 0; i < G.ni; i++) {
          const long long c = i + G.ni * (j + G.nj * k), q = i + G.ni * k;
          for (int e = 0; e < 64; e++) {
            const double v = christ[64 * c + e];
            if (j == 0) ct[e * G.P2 + q] = v;
            else if (std::memcmp(&v, &ct[e * G.P2 + q], 8)) { set_error("Christoffel symbols depend on j"); return false; }
          }
        }
    HIPCHK(hipMemcpy(m->chr, ct.data(), ct.size() * 8, hipMemcpyHostToDevice));
  }
  // 1D grids; dxs[d] = rec2's dummy dx*b(d) = actual dx*b(d-1) (rec.f90:58), clamped at the ends
  const double* xsrc[3] = {x1, x2, x3};
  const double* dsrc[3] = {dx1b, dx2b, dx3b};
  const int nn[3] = {G.ni, G.nj, G.nk};
  for (int a = 0; a < 3; a++) {
    const int n = nn[a];
    if (!dalloc(&m->xs[a], n) || !dalloc(&m->dxb[a], n) || !dalloc(&m->dxs[a], n + 2)) return false;
    std::vector<double> db(n, 0.0), ds(n + 2, 0.0);
    for (int q = 0; q < n - 1; q++) db[q] = dsrc[a][q];
    for (int d = 0; d <= n + 1; d++) {
      int q = d - 1;  // 1-based actual index
      if (q < 1) q = 1;
      if (q > n - 1) q = n - 1;
      ds[d] = (n > 1) ? dsrc[a][q - 1] : 0.0;
    }
    HIPCHK(hipMemcpy(m->xs[a], xsrc[a], n * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(m->dxb[a], db.data(), n * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(m->dxs[a], ds.data(), (n + 2) * 8, hipMemcpyHostToDevice));
  }
  // r-power tables from the host libm, as the reference computes them (correction.f90:318-319)
  std::vector<double> r15(G.ni), r25(G.ni);
  for (int i = 0; i < G.ni; i++) { r15[i] = std::pow(x1[i], -1.5); r25[i] = std::pow(x1[i], -2.5); }
  HIPCHK(hipMemcpy(m->r15, r15.data(), G.ni * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(m->r25, r25.data(), G.ni * 8, hipMemcpyHostToDevice));
  int bch[54];
  const int* bsrc[6] = {p->iboux1in, p->iboux1ot, p->iboux2in, p->iboux2ot, p->iboux3in, p->iboux3ot};
  for (int a = 0; a < 6; a++) for (int q = 0; q < 9; q++) bch[a * 9 + q] = bsrc[a][q];
  HIPCHK(hipMemcpy(m->bc, bch, sizeof bch, hipMemcpyHostToDevice));
  HIPCHK(hipStreamSynchronize(m->st));
  return true;
}

struct StageTap {  // optional host copies of a stage's intermediates (reference layouts)
  double *ww = nullptr, *wpre = nullptr;
};
bool copy_planes_out(raishin_model* m, const double* planes, int nplanes, int stride_out, int offset_out,
                     double* host) {
  // planes[n*N + c] -> host[offset_out + stride_out*(n + nplanes*c)] via a staging buffer
  const long long N = m->G.N;
  std::vector<double> tmp((size_t)nplanes * N);
  HIPCHK(hipMemcpyAsync(tmp.data(), planes, (size_t)nplanes * N * 8, hipMemcpyDeviceToHost, m->st));
  HIPCHK(hipStreamSynchronize(m->st));
  for (long long c = 0; c < N; c++)
    for (int n = 0; n < nplanes; n++) host[offset_out + (size_t)stride_out * (n + (size_t)nplanes * c)] = tmp[(size_t)n * N + c];
  return true;
}

// One RK stage: sweeps -> flux-CT -> update (RK + recov) -> fix-up -> ghosts/floors/conserved.
bool run_stage(raishin_model* m, int rkstage, const double* Ua, const double* Ub, double* Uout,
               const StageTap* tap = nullptr) {
  const Geo& G = m->G;
  const long long N = G.N;
  double* F1 = m->F;
  double* F2 = m->F + 9 * N;
  double* F3 = m->F + 18 * N;
  { TimedLaunch t(m, K_SWEEP1); HIPCHK(launch_sweep(G, 1, m->p.irec, m->W, m->met[1], m->dxs[0], m->xs[0], m->p.gam, F1, m->CT, m->st)); }
  { TimedLaunch t(m, K_SWEEP2); HIPCHK(launch_sweep(G, 2, m->p.irec, m->W, m->met[2], m->dxs[1], m->xs[1], m->p.gam, F2, m->CT, m->st)); }
  { TimedLaunch t(m, K_SWEEP3); HIPCHK(launch_sweep(G, 3, m->p.irec, m->W, m->met[3], m->dxs[2], m->xs[2], m->p.gam, F3, m->CT, m->st)); }
  { TimedLaunch t(m, K_FLUXCT); HIPCHK(launch_fluxct(G, m->CT, F1, F2, F3, m->st)); }
  if (tap && tap->ww) {
    // ww(d,n,i,j,k) as after ct()
    for (int d = 0; d < 3; d++)
      if (!copy_planes_out(m, m->F + (long long)d * 9 * N, 9, 3, d, tap->ww)) return false;
  }
  StageArgs A{};
  A.W = m->W; A.Ua = Ua; A.Ub = Ub; A.F1 = F1; A.F2 = F2; A.F3 = F3;
  A.metC = m->met[0]; A.chr = m->chr; A.dx1b = m->dxb[0]; A.dx2b = m->dxb[1]; A.dx3b = m->dxb[2];
  A.r15 = m->r15; A.r25 = m->r25; A.dtbits = m->dtbits; A.Wpre = m->Wpre; A.wgout = m->wgout; A.Ukeep = Uout;
  A.cnt = m->cnt; A.carry = m->carry; A.cflag = m->cflag; A.carry_cap = m->carry_cap;
  A.gam = m->p.gam; A.cfl = m->p.cfl; A.dmin = m->p.dmin; A.pmin = m->p.pmin; A.iter = m->p.iter;
  { TimedLaunch t(m, K_UPDATE); HIPCHK(launch_update(G, rkstage, A, m->st)); }
  FixArgs X{};
  X.W = m->W; X.Ustage = Uout; X.metC = m->met[0]; X.r15 = m->r15; X.r25 = m->r25; X.Wpre = m->Wpre;
  X.wgout = m->wgout; X.carry = m->carry; X.cnt = m->cnt; X.cflag = m->cflag; X.pend = m->pend;
  X.carry_cap = m->carry_cap; X.gam = m->p.gam;
  X.dmin = m->p.dmin; X.pmin = m->p.pmin; X.iter = m->p.iter;
  { TimedLaunch t(m, K_CARRY); HIPCHK(launch_recov_carry(G, X, m->st)); }
  if (tap && tap->wpre && !copy_planes_out(m, m->Wpre, 9, 1, 0, tap->wpre)) return false;
  PostArgs P{};
  P.Wpre = m->Wpre; P.Wold = m->W; P.Wo = m->need_wo ? m->Wo : m->W; P.metC = m->met[0]; P.r15 = m->r15;
  P.r25 = m->r25; P.bc = m->bc; P.Wout = m->Wb; P.Uout = Uout; P.cnt = m->cnt; P.gam = m->p.gam;
  P.dmin = m->p.dmin; P.pmin = m->p.pmin; P.pmax = m->p.pmax; P.model = m->p.model;
  { TimedLaunch t(m, K_POST); HIPCHK(launch_post(G, P, m->st)); }
  std::swap(m->W, m->Wb);
  return true;
}

bool run_cfl(raishin_model* m) {
  CflArgs C{};
  C.W = m->W; C.metC = m->met[0]; C.dx1b = m->dxb[0]; C.dx2b = m->dxb[1]; C.dx3b = m->dxb[2];
  C.dtbits = m->dtbits; C.blast = m->cfl_last; C.bmin = m->cfl_min; C.gam = m->p.gam;
  { TimedLaunch t(m, K_CFL); HIPCHK(launch_cfl(m->G, C, m->st)); }
  return true;
}

bool run_step(raishin_model* m) {
  const long long N = m->G.N;
  if (!run_cfl(m)) return false;
  HIPCHK(launch_advance_time(m->dtbits, m->p.cfl, m->tdt, m->st));
  if (m->need_wo) HIPCHK(hipMemcpyAsync(m->Wo, m->W, 9 * N * 8, hipMemcpyDeviceToDevice, m->st));
  if (!run_stage(m, 1, m->Ucur, nullptr, m->Uh)) return false;
  if (m->p.irkt == 2) {
    if (!run_stage(m, 2, m->Ucur, m->Uh, m->Unext)) return false;
  } else {
    if (!run_stage(m, 3, m->Ucur, m->Uh, m->Us)) return false;
    if (!run_stage(m, 4, m->Ucur, m->Us, m->Unext)) return false;
  }
  std::swap(m->Ucur, m->Unext);
  return true;
}

raishin_pram g_pram;
bool g_pram_set = false;

}  // namespace

extern "C" {

int raishin_last_error(char* buf, int len) {
  std::lock_guard<std::mutex> lk(g_err_mu);
  if (g_err.empty()) return 0;
  if (buf && len > 0) {
    std::strncpy(buf, g_err.c_str(), (size_t)len - 1);
    buf[len - 1] = 0;
  }
  return 1;
}

void raishin_clear_error(void) {
  std::lock_guard<std::mutex> lk(g_err_mu);
  g_err.clear();
}

raishin_model* raishin_model_create(const raishin_pram* p, const double* x1, const double* x2, const double* x3,
                                    const double* dx1b, const double* dx2b, const double* dx3b, const double* gcov,
                                    const double* gcon, const double* gcovi, const double* gconi, const double* gcovj,
                                    const double* gconj, const double* gcovk, const double* gconk, const double* detg,
                                    const double* detgi, const double* detgj, const double* detgk,
                                    const double* christ) {
  if (!p || !x1 || !x2 || !x3 || !gcov || !gcon || !christ) { set_error("raishin_model_create: null argument"); return nullptr; }
  raishin_model* m = new raishin_model();
  const double* gcv[4] = {gcov, gcovi, gcovj, gcovk};
  const double* gcn[4] = {gcon, gconi, gconj, gconk};
  const double* dtg[4] = {detg, detgi, detgj, detgk};
  if (!build_model(m, p, x1, x2, x3, dx1b, dx2b, dx3b, gcv, gcn, dtg, christ)) {
    dfree_all(m);
    if (m->st) (void)hipStreamDestroy(m->st);
    delete m;
    return nullptr;
  }
  return m;
}

void raishin_model_destroy(raishin_model* m) {
  if (!m) return;
  (void)hipStreamSynchronize(m->st);
  for (int k = 0; k < K_NKINDS; k++)
    for (auto& pr : m->ev[k]) { (void)hipEventDestroy(pr.first); (void)hipEventDestroy(pr.second); }
  dfree_all(m);
  (void)hipStreamDestroy(m->st);
  delete m;
}

void raishin_model_set_state(raishin_model* m, const double* uri, const double* uu, double time) {
  const long long N = m->G.N;
  auto go = [&]() -> bool {
    if (!m->stage_aos && !dalloc(&m->stage_aos, 9 * N)) return false;
    if (!upload(m->stage_aos, uri, 9 * N, m->st)) return false;
    HIPCHK(launch_aos_to_soa(m->stage_aos, m->W, N, m->st));
    if (uu) {
      if (!upload(m->stage_aos, uu, 9 * N, m->st)) return false;
      HIPCHK(launch_aos_to_soa(m->stage_aos, m->Ucur, N, m->st));
    } else {
      HIPCHK(launch_caluu1(m->G, m->W, m->met[0], m->p.gam, m->Ucur, m->st));  // main.f90:530
    }
    double tdt[2] = {time, 0.0};
    HIPCHK(hipMemcpyAsync(m->tdt, tdt, 16, hipMemcpyHostToDevice, m->st));
    // a new run: ww starts from zero (main.f90:374-376); hlle keeps its old
    // value wherever the wave speeds are NaN
    HIPCHK(hipMemsetAsync(m->F, 0, (size_t)27 * N * 8, m->st));
    HIPCHK(hipStreamSynchronize(m->st));
    return true;
  };
  (void)go();
}

void raishin_model_get_state(raishin_model* m, double* uri, double* uu, double* time) {
  const long long N = m->G.N;
  auto go = [&]() -> bool {
    if (!m->stage_aos && !dalloc(&m->stage_aos, 9 * N)) return false;
    if (uri) {
      HIPCHK(launch_soa_to_aos(m->W, m->stage_aos, N, m->st));
      HIPCHK(hipMemcpyAsync(uri, m->stage_aos, 9 * N * 8, hipMemcpyDeviceToHost, m->st));
      HIPCHK(hipStreamSynchronize(m->st));
    }
    if (uu) {
      HIPCHK(launch_soa_to_aos(m->Ucur, m->stage_aos, N, m->st));
      HIPCHK(hipMemcpyAsync(uu, m->stage_aos, 9 * N * 8, hipMemcpyDeviceToHost, m->st));
      HIPCHK(hipStreamSynchronize(m->st));
    }
    if (time) {
      double tdt[2];
      HIPCHK(hipMemcpyAsync(tdt, m->tdt, 16, hipMemcpyDeviceToHost, m->st));
      HIPCHK(hipStreamSynchronize(m->st));
      *time = tdt[0];
    }
    return true;
  };
  (void)go();
}

void raishin_model_step(raishin_model* m, int nsteps, double* time, double* dt_last) {
  for (int s = 0; s < nsteps; s++)
    if (!run_step(m)) return;
  if (time || dt_last) {
    double tdt[2] = {0.0, 0.0};
    if (hipMemcpyAsync(tdt, m->tdt, 16, hipMemcpyDeviceToHost, m->st) != hipSuccess ||
        hipStreamSynchronize(m->st) != hipSuccess) {
      set_error("raishin_model_step: synchronisation failed");
      return;
    }
    if (time) *time = tdt[0];
    if (dt_last) *dt_last = tdt[1];
  }
}

void raishin_model_cdtcfl4(raishin_model* m, double* dtcfl, double* dt1, double* dt2, double* dt3) {
  if (!run_cfl(m)) return;
  unsigned long long b[4];
  if (hipMemcpyAsync(b, m->dtbits, 32, hipMemcpyDeviceToHost, m->st) != hipSuccess ||
      hipStreamSynchronize(m->st) != hipSuccess) {
    set_error("raishin_model_cdtcfl4: copy failed");
    return;
  }
  double v[4];
  std::memcpy(v, b, 32);
  if (dtcfl) *dtcfl = v[0];
  if (dt1) *dt1 = v[1];
  if (dt2) *dt2 = v[2];
  if (dt3) *dt3 = v[3];
}

void raishin_model_counters(raishin_model* m, raishin_counters* c) {
  unsigned long long h[NCOUNT];
  if (hipMemcpyAsync(h, m->cnt, sizeof h, hipMemcpyDeviceToHost, m->st) != hipSuccess ||
      hipStreamSynchronize(m->st) != hipSuccess) {
    set_error("raishin_model_counters: copy failed");
    return;
  }
  static_assert(sizeof(raishin_counters) == sizeof h, "counter layout");
  std::memcpy(c, h, sizeof h);
}

void raishin_model_reset_counters(raishin_model* m) {
  if (hipMemsetAsync(m->cnt, 0, NCOUNT * 8, m->st) != hipSuccess) set_error("reset counters failed");
}

void raishin_model_sync(raishin_model* m) {
  if (hipStreamSynchronize(m->st) != hipSuccess) set_error("raishin_model_sync failed");
}

void* raishin_model_stream(raishin_model* m) { return (void*)m->st; }

void raishin_model_debug_stage(raishin_model* m, int stage, double* ww, double* wpre) {
  // stage 1 also runs the CFL reduction and time advance of a step
  StageTap tap;
  tap.ww = ww;
  tap.wpre = wpre;
  auto go = [&]() -> bool {
    if (stage == 1) {
      if (!run_cfl(m)) return false;
      HIPCHK(launch_advance_time(m->dtbits, m->p.cfl, m->tdt, m->st));
      if (m->need_wo) HIPCHK(hipMemcpyAsync(m->Wo, m->W, 9 * m->G.N * 8, hipMemcpyDeviceToDevice, m->st));
      return run_stage(m, 1, m->Ucur, nullptr, m->Uh, &tap);
    }
    if (stage == 2) {
      if (m->p.irkt == 2) return run_stage(m, 2, m->Ucur, m->Uh, m->Unext, &tap);
      return run_stage(m, 3, m->Ucur, m->Uh, m->Us, &tap);
    }
    return run_stage(m, 4, m->Ucur, m->Us, m->Unext, &tap);
  };
  if (go() && stage == m->p.irkt) std::swap(m->Ucur, m->Unext);
}

void raishin_model_set_timing(raishin_model* m, int on) {
  resolve_timing(m);
  for (int k = 0; k < K_NKINDS; k++) { m->ev_ms[k] = 0.0; m->ev_n[k] = 0; }
  m->timing = on != 0;
}

int raishin_model_get_timing(raishin_model* m, double* ms, long long* launches, int nkinds) {
  resolve_timing(m);
  const int n = nkinds < K_NKINDS ? nkinds : K_NKINDS;
  for (int k = 0; k < n; k++) {
    if (ms) ms[k] = m->ev_ms[k];
    if (launches) launches[k] = m->ev_n[k];
  }
  return K_NKINDS;
}

void raishin_init(const raishin_pram* p) {
  g_pram = *p;
  g_pram_set = true;
}

void raishin_cdtcfl4(const double* uri, const double* gcov, const double* gcon, double* dtcfl, const double* dx1,
                     const double* dx2, const double* dx3, double* dt1, double* dt2, double* dt3, const int* nm1,
                     const int* is1, const int* ie1, const int* js1, const int* je1, const int* ks1, const int* ke1) {
  if (!g_pram_set) { set_error("raishin_cdtcfl4: raishin_init not called"); return; }
  raishin_pram p = g_pram;
  if (*is1 != 1 || *js1 != 1 || *ks1 != 1 || *ie1 != p.imax || *je1 != p.jmax || *ke1 != p.kmax) {
    set_error("raishin_cdtcfl4: one rank spans the grid (is1=1..ie1=imax)");
    return;
  }
  if (*nm1 != ghost_nm0(p.irec) + 1) { set_error("raishin_cdtcfl4: nm1 does not match irec/ict"); return; }
  // A model without christ/face metrics: reuse the centre arrays for every location.
  std::vector<double> christ((size_t)64 * p.imax * p.jmax * p.kmax, 0.0);
  std::vector<double> detg((size_t)p.imax * p.jmax * p.kmax, 1.0);
  raishin_model* m = raishin_model_create(&p, dx1, dx2, dx3, dx1, dx2, dx3, gcov, gcon, gcov, gcon, gcov, gcon, gcov,
                                          gcon, detg.data(), detg.data(), detg.data(), detg.data(), christ.data());
  if (!m) return;
  raishin_model_set_state(m, uri, uri, 0.0);
  raishin_model_cdtcfl4(m, dtcfl, dt1, dt2, dt3);
  raishin_model_destroy(m);
}

}  // extern "C"
