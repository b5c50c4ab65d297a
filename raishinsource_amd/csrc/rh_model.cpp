// Resident model internals: set-up of device buffers and (i,k) tables, and the
// pieces of an RK stage.  Host code; kernels live in rh_kernels.hip.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "rh_model.hpp"

using namespace rh;

namespace rhm {

std::mutex g_err_mu;
thread_local size_t g_dalloc_bytes = 0;
std::string g_err;

void set_error(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  std::lock_guard<std::mutex> lk(g_err_mu);
  if (g_err.empty()) g_err = buf;  // sticky: keep the first error
  (void)hipGetLastError();         // recorded: later launches on this thread report their own errors
}



int ghost_nm0(int irec) {
  // main.f90:179-211
  if (irec == 1 || irec == 2 || irec == 3 || irec == 11 || irec == 12 || irec == 13) return 2;
  if (irec == 4 || irec == 6 || irec == 7 || irec == 8 || irec == 9 || irec == 10) return 3;
  if (irec == 5) return 4;
  return 1;
}

bool check_pram(const raishin_pram& p) {
  if (p.imax < 1 || p.jmax < 1 || p.kmax < 1) { set_error("bad grid %d %d %d", p.imax, p.jmax, p.kmax); return false; }
  if (p.irec != 1 && p.irec != 5 && p.irec != 6 && p.irec != 7) {
    set_error("irec=%d not built (1 MC, 5 PPM, 6 MP5, 7 WENO5)", p.irec);
    return false;
  }
  if (p.irkt != 2 && p.irkt != 3) { set_error("irkt=%d not supported", p.irkt); return false; }
  if (p.ict != 1) { set_error("ict=%d not built (1 = flux-CT)", p.ict); return false; }
  if (p.ihll != 1) { set_error("ihll=%d not built (1 = HLLE)", p.ihll); return false; }
  if (p.icha < 0 || p.icha > 2) {
    set_error("icha=%d not built (0 Gammie, 1 Leismann, 2 Del Zanna speeds)", p.icha);
    return false;
  }
  if (p.iwvec != 1 && p.iwvec != 2) { set_error("iwvec=%d not built (1 recov1a, 2 recov2)", p.iwvec); return false; }
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


void dfree_all(raishin_model* m) {
  double** bufs[] = {&m->W, &m->Wb, &m->Wpre, &m->Wo, &m->Ucur, &m->Uh, &m->Us, &m->Unext, &m->F, &m->CT, &m->E,
                     &m->wgout, &m->stage_aos, &m->chr, &m->r15, &m->r25, &m->tdt};
  for (auto b : bufs) if (*b) { (void)hipFree(*b); *b = nullptr; }
  for (int a = 0; a < 4; a++) if (m->met[a]) { (void)hipFree(m->met[a]); m->met[a] = nullptr; }
  for (int a = 0; a < 3; a++) {
    if (m->ppmc[a]) (void)hipFree(m->ppmc[a]);
    m->ppmc[a] = nullptr;
    if (m->xs[a]) (void)hipFree(m->xs[a]);
    if (m->dxb[a]) (void)hipFree(m->dxb[a]);
    if (m->dxs[a]) (void)hipFree(m->dxs[a]);
    if (m->xa[a]) (void)hipFree(m->xa[a]);
    m->xs[a] = m->dxb[a] = m->dxs[a] = m->xa[a] = nullptr;
  }
  if (m->bc) (void)hipFree(m->bc);
  if (m->dtbits) (void)hipFree(m->dtbits);
  if (m->cnt) (void)hipFree(m->cnt);
  if (m->cpart) (void)hipFree(m->cpart);
  m->cpart = nullptr;
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
// The reference arrays are checked to be independent of j, bit for bit; nj
// is the number of phi rows the arrays hold (jmax, or 1 for one row).
bool metric_tables(const raishin_pram& p, long long nj, const double* gcov, const double* gcon, const double* detg,
                   std::vector<double>& out, const char* what) {
  const long long ni = p.imax, nk = p.kmax, P2 = ni * nk;
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
    if (!((m->timing >> kind) & 1u)) return;
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

// Host-to-device copy ordered on the model's stream, complete on return.  No
// copy here goes through the legacy default stream: models are built on
// worker threads while other streams may be capturing a graph.
bool h2d_sync(void* dst, const void* src, size_t bytes, hipStream_t st) {
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  return true;
}

// interp's coefficients that depend only on the cell widths (rec.f90:1353-1355,
// 1427-1436), per index q of the dummy width array ds (n+2 entries), in the
// reference's expressions and evaluation order: planes 0..2 t1a,t1b,t1c of
// dma at q; planes 3..6 of equation (51) at q+1/2: dB/(dB+dC), t3a*t3b*t3c,
// t3a*dB*(...), t3a*dC*(...) with (dA,dB,dC,dD) = ds(q-1..q+2).
std::vector<double> ppm_coefficients(const std::vector<double>& ds) {
  const long long n2 = (long long)ds.size();
  std::vector<double> pc(7 * n2, 0.0);
  for (long long q = 1; q + 1 < n2; q++) {
    const double dl = ds[q - 1], dc = ds[q], dr = ds[q + 1];
    pc[q] = dc / (dl + dc + dr);
    pc[n2 + q] = (2.0 * dl + dc) / (dr + dc);
    pc[2 * n2 + q] = (dc + 2.0 * dr) / (dl + dc);
  }
  for (long long q = 1; q + 2 < n2; q++) {
    const double dA = ds[q - 1], dB = ds[q], dC = ds[q + 1], dD = ds[q + 2];
    const double t3a = 1.0 / (dA + dB + dC + dD);
    const double t3b = (2.0 * dC * dB) / (dB + dC);
    const double t3c = ((dA + dB) / (2.0 * dB + dC)) - ((dD + dC) / (2.0 * dC + dB));
    pc[3 * n2 + q] = dB / (dB + dC);
    pc[4 * n2 + q] = t3a * t3b * t3c;
    pc[5 * n2 + q] = t3a * dB * ((dA + dB) / (2.0 * dB + dC));
    pc[6 * n2 + q] = t3a * dC * ((dC + dD) / (dB + 2.0 * dC));
  }
  return pc;
}

static bool build_model_(raishin_model* m, const raishin_pram* p, const double* x1, const double* x2,
                         const double* x3, const double* dx1b, const double* dx2b, const double* dx3b,
                         const double* const gcv[4], const double* const gcn[4], const double* const dtg[4],
                         const double* christ, int metric_rows, const double* x1a, const double* x3a);

bool build_model(raishin_model* m, const raishin_pram* p, const double* x1, const double* x2, const double* x3,
                 const double* dx1b, const double* dx2b, const double* dx3b, const double* const gcv[4],
                 const double* const gcn[4], const double* const dtg[4], const double* christ, int metric_rows,
                 const double* x1a, const double* x3a) {
  const size_t bytes0 = g_dalloc_bytes;
  const bool ok =
      build_model_(m, p, x1, x2, x3, dx1b, dx2b, dx3b, gcv, gcn, dtg, christ, metric_rows, x1a, x3a);
  m->dev_bytes = g_dalloc_bytes - bytes0;
  return ok;
}

static bool build_model_(raishin_model* m, const raishin_pram* p, const double* x1, const double* x2,
                         const double* x3, const double* dx1b, const double* dx2b, const double* dx3b,
                         const double* const gcv[4], const double* const gcn[4], const double* const dtg[4],
                         const double* christ, int metric_rows, const double* x1a,
                         const double* x3a) {
  m->p = *p;
  if (!check_pram(*p)) return false;
  Geo& G = m->G;
  G.ni = p->imax; G.nj = p->jmax; G.nk = p->kmax;
  G.nm0 = ghost_nm0(p->irec); G.nm1 = G.nm0 + 1;
  G.icha = p->icha;
  G.iwvec = p->iwvec;
  G.N = (long long)G.ni * G.nj * G.nk;
  G.P2 = (long long)G.ni * G.nk;
  G.njs = G.nj;
  G.jfix = 0;
  const long long N = G.N;
  HIPCHK(hipGetDevice(&m->dev));
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
      !dalloc(&m->carry, 2 + m->carry_cap) || !dalloc(&m->tdt, 2) || !dalloc(&m->cflag, N) ||
      !dalloc(&m->pend, m->carry_cap) || !dalloc(&m->cfl_last, 4 * cfl_blocks(G)) ||
      !dalloc(&m->cfl_min, 4 * cfl_blocks(G)))
    return false;
  if (m->need_wo && !dalloc(&m->Wo, 9 * N)) return false;
  if ((p->irec == 5 || p->irec == 1) && !dalloc(&m->E, 18 * N)) return false;  // face values
  if (!dalloc(&m->stage_aos, 9 * N)) return false;            // set_state / get_state staging
  m->crows = ((N + 255) / 256) * 4;  // wavefronts of the largest launch (every cell, 256 per workgroup)
  if (!dalloc(&m->cpart, (size_t)m->crows * NCOUNT)) return false;
  HIPCHK(hipMemsetAsync(m->cpart, 0, (size_t)m->crows * NCOUNT * 8, m->st));
  HIPCHK(hipMemsetAsync(m->cnt, 0, NCOUNT * 8, m->st));
  HIPCHK(hipMemsetAsync(m->carry, 0, (2 + m->carry_cap) * 8, m->st));
  HIPCHK(hipMemsetAsync(m->tdt, 0, 16, m->st));
  HIPCHK(hipMemsetAsync(m->W, 0, 9 * N * 8, m->st));
  // the reference zeroes its flux array once (main.f90:372-381); an interface
  // whose speeds are unordered (NaN) keeps its previous flux (hll.f90:86-115)
  HIPCHK(hipMemsetAsync(m->F, 0, 27 * N * 8, m->st));
  HIPCHK(hipMemsetAsync(m->CT, 0, 6 * N * 8, m->st));
  HIPCHK(hipMemsetAsync(m->Wpre, 0, 9 * N * 8, m->st));
  HIPCHK(hipMemsetAsync(m->cflag, 0, N, m->st));
  // metric tables
  if (!gcv) {
    // from the grid on the device (k_metric_tables); its inputs are staged in
    // wgout, which the recovery does not use before the first step
    if (p->metric != 303 && p->metric != 1) {
      set_error("metric=%d: the device tables are built for metric 303 and 1", p->metric);
      return false;
    }
    if (!x1a || !x3a) { set_error("raishin_model_create_grid: face positions x1a, x3a missing"); return false; }
    for (int a = 0; a < 4; a++)
      if (!dalloc(&m->met[a], 33 * G.P2)) return false;
    const long long need = 2LL * G.ni + 9LL * G.nk + 1;
    if (need > N) { set_error("grid too small to stage the metric inputs"); return false; }
    std::vector<double> tr(8 * (size_t)G.nk);
    theta_trig_table(*p, x3, x3a, tr.data());
    double* sc = m->wgout;
    int* sym = reinterpret_cast<int*>(sc + 2LL * G.ni + 9LL * G.nk);
    const int one = 1;
    if (!h2d_sync(sc, x1, G.ni * 8, m->st) || !h2d_sync(sc + G.ni, x1a, G.ni * 8, m->st) ||
        !h2d_sync(sc + 2LL * G.ni, x3, G.nk * 8, m->st) || !h2d_sync(sc + 2LL * G.ni + G.nk, tr.data(), tr.size() * 8, m->st) ||
        !h2d_sync(sym, &one, 4, m->st))
      return false;
    HIPCHK(launch_metric_tables(G.ni, G.nk, p->metric, p->akm, sc, sc + G.ni, sc + 2LL * G.ni, x2[0],
                                sc + 2LL * G.ni + G.nk, m->met, m->chr, sym, m->st));
    int symh = 0;
    HIPCHK(hipMemcpyAsync(&symh, sym, 4, hipMemcpyDeviceToHost, m->st));
    HIPCHK(hipStreamSynchronize(m->st));
    m->chr_sym = symh != 0;
  } else {
  const char* names[4] = {"centre", "i-face", "j-face", "k-face"};
  std::vector<double> tab;
  for (int a = 0; a < 4; a++) {
    if (!metric_tables(*p, metric_rows, gcv[a], gcn[a], dtg[a], tab, names[a])) return false;
    if (!dalloc(&m->met[a], tab.size())) return false;
    if (!h2d_sync(m->met[a], tab.data(), tab.size() * 8, m->st)) return false;
  }
  {
    std::vector<double> ct(64 * G.P2);
    const long long mr = metric_rows;
    for (long long k = 0; k < G.nk; k++)
      for (long long j = 0; j < mr; j++)
        for (long long i = 0; i < G.ni; i++) {
          const long long c = i + G.ni * (j + mr * k), q = i + G.ni * k;
          for (int e = 0; e < 64; e++) {
            const double v = christ[64 * c + e];
            if (j == 0) ct[e * G.P2 + q] = v;
            else if (std::memcmp(&v, &ct[e * G.P2 + q], 8)) { set_error("Christoffel symbols depend on j"); return false; }
          }
        }
    if (!h2d_sync(m->chr, ct.data(), ct.size() * 8, m->st)) return false;
    // Gamma^l_{ab} = Gamma^l_{ba} bit for bit lets the source kernel read each pair once
    m->chr_sym = true;
    for (int l = 0; l < 4 && m->chr_sym; l++)
      for (int a = 0; a < 4 && m->chr_sym; a++)
        for (int b = a + 1; b < 4 && m->chr_sym; b++)
          m->chr_sym = std::memcmp(&ct[(size_t)(l + 4 * a + 16 * b) * G.P2], &ct[(size_t)(l + 4 * b + 16 * a) * G.P2],
                                   (size_t)G.P2 * 8) == 0;
  }
  }
  // 1D grids; dxs[d] = rec2's dummy dx*b(d) = actual dx*b(d-1) (rec.f90:58), clamped at the ends
  const double* xsrc[3] = {x1, x2, x3};
  const double* dsrc[3] = {dx1b, dx2b, dx3b};
  const int nn[3] = {G.ni, G.nj, G.nk};
  for (int a = 0; a < 3; a++) {
    const int n = nn[a];
    if (!dalloc(&m->xs[a], n) || !dalloc(&m->dxb[a], n) || !dalloc(&m->dxs[a], n + 2) || !dalloc(&m->xa[a], n))
      return false;
    // x*a(i) = (x*(i) + x*(i+1))/2 as grid.f90 builds it (the last entry is never read)
    std::vector<double> fa(n, 0.0);
    for (int q = 0; q + 1 < n; q++) fa[q] = 0.5 * (xsrc[a][q] + xsrc[a][q + 1]);
    if (n > 1) fa[n - 1] = fa[n - 2] + (xsrc[a][n - 1] - xsrc[a][n - 2]);
    if (!h2d_sync(m->xa[a], fa.data(), n * 8, m->st)) return false;
    std::vector<double> db(n, 0.0), ds(n + 2, 0.0);
    for (int q = 0; q < n - 1; q++) db[q] = dsrc[a][q];
    for (int d = 0; d <= n + 1; d++) {
      int q = d - 1;  // 1-based actual index
      if (q < 1) q = 1;
      if (q > n - 1) q = n - 1;
      ds[d] = (n > 1) ? dsrc[a][q - 1] : 0.0;
    }
    if (!h2d_sync(m->xs[a], xsrc[a], n * 8, m->st)) return false;
    if (!h2d_sync(m->dxb[a], db.data(), n * 8, m->st)) return false;
    if (!h2d_sync(m->dxs[a], ds.data(), (n + 2) * 8, m->st)) return false;
    if (p->irec == 5) {
      const std::vector<double> pc = ppm_coefficients(ds);
      if (!dalloc(&m->ppmc[a], pc.size())) return false;
      if (!h2d_sync(m->ppmc[a], pc.data(), pc.size() * 8, m->st)) return false;
    }
  }
  // r-power tables from the host libm, as the reference computes them (correction.f90:318-319)
  std::vector<double> r15(G.ni), r25(G.ni);
  for (int i = 0; i < G.ni; i++) { r15[i] = std::pow(x1[i], -1.5); r25[i] = std::pow(x1[i], -2.5); }
  if (!h2d_sync(m->r15, r15.data(), G.ni * 8, m->st)) return false;
  if (!h2d_sync(m->r25, r25.data(), G.ni * 8, m->st)) return false;
  int bch[54];
  const int* bsrc[6] = {p->iboux1in, p->iboux1ot, p->iboux2in, p->iboux2ot, p->iboux3in, p->iboux3ot};
  for (int a = 0; a < 6; a++) for (int q = 0; q < 9; q++) bch[a * 9 + q] = bsrc[a][q];
  if (!h2d_sync(m->bc, bch, sizeof bch, m->st)) return false;
  HIPCHK(hipStreamSynchronize(m->st));
  return true;
}

bool copy_planes_out(raishin_model* m, const void* planes, int nplanes, int stride_out, int offset_out,
                     double* host, bool f32 = false) {
  // planes[n*N + c] -> host[offset_out + stride_out*(n + nplanes*c)] via a
  // staging buffer; a collapsed model's row goes to every phi row.  f32: the
  // planes hold floats (the mixed form's fluxes), widened exactly.
  const Geo& G = m->G;
  const long long N = G.N;
  std::vector<double> tmp((size_t)nplanes * N);
  if (f32) {
    std::vector<float> tf((size_t)nplanes * N);
    HIPCHK(hipMemcpyAsync(tf.data(), planes, tf.size() * sizeof(float), hipMemcpyDeviceToHost, m->st));
    HIPCHK(hipStreamSynchronize(m->st));
    for (size_t e = 0; e < tf.size(); e++) tmp[e] = tf[e];
  } else {
    HIPCHK(hipMemcpyAsync(tmp.data(), planes, (size_t)nplanes * N * 8, hipMemcpyDeviceToHost, m->st));
    HIPCHK(hipStreamSynchronize(m->st));
  }
  const long long ni = G.ni, nj = G.nj, nk = G.nk;
  for (long long k = 0; k < nk; k++)
    for (long long j = 0; j < nj; j++)
      for (long long i = 0; i < ni; i++) {
        const long long cs = i + ni * ((G.jfix ? 0 : j) + (long long)G.njs * k);
        const long long cf = i + ni * (j + nj * k);
        for (int n = 0; n < nplanes; n++)
          host[offset_out + (size_t)stride_out * (n + (size_t)nplanes * cf)] = tmp[(size_t)n * N + cs];
      }
  return true;
}

bool collapsible(const raishin_model* m) {
  const raishin_pram& p = m->p;
  const Geo& G = m->G;
  if (!m->allow_collapse || G.nj != 2 * G.nm1 + 1 || m->need_wo) return false;
  for (int q = 0; q < 9; q++) {
    // periodic phi; r and theta boundaries that are not periodic (two
    // periodic axes leave corner ghosts that no pass writes)
    if (p.iboux2in[q] != 1 || p.iboux2ot[q] != 1) return false;
    if (p.iboux1in[q] == 1 || p.iboux1ot[q] == 1 || p.iboux3in[q] == 1 || p.iboux3ot[q] == 1) return false;
  }
  return true;
}

bool rows_equal(const Geo& G, const double* a, int j0) {
  const size_t row = 9 * (size_t)G.ni;
  for (size_t k = 0; k < (size_t)G.nk; k++) {
    const double* r0 = a + row * ((size_t)(j0 - 1) + (size_t)G.nj * k);
    for (size_t j = 0; j < (size_t)G.nj; j++)
      if (std::memcmp(a + row * (j + (size_t)G.nj * k), r0, row * sizeof(double))) return false;
  }
  return true;
}

void set_rows(raishin_model* m, bool collapse) {
  Geo& G = m->G;
  const int jfix = collapse ? G.nm1 + 1 : 0;
  if (G.jfix == jfix) return;
  G.jfix = jfix;
  G.njs = collapse ? 1 : G.nj;
  G.N = (long long)G.ni * G.njs * G.nk;
  m->geo_gen++;
  if (m->gexec) { (void)hipGraphExecDestroy(m->gexec); m->gexec = nullptr; }
  if (m->graph) { (void)hipGraphDestroy(m->graph); m->graph = nullptr; }
}

void pack_row(const Geo& G, const double* full, std::vector<double>& out) {
  const size_t row = 9 * (size_t)G.ni;
  out.resize(row * G.nk);
  for (size_t k = 0; k < (size_t)G.nk; k++)
    std::memcpy(&out[row * k], full + row * ((size_t)(G.jfix - 1) + (size_t)G.nj * k), row * sizeof(double));
}

void unpack_row(const Geo& G, const std::vector<double>& in, double* full) {
  const size_t row = 9 * (size_t)G.ni;
  for (size_t k = 0; k < (size_t)G.nk; k++)
    for (size_t j = 0; j < (size_t)G.nj; j++)
      std::memcpy(full + row * (j + (size_t)G.nj * k), &in[row * k], row * sizeof(double));
}

// Plane p of the flux buffer F: 8-byte elements, or 4-byte ones in the mixed
// form (the buffer keeps its FP64 size; floats use its first half).
static void* flux_plane(const raishin_model* m, double* F, long long p) {
  return m->mixed ? (void*)((float*)F + p * m->G.N) : (void*)(F + p * m->G.N);
}

bool stage_sweeps(raishin_model* m, const double* W, double* F) {
  const Geo& G = m->G;
  const int irec = m->p.irec;
  for (int d = 1; d <= 3; d++) {
    if (irec == 5 && !G.ppm_split) {
      TimedLaunch t(m, K_SWEEP1 + d - 1);
      HIPCHK(launch_psweep(G, d, W, m->met[d], m->dxs[d - 1], m->xs[d - 1], m->ppmc[d - 1], m->p.gam,
                           flux_plane(m, F, (d - 1) * 9), m->CT, m->mixed, m->st));
      continue;
    }
    if (irec == 5 || irec == 1) {
      TimedLaunch t(m, K_EDGES);
      const double* tab = irec == 5 ? m->ppmc[d - 1] : m->xa[d - 1];
      HIPCHK(launch_edges(G, d, irec, W, m->dxs[d - 1], m->xs[d - 1], tab, m->p.gam, m->E, m->mixed, m->st));
    }
    TimedLaunch t(m, K_SWEEP1 + d - 1);
    HIPCHK(launch_sweep(G, d, irec, W, m->met[d], m->E, m->p.gam, flux_plane(m, F, (d - 1) * 9), m->CT, m->mixed,
                        m->st));
  }
  return true;
}

// The mixed form's FP64 copies of flux-CT's six averages: the flux buffer's
// second half, which the float planes leave unused.
static double* ct_fp64(const raishin_model* m, double* F) { return m->mixed ? F + 14 * m->G.N : nullptr; }

bool stage_ct(raishin_model* m, double* F) {
  TimedLaunch t(m, K_FLUXCT);
  HIPCHK(launch_fluxct(m->G, m->CT, flux_plane(m, F, 0), flux_plane(m, F, 9), flux_plane(m, F, 18), m->mixed,
                       ct_fp64(m, F), m->st));
  return true;
}

bool stage_fluxes(raishin_model* m, const double* W, double* F) { return stage_sweeps(m, W, F) && stage_ct(m, F); }

bool stage_update(raishin_model* m, int rkstage, const double* W, const double* Ua, const double* Ub, double* Uout,
                  const double* F, const HostDt* hd) {
  const Geo& G = m->G;
  StageArgs A{};
  A.W = W; A.Ua = Ua; A.Ub = Ub;
  A.F1 = flux_plane(m, const_cast<double*>(F), 0);
  A.F2 = flux_plane(m, const_cast<double*>(F), 9);
  A.F3 = flux_plane(m, const_cast<double*>(F), 18);
  A.f32flux = m->mixed;
  A.CTd = ct_fp64(m, const_cast<double*>(F));
  A.metC = m->met[0]; A.chr = m->chr; A.dx1b = m->dxb[0]; A.dx2b = m->dxb[1]; A.dx3b = m->dxb[2];
  A.r15 = m->r15; A.r25 = m->r25; A.dtbits = m->dtbits; A.Wpre = m->Wpre; A.wgout = m->wgout; A.Ukeep = Uout;
  A.cnt = m->cnt; A.cpart = m->cpart; A.carry = m->carry; A.cflag = m->cflag; A.carry_cap = m->carry_cap;
  A.gam = m->p.gam; A.cfl = m->p.cfl; A.dmin = m->p.dmin; A.pmin = m->p.pmin; A.iter = m->p.iter;
  A.chr_sym = m->chr_sym;
  if (hd) { A.ak1 = hd->ak1; A.ak2 = hd->ak2; A.ak3 = hd->ak3; A.dt_host = hd->dt; }
  { TimedLaunch t(m, K_UPDATE); HIPCHK(launch_update(G, rkstage, A, m->st)); }
  { TimedLaunch t(m, K_RECOV); HIPCHK(launch_recov(G, A, m->st)); }
  FixArgs X{};
  X.W = W; X.Ustage = Uout; X.metC = m->met[0]; X.r15 = m->r15; X.r25 = m->r25; X.Wpre = m->Wpre;
  X.wgout = m->wgout; X.carry = m->carry; X.cnt = m->cnt; X.cflag = m->cflag; X.pend = m->pend;
  X.carry_cap = m->carry_cap; X.gam = m->p.gam;
  X.dmin = m->p.dmin; X.pmin = m->p.pmin; X.iter = m->p.iter;
  { TimedLaunch t(m, K_CARRY); HIPCHK(launch_recov_carry(G, X, m->st)); }
  return true;
}

bool stage_post(raishin_model* m, const double* Wold, const double* Wo, double* Wout, double* Uout) {
  PostArgs P{};
  P.Wpre = m->Wpre; P.Wghost = m->Wpre; P.Wold = Wold; P.Wo = Wo; P.metC = m->met[0]; P.r15 = m->r15;
  P.r25 = m->r25; P.bc = m->bc; P.Wout = Wout; P.Uout = Uout; P.cnt = m->cnt; P.cpart = m->cpart; P.gam = m->p.gam;
  P.dmin = m->p.dmin; P.pmin = m->p.pmin; P.pmax = m->p.pmax; P.model = m->p.model;
  { TimedLaunch t(m, K_POST); HIPCHK(launch_post(m->G, P, m->st)); }
  return true;
}

// One RK stage of the model path: sweeps -> flux-CT -> update (RK + recov) ->
// fix-up -> ghosts/floors/conserved.
bool run_stage(raishin_model* m, int rkstage, const double* Ua, const double* Ub, double* Uout,
               const StageTap* tap) {
  if (!stage_fluxes(m, m->W, m->F)) return false;
  if (tap && tap->ww) {
    // ww(d,n,i,j,k) as after ct()
    for (int d = 0; d < 3; d++)
      if (!copy_planes_out(m, flux_plane(m, m->F, (long long)d * 9), 9, 3, d, tap->ww, m->mixed)) return false;
  }
  if (!stage_update(m, rkstage, m->W, Ua, Ub, Uout, m->F, nullptr)) return false;
  if (tap && tap->wpre && !copy_planes_out(m, m->Wpre, 9, 1, 0, tap->wpre)) return false;
  if (!stage_post(m, m->W, m->need_wo ? m->Wo : m->W, m->Wb, Uout)) return false;
  std::swap(m->W, m->Wb);
  return true;
}

bool copy_fluxes_out(raishin_model* m, double* ww, double* ct) {
  for (int d = 0; ww && d < 3; d++)
    if (!copy_planes_out(m, flux_plane(m, m->F, (long long)d * 9), 9, 3, d, ww, m->mixed)) return false;
  return !ct || copy_planes_out(m, m->CT, 6, 1, 0, ct, m->mixed);
}

bool run_cfl(raishin_model* m, const double* W) {
  CflArgs C{};
  C.W = W; C.metC = m->met[0]; C.dx1b = m->dxb[0]; C.dx2b = m->dxb[1]; C.dx3b = m->dxb[2];
  C.dtbits = m->dtbits; C.blast = m->cfl_last; C.bmin = m->cfl_min; C.gam = m->p.gam;
  { TimedLaunch t(m, K_CFL); HIPCHK(launch_cfl(m->G, C, m->st)); }
  return true;
}

bool cdtcfl4_on(raishin_model* m, const double* W, double* dtcfl, double* dt1, double* dt2, double* dt3) {
  if (!run_cfl(m, W)) return false;
  unsigned long long b[4];
  HIPCHK(hipMemcpyAsync(b, m->dtbits, 32, hipMemcpyDeviceToHost, m->st));
  HIPCHK(hipStreamSynchronize(m->st));
  double v[4];
  std::memcpy(v, b, 32);
  if (dtcfl) *dtcfl = v[0];
  if (dt1) *dt1 = v[1];
  if (dt2) *dt2 = v[2];
  if (dt3) *dt3 = v[3];
  return true;
}

bool run_step(raishin_model* m) {
  const long long N = m->G.N;
  if (!run_cfl(m, m->W)) return false;
  HIPCHK(launch_advance_time(m->G, m->dtbits, m->p.cfl, m->tdt, m->st));
  if (m->need_wo) HIPCHK(launch_copy(m->G, m->Wo, m->W, 9 * N, m->st));
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


}  // namespace rhm
