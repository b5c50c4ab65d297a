// C-ABI of the MI355X-native RAISHIN GRMHD step: resident models (the
// production path).  The reference-named entry points are in rh_ref.cpp.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "rh_model.hpp"

using namespace rh;
using namespace rhm;

namespace rhm {
extern std::mutex g_err_mu;
extern std::string g_err;
}

#define HIPCHK_VOID(x)                                                              \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) {                                                         \
      ::rhm::set_error("%s:%d %s: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
      return;                                                                       \
    }                                                                               \
  } while (0)

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

static raishin_model* create_model(const raishin_pram* p, const double* x1, const double* x2, const double* x3,
                                   const double* dx1b, const double* dx2b, const double* dx3b, const double* gcov,
                                   const double* gcon, const double* gcovi, const double* gconi, const double* gcovj,
                                   const double* gconj, const double* gcovk, const double* gconk, const double* detg,
                                   const double* detgi, const double* detgj, const double* detgk, const double* christ,
                                   int metric_rows) {
  if (!p || !x1 || !x2 || !x3 || !gcov || !gcon || !christ) { set_error("raishin_model_create: null argument"); return nullptr; }
  raishin_model* m = new raishin_model();
  const double* gcv[4] = {gcov, gcovi, gcovj, gcovk};
  const double* gcn[4] = {gcon, gconi, gconj, gconk};
  const double* dtg[4] = {detg, detgi, detgj, detgk};
  if (!build_model(m, p, x1, x2, x3, dx1b, dx2b, dx3b, gcv, gcn, dtg, christ, metric_rows)) {
    dfree_all(m);
    if (m->st) (void)hipStreamDestroy(m->st);
    delete m;
    return nullptr;
  }
  return m;
}

raishin_model* raishin_model_create(const raishin_pram* p, const double* x1, const double* x2, const double* x3,
                                    const double* dx1b, const double* dx2b, const double* dx3b, const double* gcov,
                                    const double* gcon, const double* gcovi, const double* gconi, const double* gcovj,
                                    const double* gconj, const double* gcovk, const double* gconk, const double* detg,
                                    const double* detgi, const double* detgj, const double* detgk,
                                    const double* christ) {
  return create_model(p, x1, x2, x3, dx1b, dx2b, dx3b, gcov, gcon, gcovi, gconi, gcovj, gconj, gcovk, gconk, detg,
                      detgi, detgj, detgk, christ, p ? p->jmax : 0);
}

raishin_model* raishin_model_create_ik(const raishin_pram* p, const double* x1, const double* x2, const double* x3,
                                       const double* dx1b, const double* dx2b, const double* dx3b, const double* gcov,
                                       const double* gcon, const double* gcovi, const double* gconi,
                                       const double* gcovj, const double* gconj, const double* gcovk,
                                       const double* gconk, const double* detg, const double* detgi,
                                       const double* detgj, const double* detgk, const double* christ) {
  return create_model(p, x1, x2, x3, dx1b, dx2b, dx3b, gcov, gcon, gcovi, gconi, gcovj, gconj, gcovk, gconk, detg,
                      detgi, detgj, detgk, christ, 1);
}

raishin_model* raishin_model_create_grid(const raishin_pram* p, const double* x1, const double* x2,
                                         const double* x3, const double* dx1b, const double* dx2b,
                                         const double* dx3b, const double* x1a, const double* x3a) {
  if (!p || !x1 || !x2 || !x3 || !x1a || !x3a) { set_error("raishin_model_create_grid: null argument"); return nullptr; }
  raishin_model* m = new raishin_model();
  const double* none[4] = {nullptr, nullptr, nullptr, nullptr};
  if (!build_model(m, p, x1, x2, x3, dx1b, dx2b, dx3b, nullptr, none, none, nullptr, 1, x1a, x3a)) {
    dfree_all(m);
    if (m->st) (void)hipStreamDestroy(m->st);
    delete m;
    return nullptr;
  }
  return m;
}

int raishin_model_tables(raishin_model* m, int which, double* out) {
  if (!m || !out || which < 0 || which > 4) { set_error("raishin_model_tables: bad argument"); return -1; }
  const double* src = which < 4 ? m->met[which] : m->chr;
  const size_t n = (size_t)(which < 4 ? 33 : 64) * (size_t)m->G.P2;
  if (hipMemcpyAsync(out, src, n * 8, hipMemcpyDeviceToHost, m->st) != hipSuccess ||
      hipStreamSynchronize(m->st) != hipSuccess) {
    set_error("raishin_model_tables: copy failed");
    return -1;
  }
  return 0;
}

void raishin_model_destroy(raishin_model* m) {
  if (!m) return;
  (void)hipStreamSynchronize(m->st);
  if (m->gexec) (void)hipGraphExecDestroy(m->gexec);
  if (m->graph) (void)hipGraphDestroy(m->graph);
  for (int k = 0; k < K_NKINDS; k++)
    for (auto& pr : m->ev[k]) { (void)hipEventDestroy(pr.first); (void)hipEventDestroy(pr.second); }
  dfree_all(m);
  (void)hipStreamDestroy(m->st);
  delete m;
}

void raishin_model_set_state(raishin_model* m, const double* uri, const double* uu, double time) {
  (void)rhm::model_set_state(m, uri, uu, time);
}

void raishin_model_get_state(raishin_model* m, double* uri, double* uu, double* time) {
  (void)rhm::model_get_state(m, uri, uu, time);
}

}  // extern "C"

namespace rhm {

// raishin_model_set_state / get_state with their outcome: false when this
// call failed (the error is set), whatever error an earlier call left.
bool model_set_state(raishin_model* m, const double* uri, const double* uu, double time) {
  // one stored phi row when every row of the state is the same (and stays
  // so: periodic phi with one interior cell, r and theta not periodic)
  const int j0 = m->G.nm1 + 1;
  const bool collapse = collapsible(m) && rows_equal(m->G, uri, j0) && (!uu || rows_equal(m->G, uu, j0));
  HIPCHK(hipStreamSynchronize(m->st));
  set_rows(m, collapse);
  const long long N = m->G.N;
  std::vector<double> packed;
  auto src = [&](const double* full) -> const double* {
    if (!m->G.jfix) return full;
    pack_row(m->G, full, packed);
    return packed.data();
  };
  auto go = [&]() -> bool {
    if (!m->stage_aos && !dalloc(&m->stage_aos, 9 * N)) return false;
    if (!upload(m->stage_aos, src(uri), 9 * N, m->st)) return false;
    HIPCHK(launch_aos_to_soa(m->stage_aos, m->W, N, m->st));
    HIPCHK(hipStreamSynchronize(m->st));  // the packed row is reused below
    if (uu) {
      if (!upload(m->stage_aos, src(uu), 9 * N, m->st)) return false;
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
  return go();
}

bool model_get_state(raishin_model* m, double* uri, double* uu, double* time) {
  const long long N = m->G.N;
  std::vector<double> packed;
  // a collapsed model's row goes to every phi row of the host array
  auto fetch = [&](const double* planes, double* full) -> bool {
    HIPCHK(launch_soa_to_aos(planes, m->stage_aos, N, m->st));
    if (!m->G.jfix) {
      HIPCHK(hipMemcpyAsync(full, m->stage_aos, 9 * N * 8, hipMemcpyDeviceToHost, m->st));
      HIPCHK(hipStreamSynchronize(m->st));
      return true;
    }
    packed.resize(9 * (size_t)N);
    HIPCHK(hipMemcpyAsync(packed.data(), m->stage_aos, 9 * N * 8, hipMemcpyDeviceToHost, m->st));
    HIPCHK(hipStreamSynchronize(m->st));
    unpack_row(m->G, packed, full);
    return true;
  };
  auto go = [&]() -> bool {
    if (!m->stage_aos && !dalloc(&m->stage_aos, 9 * N)) return false;
    if (uri && !fetch(m->W, uri)) return false;
    if (uu && !fetch(m->Ucur, uu)) return false;
    if (time) {
      double tdt[2];
      HIPCHK(hipMemcpyAsync(tdt, m->tdt, 16, hipMemcpyDeviceToHost, m->st));
      HIPCHK(hipStreamSynchronize(m->st));
      *time = tdt[0];
    }
    return true;
  };
  return go();
}

}  // namespace rhm

extern "C" {

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

}  // extern "C"

namespace {

// Capture two steps into a graph.  Each stage swaps the primitive buffers and
// each step swaps the conserved buffers once on the host; after two steps
// every pointer is back where it started, so the graph replays.  An eager
// step in between leaves the conserved buffers (and, for RK3, the primitive
// buffers) swapped, and the graph is captured again.
bool capture_pair(raishin_model* m) {
  if (m->gexec) { (void)hipGraphExecDestroy(m->gexec); m->gexec = nullptr; }
  if (m->graph) { (void)hipGraphDestroy(m->graph); m->graph = nullptr; }
  const unsigned timing = m->timing;
  m->timing = 0;
  HIPCHK(hipStreamBeginCapture(m->st, hipStreamCaptureModeThreadLocal));
  const double* W0 = m->W;
  const double* U0 = m->Ucur;
  const bool ok = run_step(m) && run_step(m);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(m->st, &g);
  m->timing = timing;
  if (!ok || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    if (e != hipSuccess) set_error("raishin_model_step_graph: capture failed: %s", hipGetErrorString(e));
    return false;
  }
  if (m->W != W0 || m->Ucur != U0) {
    (void)hipGraphDestroy(g);
    set_error("raishin_model_step_graph: buffers did not return");
    return false;
  }
  HIPCHK(hipGraphInstantiate(&m->gexec, g, nullptr, nullptr, 0));
  m->graph = g;
  m->graph_W = m->W;
  m->graph_U = m->Ucur;
  m->graph_mixed = m->mixed;
  return true;
}

bool read_time(raishin_model* m, double* time, double* dt_last, const char* who) {
  if (!time && !dt_last) return true;
  double tdt[2] = {0.0, 0.0};
  if (hipMemcpyAsync(tdt, m->tdt, 16, hipMemcpyDeviceToHost, m->st) != hipSuccess ||
      hipStreamSynchronize(m->st) != hipSuccess) {
    set_error("%s: synchronisation failed", who);
    return false;
  }
  if (time) *time = tdt[0];
  if (dt_last) *dt_last = tdt[1];
  return true;
}

// The flux buffers keep the last fluxes written (a sweep reads its slot back
// where the wave speeds are unordered), and the mixed form stores them as
// float: a change of form converts them in place, nine planes at a time,
// staged through the recovered-primitive buffer (every stage rewrites all of
// it before reading it, so it holds nothing between steps).  Narrowing walks
// the chunks upwards and widening downwards, so a chunk's output overlaps
// only chunks already staged.  Everything runs in order on the model's
// stream; a batch that steps the model orders the member's stream after its
// own work, so no device-wide wait and no allocation is needed.
bool retype_fluxes(raishin_model* m, bool to_float) {
  if (!m->F || !m->CT) return true;
  const long long N = m->G.N;
  const size_t es = to_float ? sizeof(double) : sizeof(float);  // element size as stored now
  const size_t ds = to_float ? sizeof(float) : sizeof(double);  // element size after
  const struct { double* p; int planes; } bufs[] = {{m->F, 27}, {m->CT, 6}};
  for (const auto& b : bufs) {
    const int nch = (b.planes + 8) / 9;
    for (int q = 0; q < nch; q++) {
      const int c = to_float ? q : nch - 1 - q;
      const long long p0 = 9LL * c, np = std::min(9, b.planes - 9 * c), n = np * N;
      char* base = reinterpret_cast<char*>(b.p);
      HIPCHK(hipMemcpyAsync(m->Wpre, base + (size_t)p0 * N * es, (size_t)n * es, hipMemcpyDeviceToDevice, m->st));
      HIPCHK(launch_retype(m->Wpre, base + (size_t)p0 * N * ds, n, to_float, m->st));
    }
  }
  return true;
}

bool graph_ready(raishin_model* m) {
  if (m->gexec && m->graph_W == m->W && m->graph_U == m->Ucur && m->graph_mixed == m->mixed) return true;
  return capture_pair(m);
}

}  // namespace

extern "C" {

void raishin_model_step_graph(raishin_model* m, int npairs, double* time, double* dt_last) {
  if (!graph_ready(m)) return;
  for (int q = 0; q < npairs; q++)
    if (hipGraphLaunch(m->gexec, m->st) != hipSuccess) { set_error("raishin_model_step_graph: launch failed"); return; }
  (void)read_time(m, time, dt_last, "raishin_model_step_graph");
}

void raishin_models_step_graph(raishin_model* const* ms, int nmodels, int npairs) {
  for (int a = 0; a < nmodels; a++)
    if (!graph_ready(ms[a])) return;
  // round robin: every model's next graph is queued before any model's second
  for (int q = 0; q < npairs; q++)
    for (int a = 0; a < nmodels; a++)
      if (hipGraphLaunch(ms[a]->gexec, ms[a]->st) != hipSuccess) {
        set_error("raishin_models_step_graph: launch failed");
        return;
      }
}

void raishin_model_cdtcfl4(raishin_model* m, double* dtcfl, double* dt1, double* dt2, double* dt3) {
  (void)cdtcfl4_on(m, m->W, dtcfl, dt1, dt2, dt3);
}

void raishin_model_counters(raishin_model* m, raishin_counters* c) {
  unsigned long long h[NCOUNT];
  if (launch_sum_counters(m->cpart, m->crows, m->cnt, m->st) != hipSuccess ||
      hipMemcpyAsync(h, m->cnt, sizeof h, hipMemcpyDeviceToHost, m->st) != hipSuccess ||
      hipStreamSynchronize(m->st) != hipSuccess) {
    set_error("raishin_model_counters: copy failed");
    return;
  }
  static_assert(sizeof(raishin_counters) == sizeof h, "counter layout");
  std::memcpy(c, h, sizeof h);
}

void raishin_model_reset_counters(raishin_model* m) {
  if (hipMemsetAsync(m->cnt, 0, NCOUNT * 8, m->st) != hipSuccess ||
      hipMemsetAsync(m->cpart, 0, (size_t)m->crows * NCOUNT * 8, m->st) != hipSuccess)
    set_error("reset counters failed");
}

void raishin_model_sync(raishin_model* m) {
  if (hipStreamSynchronize(m->st) != hipSuccess) set_error("raishin_model_sync failed");
}

void* raishin_model_stream(raishin_model* m) { return (void*)m->st; }

long long raishin_model_device_bytes(raishin_model* m) { return m ? (long long)m->dev_bytes : 0; }

void raishin_model_get_pram(raishin_model* m, raishin_pram* p) {
  if (m && p) *p = m->p;
}

void raishin_model_allow_collapse(raishin_model* m, int allow) {
  if (m) m->allow_collapse = allow != 0;
}

int raishin_model_stored_rows(raishin_model* m) { return m ? m->G.njs : 0; }

void raishin_model_debug_stage(raishin_model* m, int stage, double* ww, double* wpre) {
  // stage 1 also runs the CFL reduction and time advance of a step
  StageTap tap;
  tap.ww = ww;
  tap.wpre = wpre;
  auto go = [&]() -> bool {
    if (stage == 1) {
      if (!run_cfl(m, m->W)) return false;
      HIPCHK(launch_advance_time(m->G, m->dtbits, m->p.cfl, m->tdt, m->st));
      if (m->need_wo) HIPCHK(launch_copy(m->G, m->Wo, m->W, 9 * m->G.N, m->st));
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

void raishin_model_debug_fluxes(raishin_model* m, double* ww, double* ct) {
  (void)copy_fluxes_out(m, ww, ct);
}

void raishin_model_set_precision(raishin_model* m, int mode) {
  if (mode != 0 && mode != 1) { set_error("raishin_model_set_precision: mode %d (0 FP64, 1 mixed)", mode); return; }
  const bool mixed = mode == 1;
  if (mixed != m->mixed && !retype_fluxes(m, mixed)) return;
  m->mixed = mixed;
}

void raishin_model_set_timing(raishin_model* m, int on) {
  raishin_model_set_timing_mask(m, on ? (1u << K_NKINDS) - 1u : 0u);
}

void raishin_model_set_timing_mask(raishin_model* m, unsigned mask) {
  resolve_timing(m);
  for (int k = 0; k < K_NKINDS; k++) { m->ev_ms[k] = 0.0; m->ev_n[k] = 0; }
  m->timing = mask & ((1u << K_NKINDS) - 1u);
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

int raishin_model_set_block_sizes(raishin_model* m, const int* bs, int n) {
  if (!m || !bs || n != TS_N) { set_error("raishin_model_set_block_sizes: expected %d sizes", (int)TS_N); return -1; }
  for (int q = 0; q < TS_N; q++)
    if (bs[q] != 0 && (bs[q] < 64 || bs[q] > 256 || bs[q] % 64)) {
      set_error("raishin_model_set_block_sizes: size %d of slot %d is not 0 or 64, 128, 192, 256", bs[q], q);
      return -1;
    }
  for (int q = 0; q < TS_N; q++) m->G.bs[q] = (unsigned short)bs[q];
  m->geo_gen++;
  // a captured graph holds the old launch shapes
  if (m->gexec) { (void)hipGraphExecDestroy(m->gexec); m->gexec = nullptr; }
  if (m->graph) { (void)hipGraphDestroy(m->graph); m->graph = nullptr; }
  return 0;
}

int raishin_model_get_block_sizes(raishin_model* m, int* bs, int n) {
  if (!m || !bs || n != TS_N) { set_error("raishin_model_get_block_sizes: expected %d sizes", (int)TS_N); return -1; }
  for (int q = 0; q < TS_N; q++) bs[q] = slot_threads(m->G, q);
  return 0;
}

int raishin_model_set_sweep_tiles(raishin_model* m, int split, int seg_j, int seg_k) {
  if (!m || (split != 0 && split != 1) || (seg_j != 0 && !psweep_seg_ok(seg_j)) ||
      (seg_k != 0 && !psweep_seg_ok(seg_k))) {
    set_error("raishin_model_set_sweep_tiles: split must be 0 or 1 and seg_j, seg_k 0 or a power of two "
              "from 2 to 64");
    return -1;
  }
  m->G.ppm_split = (unsigned short)split;
  m->G.seg[0] = (unsigned short)seg_j;
  m->G.seg[1] = (unsigned short)seg_k;
  m->geo_gen++;
  if (m->gexec) { (void)hipGraphExecDestroy(m->gexec); m->gexec = nullptr; }
  if (m->graph) { (void)hipGraphDestroy(m->graph); m->graph = nullptr; }
  return 0;
}

int raishin_model_set_divb_form(raishin_model* m, int form) {
  if (!m || (form != 0 && form != 1)) {
    set_error("raishin_model_set_divb_form: form must be 0 (the reference's fluxes) or 1 (zero normal-B flux)");
    return -1;
  }
  m->G.divb0 = (unsigned short)form;
  m->geo_gen++;
  if (m->gexec) { (void)hipGraphExecDestroy(m->gexec); m->gexec = nullptr; }
  if (m->graph) { (void)hipGraphDestroy(m->graph); m->graph = nullptr; }
  return 0;
}

int raishin_model_get_sweep_tiles(raishin_model* m, int* split, int* seg_j, int* seg_k) {
  if (!m) { set_error("raishin_model_get_sweep_tiles: null model"); return -1; }
  if (split) *split = m->G.ppm_split;
  if (seg_j) *seg_j = psweep_segment(m->G, 2);
  if (seg_k) *seg_k = psweep_segment(m->G, 3);
  return 0;
}

namespace {
// Milliseconds per step of each timed kind over `steps` steps from the state
// (after one warm-up step); false on failure.
bool time_kinds(raishin_model* m, const double* uri, const double* uu, double time, int steps, double* ms) {
  raishin_model_set_state(m, uri, uu, time);
  if (!run_step(m)) return false;
  raishin_model_set_timing(m, 1);
  for (int s = 0; s < steps; s++)
    if (!run_step(m)) return false;
  long long cnt[K_NKINDS];
  raishin_model_get_timing(m, ms, cnt, K_NKINDS);
  raishin_model_set_timing(m, 0);
  for (int k = 0; k < K_NKINDS; k++) ms[k] /= steps;
  return true;
}
}  // namespace

int raishin_model_autotune(raishin_model* m, const double* uri, const double* uu, double time, int steps) {
  if (!m || !uri || steps < 1) { set_error("raishin_model_autotune: bad arguments"); return -1; }
  // timed kind of each tuned slot
  const int kind_of[TS_N] = {K_EDGES, K_SWEEP1, K_SWEEP2, K_SWEEP3, K_FLUXCT, K_UPDATE, K_RECOV, K_POST};
  double best_ms[TS_N];
  int best[TS_N];
  for (int q = 0; q < TS_N; q++) { best_ms[q] = 1e300; best[q] = 256; }
  const unsigned timing = m->timing;
  for (int cand : TUNE_CANDIDATES) {
    int bs[TS_N];
    for (int q = 0; q < TS_N; q++) bs[q] = cand;
    if (raishin_model_set_block_sizes(m, bs, TS_N) != 0) return -1;
    raishin_model_set_state(m, uri, uu, time);
    if (!run_step(m)) return -1;  // warm-up
    raishin_model_set_timing(m, 1);
    for (int s = 0; s < steps; s++)
      if (!run_step(m)) return -1;
    double ms[K_NKINDS];
    long long cnt[K_NKINDS];
    raishin_model_get_timing(m, ms, cnt, K_NKINDS);
    raishin_model_set_timing(m, 0);
    for (int q = 0; q < TS_N; q++) {
      const int k = kind_of[q];
      const double avg = cnt[k] ? ms[k] / (double)cnt[k] : 1e300;
      if (avg < best_ms[q]) { best_ms[q] = avg; best[q] = cand; }
    }
  }
  if (raishin_model_set_block_sizes(m, best, TS_N) != 0) return -1;
  // PPM sweep tiles: the fused sweeps with each tile height along j and k,
  // the fastest per direction (the split form, slower at every measured
  // size, is left to explicit choice: members of one batch must agree on it)
  if (m->p.irec == 5) {
    const int segs[] = {8, 16, 32};
    double ms[K_NKINDS];
    double bj = 1e300, bk = 1e300;
    int sj = PSWEEP_SEG_DEFAULT, sk = PSWEEP_SEG_DEFAULT;
    for (int sgm : segs) {
      if (raishin_model_set_sweep_tiles(m, 0, sgm, sgm) != 0 || !time_kinds(m, uri, uu, time, steps, ms)) return -1;
      if (ms[K_SWEEP2] < bj) { bj = ms[K_SWEEP2]; sj = sgm; }
      if (ms[K_SWEEP3] < bk) { bk = ms[K_SWEEP3]; sk = sgm; }
    }
    if (raishin_model_set_sweep_tiles(m, 0, sj, sk) != 0) return -1;
  }
  // start again from the given state, as if no step had been taken
  raishin_model_set_state(m, uri, uu, time);
  raishin_model_reset_counters(m);
  raishin_model_set_timing_mask(m, timing);
  return 0;
}

}  // extern "C"
