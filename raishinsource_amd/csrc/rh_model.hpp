// Internals of a resident model, shared by the model C-ABI (rh_capi.cpp) and
// the reference-named entry points (rh_ref.cpp).  Host code.
#pragma once
#include <hip/hip_runtime.h>

#include <utility>
#include <vector>

#include "../../include/raishin_hip.h"
#include "rh_kernels.hpp"

#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) {                                                        \
      ::rhm::set_error("%s:%d %s: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
      return false;                                                                \
    }                                                                              \
  } while (0)

// Kernel kinds timed with HIP events when timing is on (raishin_model_set_timing).
enum { K_SWEEP1 = 0, K_SWEEP2, K_SWEEP3, K_FLUXCT, K_UPDATE, K_CARRY, K_POST, K_CFL, K_EDGES, K_RECOV, K_NKINDS };

struct raishin_model {
  raishin_pram p;
  int dev = 0;  // the device the model was created on
  // per-kind HIP event pairs recorded around launches of the kinds whose bit
  // is set in timing (0: none); resolved on read
  unsigned timing = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[K_NKINDS];
  size_t ev_used[K_NKINDS] = {0};
  double ev_ms[K_NKINDS] = {0};
  long long ev_n[K_NKINDS] = {0};
  rh::Geo G;
  hipStream_t st = nullptr;
  bool need_wo = false;
  bool mixed = false;    // reconstruction + face fluxes in float (raishin_model_set_precision)
  bool chr_sym = false;  // Christoffel tables symmetric in the lower indices, bit for bit
  // state (9 planes each)
  double *W = nullptr, *Wb = nullptr, *Wpre = nullptr, *Wo = nullptr;
  double *Ucur = nullptr, *Uh = nullptr, *Us = nullptr, *Unext = nullptr;
  // F: ww(3,9,...) (27 planes); CT: the six transverse B flux copies ct.f90 averages
  double *F = nullptr, *CT = nullptr, *wgout = nullptr, *stage_aos = nullptr;
  double* E = nullptr;  // face values of one direction (18 planes): PPM (irec 5) and MC (irec 1)
  double* ppmc[3] = {nullptr, nullptr, nullptr};  // PPM width coefficients per direction (irec = 5)
  // geometry
  double *met[4] = {nullptr, nullptr, nullptr, nullptr};  // centre, i-face, j-face, k-face
  double *chr = nullptr;
  double *xs[3] = {nullptr, nullptr, nullptr}, *dxb[3] = {nullptr, nullptr, nullptr}, *dxs[3] = {nullptr, nullptr, nullptr};
  double* xa[3] = {nullptr, nullptr, nullptr};  // cell-face positions x*a (MC reconstruction)
  double *r15 = nullptr, *r25 = nullptr;
  int* bc = nullptr;
  unsigned long long *dtbits = nullptr, *cnt = nullptr, *carry = nullptr;
  unsigned long long* cpart = nullptr;  // per-wavefront event rows, folded into cnt on read
  long long crows = 0;
  unsigned char *cflag = nullptr, *pend = nullptr;
  long long carry_cap = 0;  // room for every interior cell: the flagged list cannot overflow
  long long* cfl_last = nullptr;
  double* cfl_min = nullptr;
  double* tdt = nullptr;  // [0] time [1] last dt
  // two steps captured as a graph (raishin_model_step_graph); valid while the
  // buffer pointers and the precision mode are those it was captured with
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  const double* graph_W = nullptr;
  const double* graph_U = nullptr;
  bool graph_mixed = false;
  size_t dev_bytes = 0;  // device memory allocated by build_model
  // one stored phi row when the state allows it (raishin_model_set_state);
  // geo_gen changes whenever G does, so batch plans can tell
  bool allow_collapse = true;
  unsigned geo_gen = 0;
};

namespace rhm {

constexpr int NCOUNT = 11;

void set_error(const char* fmt, ...);
int ghost_nm0(int irec);
bool check_pram(const raishin_pram& p);

extern thread_local size_t g_dalloc_bytes;  // running total of dalloc sizes on this thread

template <typename T>
bool dalloc(T** p, size_t n) {
  HIPCHK(hipMalloc((void**)p, sizeof(T) * (n ? n : 1)));
  g_dalloc_bytes += sizeof(T) * (n ? n : 1);
  return true;
}
void dfree_all(raishin_model* m);
bool upload(double* dst, const double* src, size_t n, hipStream_t st);
// One stored phi row (Geo::jfix): allowed for the model's configuration,
// states whose rows are bit-identical, switching the geometry, and moving a
// row between a full host array and a packed (9, ni, nk) one.
bool collapsible(const raishin_model* m);
bool rows_equal(const rh::Geo& G, const double* a, int j0);
void set_rows(raishin_model* m, bool collapse);
void pack_row(const rh::Geo& G, const double* full, std::vector<double>& out);
void unpack_row(const rh::Geo& G, const std::vector<double>& in, double* full);
void resolve_timing(raishin_model* m);
// raishin_model_set_state / get_state reporting this call's outcome (false:
// failed, error set), for callers inside the library
bool model_set_state(raishin_model* m, const double* uri, const double* uu, double time);
bool model_get_state(raishin_model* m, double* uri, double* uu, double* time);
// metric_rows: the phi rows the metric arrays hold, jmax (the reference's
// arrays, checked to be independent of phi) or 1 (the row of an
// axisymmetric metric).  gcv == nullptr: no metric arrays; the tables are
// computed on the device from the grid (x1a, x3a: face positions).
bool build_model(raishin_model* m, const raishin_pram* p, const double* x1, const double* x2, const double* x3,
                 const double* dx1b, const double* dx2b, const double* dx3b, const double* const gcv[4],
                 const double* const gcn[4], const double* const dtg[4], const double* christ, int metric_rows,
                 const double* x1a = nullptr, const double* x3a = nullptr);
// rh_setup.cpp: per k the sin/cos inputs of the metric (8 values per k)
void theta_trig_table(const raishin_pram& p, const double* x3, const double* x3a, double* out);

// Pieces of one RK stage on explicit buffers (SoA planes on the device).
// Fluxes: the three direction sweeps and flux-CT from primitives W into F (27 planes).
bool stage_fluxes(raishin_model* m, const double* W, double* F);
bool stage_sweeps(raishin_model* m, const double* W, double* F);  // the three sweeps (hll)
bool stage_ct(raishin_model* m, double* F);                        // flux-CT in place (ct)
// RK update + sources + recovery into m->Wpre (interior), then the in-order fix-up.
// ak: device dt/dx*b arrays and dt from the caller, or all null to use the
// model's own CFL result.
struct HostDt { const double *ak1 = nullptr, *ak2 = nullptr, *ak3 = nullptr; double dt = 0.0; };
bool stage_update(raishin_model* m, int rkstage, const double* W, const double* Ua, const double* Ub, double* Uout,
                  const double* F, const HostDt* hd);
// Ghost fill, floors and conserved variables: m->Wpre -> Wout, Uout.
bool stage_post(raishin_model* m, const double* Wold, const double* Wo, double* Wout, double* Uout);
bool run_cfl(raishin_model* m, const double* W);
bool run_stage(raishin_model* m, int rkstage, const double* Ua, const double* Ub, double* Uout,
               const struct StageTap* tap = nullptr);
bool run_step(raishin_model* m);
// cdtcfl4 of primitives W: dtcfl, dt1..dt3 to the host (synchronises).
bool cdtcfl4_on(raishin_model* m, const double* W, double* dtcfl, double* dt1, double* dt2, double* dt3);

// The stored face fluxes to the host: ww(3,9,imax,jmax,kmax) and the six
// transverse-B copies of flux-CT, ct(6,imax,jmax,kmax); either may be null.
bool copy_fluxes_out(raishin_model* m, double* ww, double* ct);

struct StageTap {  // optional host copies of a stage's intermediates (reference layouts)
  double *ww = nullptr, *wpre = nullptr;
};


}  // namespace rhm
