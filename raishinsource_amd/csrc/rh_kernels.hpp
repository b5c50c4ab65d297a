// Kernel argument blocks and launchers of the GRMHD step (see rh_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <vector>
#include "rh_device.hpp"

namespace rh {

struct StageArgs {
  const double *W;                 // primitives at stage start (9 planes)
  const double *Ua, *Ub;           // RK inputs: stage1 uu; stage2 uo,uh; stage3 uo,us
  const void *F1, *F2, *F3;        // ww(1..3,:,...) after ct (9 planes each; float when f32flux)
  const double *metC, *chr;        // centre metric tables, Christoffel tables
  const double *dx1b, *dx2b, *dx3b;
  const double *r15, *r25;         // x1(i)**(-3/2), x1(i)**(-5/2) from the host libm
  const unsigned long long *dtbits;
  const double *ak1, *ak2, *ak3;   // dt/dx*b from the caller (else from dtbits)
  double dt_host;
  double *Wpre;                    // recovered primitives (interior)
  double *wgout;                   // final W per cell (inherited initial guesses)
  double *Ukeep;                   // the stage's RK result (interior cells)
  unsigned long long *cnt, *cpart; // cpart: per-wavefront event rows (count_wave)
  unsigned long long *carry;       // [0] count, [1..cap] flagged cells
  unsigned char *cflag;            // per cell: 1 while its inherited W is unresolved
  long long carry_cap;
  double gam, cfl, dmin, pmin;
  int iter;
  bool chr_sym;                    // Christoffel tables symmetric in the lower indices
  bool f32flux;                    // F1..F3 hold float (the mixed form's sweeps)
  // the mixed form: flux-CT's six averaged transverse-B fluxes in FP64
  // (planes ww(2,7) ww(3,7) ww(1,8) ww(3,8) ww(1,9) ww(2,9)), which the
  // update reads instead of their float copies in F1..F3; null in FP64
  const double *CTd;
};

struct PostArgs {
  const double *Wpre, *Wold, *Wo;
  double *Wghost;                  // = Wpre: k_bnd writes its ghost slots
  const double *metC, *r15, *r25;
  const int *bc;                   // boundary codes [axis*2+side][9]
  double *Wout, *Uout;
  unsigned long long *cnt, *cpart;
  double gam, dmin, pmin, pmax;
  int model;
};

struct CflArgs {
  const double *W, *metC, *dx1b, *dx2b, *dx3b;
  unsigned long long *dtbits;      // [0] dtcfl [1..3] dt1..dt3 (bit patterns of the doubles)
  long long *blast;                // per block and accumulator: position of the last NaN
  double *bmin;                    // per block and accumulator: minimum after it
  double gam;
  int nblocks;                     // workgroups of the k_cfl launch (set by launch_cfl)
};

struct FixArgs {
  const double *W, *Ustage, *metC, *r15, *r25;
  double *Wpre, *wgout;
  unsigned long long *carry, *cnt;
  unsigned char *cflag, *pend;
  long long carry_cap;
  double gam, dmin, pmin;
  int iter;
};

struct EdgeArgs {
  const double *W, *dxs, *xs, *pc;  // pc: PPM width coefficients, or the face positions x*a for MC
  double gam;
  void *E;                         // float or double planes, by the launch's variant
};

struct SweepArgs {
  const double *W, *metF;
  const void *E;
  double gam;
  void *F, *CTsrc;  // element type: the sweep's arithmetic type (double, or float when mixed)
};

struct PSweepArgs {
  const double *W, *metF, *dxs, *xs, *pc;  // pc: PPM width coefficients of the direction
  double gam;
  void *F, *CTsrc;  // element type: the sweep's arithmetic type
  int seg;          // interfaces per lane along j or k
};

struct CtArgs {
  const void *CTsrc;  // double, or float when mixed (as the sweeps wrote them)
  void *F1, *F2, *F3;
  double *CTd;        // mixed form: the six averages also in FP64 (StageArgs::CTd)
};

struct AdvArgs {
  const unsigned long long *dtbits;
  double cfl;
  double *tdt;
};

struct CopyArgs {
  const double *src;
  double *dst;
  long long n;
};

// Launch slots whose workgroup size is tuned per model (Geo::bs): the PPM
// edges (all directions), the three sweeps, flux-CT, the RK update, recov1a,
// and the ghost fill with floors and caluu1.  The CFL reduction keeps 256
// (its per-block partial results are sized for it).
enum TuneSlot { TS_EDGES, TS_SWEEP1, TS_SWEEP2, TS_SWEEP3, TS_FLUXCT, TS_RK, TS_RECOV, TS_POST, TS_N };
constexpr int TUNE_CANDIDATES[3] = {64, 128, 256};
inline int slot_threads(const Geo& G, int slot) { return G.bs[slot] ? (int)G.bs[slot] : 256; }

// Launch kinds of one step, in the order run_step makes them.
enum LaunchKind {
  LK_EDGES, LK_SWEEP, LK_FLUXCT, LK_RK, LK_RECOV, LK_CARRY, LK_BND, LK_POST, LK_CFL, LK_CFL_FINISH,
  LK_ADVANCE, LK_COPY, LK_PSWEEP, LK_N
};

union AnyArgs {
  EdgeArgs e;
  SweepArgs s;
  PSweepArgs ps;
  CtArgs c;
  StageArgs st;
  FixArgs fx;
  PostArgs p;
  CflArgs cf;
  AdvArgs a;
  CopyArgs cp;
};

// A launch as recorded for one model: kind, template variant (edges: dir +
// 4*mixed; sweep: dir + 4*mixed + 8*irec; rk: stage + 8*sym), workgroups.
struct LaunchRec {
  int kind, variant;
  int threads;                     // workgroup size
  long long nb;
  Geo G;
  AnyArgs A;
};

// One model's entry in the device table of a batched launch; off = its first
// workgroup (kinds with one workgroup per model: its index).
struct BatchItem {
  Geo G;
  AnyArgs A;
  long long off;
};

// While r is set, the launch_* helpers below append a record to *r instead
// of launching (thread-local; nullptr ends recording).
void set_recorder(std::vector<LaunchRec>* r);
// One launch of a kind and variant over a table of n models (total_blocks
// workgroups; for one-workgroup-per-model kinds, n).
hipError_t launch_batched(int kind, int variant, int threads, const BatchItem* items, int n, long long total_blocks,
                          hipStream_t st);
int block_threads(int kind);
// dst[q] = src[q][0] for q < n (the members' times of a batch)
hipError_t launch_gather_first(const double* const* src, double* dst, int n, hipStream_t st);
bool one_group_per_model(int kind);

// Face values of every cell along dir into E (18 planes: left edges, right
// edges): PPM (irec 5; tab = the width-only interp coefficients of the
// direction, 7 planes of n+2, ppm_coefficients) or the MC limiter (irec 1;
// tab = the face positions x*a).  mixed: the reconstruction and the face
// arithmetic in float (E holds floats), else double.
hipError_t launch_edges(const Geo& G, int dir, int irec, const double* W, const double* dxs, const double* xs,
                        const double* tab, double gam, void* E, bool mixed, hipStream_t st);
// Fluxes at the interfaces along dir; E = the PPM edges (irec 5), unused otherwise.
// F and CTsrc take doubles, or floats when mixed (every value the float
// arithmetic produced, so storing it in float loses nothing).
hipError_t launch_sweep(const Geo& G, int dir, int irec, const double* W, const double* metF, const void* E,
                        double gam, void* F, void* CTsrc, bool mixed, hipStream_t st);
// PPM (irec 5) sweep along dir with the face values computed in the same
// kernel (through LDS): reads W and the direction's PPM width coefficients
// pc, writes what launch_edges + launch_sweep write to F and CTsrc.
// seg: cells along j or k per wave of the fused sweep (64 / seg consecutive
// i per wave), a power of two from 2 to 64
constexpr int PSWEEP_SEG_DEFAULT = 16;
inline bool psweep_seg_ok(int seg) { return seg >= 2 && seg <= 64 && (seg & (seg - 1)) == 0; }
hipError_t launch_psweep(const Geo& G, int dir, const double* W, const double* metF, const double* dxs,
                         const double* xs, const double* pc, double gam, void* F, void* CTsrc, bool mixed,
                         hipStream_t st);
int psweep_segment(const Geo& G, int dir);
size_t psweep_lds_bytes(int threads, bool mixed);
// f32: CTsrc and F hold floats (the averages are formed in double either way)
hipError_t launch_fluxct(const Geo& G, const void* CTsrc, void* F1, void* F2, void* F3, bool f32, double* CTd,
                         hipStream_t st);
// n values of type double (to_float) or float (else) from src into dst, the other type
hipError_t launch_retype(const void* src, void* dst, long long n, bool to_float, hipStream_t st);
// RK update + sources into A.Ukeep (every interior cell), then recov1a into A.Wpre.
hipError_t launch_update(const Geo& G, int stage, const StageArgs& A, hipStream_t st);
hipError_t launch_recov(const Geo& G, const StageArgs& A, hipStream_t st);
hipError_t launch_recov_carry(const Geo& G, const FixArgs& A, hipStream_t st);
hipError_t launch_post(const Geo& G, const PostArgs& A, hipStream_t st);
hipError_t launch_cfl(const Geo& G, const CflArgs& A, hipStream_t st);
long long cfl_blocks(const Geo& G);
// Fold the per-wavefront event rows into cnt (and clear them).
hipError_t launch_sum_counters(unsigned long long* cpart, long long rows, unsigned long long* cnt, hipStream_t st);
hipError_t launch_advance_time(const Geo& G, const unsigned long long* dtbits, double cfl, double* tdt,
                               hipStream_t st);
// dst <- src (n doubles), stream-ordered
// The metric and Christoffel (i,k) tables of an axisymmetric metric from the
// grid, on the device (k_metric_tables); tr: per k the host's sin/cos at
// x3, x3+1e-5, x3-1e-5 and x3a.
hipError_t launch_metric_tables(int ni, int nk, int metric, double akm, const double* x1, const double* x1a,
                                const double* x3, double ph, const double* tr, double* const met[4], double* chr,
                                int* sym, hipStream_t st);
hipError_t launch_copy(const Geo& G, double* dst, const double* src, long long n, hipStream_t st);
hipError_t launch_caluu1(const Geo& G, const double* W, const double* metC, double gam, double* U, hipStream_t st);
hipError_t launch_aos_to_soa(const double* aos, double* soa, long long N, hipStream_t st);
hipError_t launch_soa_to_aos(const double* soa, double* aos, long long N, hipStream_t st);
hipError_t launch_ww_to_planes(const double* ww, double* F, long long N, hipStream_t st);
hipError_t launch_planes_to_ww(const double* F, double* ww, long long N, hipStream_t st);

}  // namespace rh
