/* raishin_hip.h -- C-ABI of the MI355X-native RAISHIN GRMHD time step.
 *
 * Drop-in boundary for the reference's time-step path (main.f90:743-1149):
 * plain pointers and sizes, no torch types.  Arrays use the reference's own
 * column-major layouts with one MPI rank (is1=js1=ks1=1, ie1=imax, ...):
 *   state  uri/uu(nv=9, imax, jmax, kmax)        main.f90:35-57
 *   metric gcov/gcon(0:3, 0:3, imax, jmax, kmax) main.f90:62-70
 *   detg(imax, jmax, kmax), christ(0:3,0:3,0:3, imax, jmax, kmax)
 *   grids  x1(imax), dx1b(imax-1), ...           main.f90:115-126
 * Integer arguments of the reference-named routines are passed by address,
 * as Fortran passes them, so an iso_c_binding interface binds them directly.
 *
 * Errors: entry points keep the reference's void signatures; a failure sets a
 * sticky error read with raishin_last_error().
 */
#ifndef RAISHIN_HIP_H
#define RAISHIN_HIP_H
#ifdef __cplusplus
extern "C" {
#endif

/* Compile-time `module pram` values of the reference (pram.f90:11-80). */
typedef struct {
  int imax, jmax, kmax;
  int irec, irkt, ict, ihll, icha, iwvec, ieos, model, metric, iter;
  double cfl, gam, akm, dmin, pmin, pmax;
  int iboux1in[9], iboux1ot[9], iboux2in[9], iboux2ot[9], iboux3in[9], iboux3ot[9];
} raishin_pram;

/* Event counts; the first five are the events recov1a reports on stdout
 * (recov.f90:859,921,1124,1130,1154), the floors are silent in the reference. */
typedef struct {
  unsigned long long recov_fallback; /* 2D Newton-Raphson failed, entropy inversion used */
  unsigned long long recov_fail;     /* "Not convergence in recov1-3": previous primitives kept */
  unsigned long long recov_neg_rho;  /* entropy branch "negative density" */
  unsigned long long recov_neg_p;    /* entropy branch "negative pressure" */
  unsigned long long recov_vsq_gt1;  /* "vsq > 1.0 at initial guess" */
  unsigned long long recov_det0;     /* "det=0 in recov1" */
  unsigned long long recov_wg_carry; /* initial W inherited from the previous cell */
  unsigned long long atm_floor;      /* atmosphere1a resets */
  unsigned long long pmm_scale;      /* pminmax2b 1.05 rescaling entered */
  unsigned long long pmm_floor;      /* pminmax2b min/max/ratio clamps */
  unsigned long long guard;          /* iteration guards hit (unbounded loops in the reference) */
} raishin_counters;

/* Sticky error: returns 0 when no error, else copies the message. */
int raishin_last_error(char *buf, int len);
void raishin_clear_error(void);

/* ---- resident model: the production path ------------------------------- */
typedef struct raishin_model raishin_model;

/* One model on the current device: pram + grid (grid.f90:2) + metric arrays
 * (coord.f90:5 coord1, coord.f90:1089 calchrist), all host arrays in the
 * reference layout.  The metric must not depend on j (phi): it is stored as
 * (i,k) tables; a j-dependent metric is rejected.  Replaces the set-up half of
 * main.f90:436-448.  Returns NULL on error. */
raishin_model *raishin_model_create(const raishin_pram *p, const double *x1, const double *x2,
                                    const double *x3, const double *dx1b, const double *dx2b,
                                    const double *dx3b, const double *gcov, const double *gcon,
                                    const double *gcovi, const double *gconi, const double *gcovj,
                                    const double *gconj, const double *gcovk, const double *gconk,
                                    const double *detg, const double *detgi, const double *detgj,
                                    const double *detgk, const double *christ);
/* The same model from the one phi row of an axisymmetric metric: gcov..gconk
 * (4,4,imax,kmax), detg..detgk (imax,kmax), christ (4,4,4,imax,kmax), as
 * raishin_coord1 / raishin_calchrist write them for a pram with jmax = 1.
 * Bit-identical tables to raishin_model_create on the full arrays, without
 * building (and first-touching) 200 doubles per cell on the host; no
 * counterpart in the reference, whose metric arrays always hold every cell. */
raishin_model *raishin_model_create_ik(const raishin_pram *p, const double *x1, const double *x2,
                                       const double *x3, const double *dx1b, const double *dx2b,
                                       const double *dx3b, const double *gcov, const double *gcon,
                                       const double *gcovi, const double *gconi, const double *gcovj,
                                       const double *gconj, const double *gcovk, const double *gconk,
                                       const double *detg, const double *detgi, const double *detgj,
                                       const double *detgk, const double *christ);
/* The same model with no metric arrays at all: the metric at centres and
 * faces, sqrt(-g) and the Christoffel symbols (metric 303 or 1, the set-up's
 * metrics) are computed on the device from the grid (grid.f90: x1, x2, x3,
 * dx*b, and the face positions x1a, x3a).  Only sin and cos run on the host
 * (per theta), so the tables equal those of raishin_coord1 and
 * raishin_calchrist bit for bit.  No counterpart in the reference; for
 * drivers that create many models (ensembles, serving). */
raishin_model *raishin_model_create_grid(const raishin_pram *p, const double *x1, const double *x2,
                                         const double *x3, const double *dx1b, const double *dx2b,
                                         const double *dx3b, const double *x1a, const double *x3a);
/* Copy one of the model's (i,k) tables to the host (planes of imax*kmax,
 * i fastest): which 0..3 the metric at centres, i-, j- and k-faces (33
 * planes: gcov(m,n) at m+4n, gcon(m,n) at 16+m+4n, sqrt(-g) at 32), 4 the
 * Christoffel symbols (64 planes: Gamma^l_{mn} at l+4m+16n).  0, or -1 with
 * the error set. */
int raishin_model_tables(raishin_model *m, int which, double *out);
void raishin_model_destroy(raishin_model *m);

/* Primitive (uri) and conserved (uu) state, reference layout (9,imax,jmax,kmax).
 * set: host -> device (the restart path of restar.f90:273-317 / the end of
 * init, main.f90:530-539); uu may be NULL, then uu = caluu1(uri) is computed
 * on the device as main.f90:530 does.  get: device -> host (before
 * store1/output1); either pointer may be NULL. */
void raishin_model_set_state(raishin_model *m, const double *uri, const double *uu, double time);
void raishin_model_get_state(raishin_model *m, double *uri, double *uu, double *time);

/* nsteps full time steps: cdtcfl4 + 2 or 3 RK stages each (main.f90:743-1149),
 * all on the device, stream-ordered; no host synchronisation inside.
 * *time and *dt_last are written when non-NULL (that synchronises). */
void raishin_model_step(raishin_model *m, int nsteps, double *time, double *dt_last);

/* 2*npairs steps through a HIP graph of two steps, captured on first use (and
 * again after raishin_model_step ran an odd number of steps or the precision
 * changed): one graph launch per two steps instead of ~40 kernel launches per
 * step.  Same results as raishin_model_step. */
void raishin_model_step_graph(raishin_model *m, int npairs, double *time, double *dt_last);
/* Many models, each on its own stream: 2*npairs steps of every model, graph
 * launches queued round robin so the models run concurrently.  Asynchronous:
 * raishin_model_sync or raishin_model_get_state waits for a model. */
void raishin_models_step_graph(raishin_model *const *models, int nmodels, int npairs);

/* Only the CFL reduction (cdtcfl4 + MIN): dtcfl and dt1..dt3 of the current state. */
void raishin_model_cdtcfl4(raishin_model *m, double *dtcfl, double *dt1, double *dt2, double *dt3);

/* The same reduction with no allocation and no host synchronisation; no
 * counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  out: 4
 * doubles in device memory, written in order on the model's stream with
 * dtcfl, dt1, dt2, dt3 (bit-identical to raishin_model_cdtcfl4).  The
 * per-workgroup partials go to one fixed 8 MiB device array per GPU (the
 * library's, held until the process ends) instead of the model's buffers:
 * models of up to 131072 * 256 CFL cells ((imax-2)(jmax-2)(kmax-2) rounded
 * up per 256); larger ones are rejected with an error.  The array needs no
 * clearing (every call writes what it reads); calls that overlap on one GPU
 * race on it, so order them (one stream, or events).  The model's device
 * must be current.  Returns 0, or -1 with raishin_last_error set. */
int raishin_model_cdtcfl4_noalloc(raishin_model *m, double *out);

void raishin_model_counters(raishin_model *m, raishin_counters *c);
void raishin_model_reset_counters(raishin_model *m);
void raishin_model_sync(raishin_model *m);
/* Device address of the stream the model's kernels run on (hipStream_t). */
void *raishin_model_stream(raishin_model *m);

/* Arithmetic of the reconstruction and of the face fluxes (wave speeds,
 * HLLE): 0 = FP64, the reference's arithmetic (default); 1 = mixed: those
 * two in FP32 on the vector ALUs, while the metric tables, flux-CT, the RK
 * accumulation, the sources, recov1a, the floors and the state stay FP64.
 * The face fluxes are stored as float in mode 1 (flux-CT reads them and
 * averages in FP64, and writes its result back as float); a change of mode
 * converts the stored fluxes in place, in order on the model's stream, with
 * no allocation and no device-wide wait.  Call it between steps.
 * Not part of the reference; its error against FP64 is measured by bench.py
 * (its "mixed" object, --mixed-steps).  Takes effect from the next stage. */
void raishin_model_set_precision(raishin_model *m, int mode);

/* Per-kernel timing with HIP events on the model's stream.  Kinds, in order:
 * 0..2 direction sweeps i,j,k; 3 flux-CT; 4 RK update + sources;
 * 5 in-order recovery fix-up; 6 ghosts/floors/caluu1; 7 CFL reduction;
 * 8 PPM face values (irec=5, all three directions); 9 recov1a.
 * set_timing(1) clears the totals and starts recording; get_timing returns
 * the number of kinds and fills total milliseconds and launch counts. */
void raishin_model_set_timing(raishin_model *m, int on);
/* As set_timing, for the kinds whose bit (1 << kind) is set in mask: the
 * events between launches cost a little, so a measurement that needs one
 * kind's launch times times that kind alone. */
void raishin_model_set_timing_mask(raishin_model *m, unsigned mask);

/* One RK stage of a step (stage 1 includes the CFL reduction), exposing its
 * intermediates in reference layouts when the pointers are non-NULL:
 * ww(3,9,imax,jmax,kmax) as after ct() (main.f90:845) and the recovered
 * primitives as after recov() (main.f90:861; interior cells).  Calling
 * stages 1..irkt in turn is one raishin_model_step. */
void raishin_model_debug_stage(raishin_model *m, int stage, double *ww, double *wpre);
/* The stored face fluxes, as the next sweep would find them: ww(3,9,imax,
 * jmax,kmax) in the layout of main.f90's ww and the six transverse-B flux
 * copies that flux-CT averages, ct(6,imax,jmax,kmax); either may be NULL.
 * In the mixed form they are stored as float and returned widened. */
void raishin_model_debug_fluxes(raishin_model *m, double *ww, double *ct);
int raishin_model_get_timing(raishin_model *m, double *ms, long long *launches, int nkinds);

/* Workgroup sizes of the tuned launches, in the order: PPM face values,
 * sweep i, sweep j, sweep k, flux-CT, RK update, recov1a, ghosts/floors/
 * caluu1 (8 entries; 0 = 256, else 64..256 in steps of 64).  Results do not
 * depend on them.  get returns the sizes in use. */
int raishin_model_set_block_sizes(raishin_model *m, const int *bs, int n);
int raishin_model_get_block_sizes(raishin_model *m, int *bs, int n);
/* Tiles of the PPM (irec = 5) sweeps.  split = 0 (default): each sweep
 * computes its face values itself, through LDS; a wave of the j (k) sweep
 * holds seg_j (seg_k) consecutive cells along j (k) by 64 / seg consecutive
 * i, a power of two from 2 to 64 (0 keeps the default of 16).  split = 1:
 * the face values of every cell go through HBM from a kernel of their own.
 * Results do not depend on them.  0, or -1 with the error set. */
int raishin_model_set_sweep_tiles(raishin_model *m, int split, int seg_j, int seg_k);
int raishin_model_get_sweep_tiles(raishin_model *m, int *split, int *seg_j, int *seg_k);
/* Face flux of the normal magnetic field component.  0 (default): the
 * reference's HLLE flux, whose dissipation term c(U_r - U_l) is nonzero for
 * the normal component (hll.f90:109-111) and which flux-CT leaves alone
 * (ct.f90:98-109), so div B drifts.  1: that flux set to zero, so flux-CT
 * keeps the corner divergence of sqrt(-g) B at FP64 round-off.  Form 1 is
 * not the reference's arithmetic and has no counterpart there.  0, or -1
 * with the error set. */
int raishin_model_set_divb_form(raishin_model *m, int form);
/* Autotuner: for each candidate size (64, 128, 256) runs one warm-up step
 * and `steps` timed steps from the state (uri, uu or NULL, time), keeps the
 * fastest size per launch slot; for PPM models it then times the fused
 * sweeps with tiles of 8, 16 and 32 cells along j and k and keeps the
 * fastest tile per direction (the split form is never chosen: it is slower
 * at every size measured, and members of a batch must agree on the form);
 * then it sets the state again and clears the counters, so the run
 * starts as if untuned.  Returns 0 or -1. */
int raishin_model_autotune(raishin_model *m, const double *uri, const double *uu, double time, int steps);

/* Axisymmetric models with one interior phi cell (jmax = 2*nm1+1), periodic
 * phi and non-periodic r and theta keep every phi row identical at all
 * times.  When raishin_model_set_state receives a state whose rows are
 * bit-identical, the model stores and computes one row (results unchanged,
 * get_state returns every row); allow = 0 keeps all rows.  stored_rows
 * returns the phi rows in use (1 or jmax). */
void raishin_model_allow_collapse(raishin_model *m, int allow);
int raishin_model_stored_rows(raishin_model *m);

/* Device memory the model holds, in bytes (state, fluxes, metric tables and
 * work arrays): what admission of a new model has to find free. */
long long raishin_model_device_bytes(raishin_model *m);

/* ------------------------------------------------------------------------
 * Batches: many resident models advanced together (no counterpart in the
 * reference, whose main loop runs one model per MPI job).  Every kernel of a
 * step is one launch for all members; each member owns a contiguous range
 * of the launch's workgroups.  Members may differ in size, metric, initial
 * data and precision, and must share irec, irkt and the boundary types that
 * decide which kernels a step makes.  Each member's results are those of
 * raishin_model_step on that member alone, bit for bit.  The calls are
 * stream-ordered with the members' own streams (the batch waits for work
 * queued on them, and later work on them waits for the batch).  Returns 0,
 * or -1 with raishin_last_error set.
 * ------------------------------------------------------------------------ */
typedef struct raishin_batch raishin_batch;
raishin_batch *raishin_batch_create(void);
void raishin_batch_destroy(raishin_batch *b);
/* Replace the members (a model appears at most once; n = 0 empties it). */
int raishin_batch_set(raishin_batch *b, raishin_model *const *models, int n);
/* nsteps steps of every member. */
int raishin_batch_step(raishin_batch *b, int nsteps);
/* 2*npairs steps of every member through a captured HIP graph of two steps
 * (captured again after the members or their buffers change). */
int raishin_batch_step_graph(raishin_batch *b, int npairs);
/* Simulation time of each member, in member order (waits for the batch). */
int raishin_batch_times(raishin_batch *b, double *times);
int raishin_batch_sync(raishin_batch *b);

#ifdef __cplusplus
}
#endif
#endif
