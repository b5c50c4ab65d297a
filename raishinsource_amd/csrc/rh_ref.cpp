// Reference-named entry points: the argument lists of the reference's
// subroutines (integers by address, arrays by base address in the reference's
// column-major layouts), for a Fortran main loop bound through iso_c_binding.
//
// State stays resident on the device between calls.  Every host array the
// calls name gets a device mirror keyed by its host address; the host copy
// goes stale when a call writes the mirror and is refreshed by raishin_pull.
// Routines whose outputs only feed later routines of the same RK stage
// (rec2 -> calcha -> hll, caltenr -> calsf -> rk* -> rk*ads* -> recov,
// bnd4 -> infchk -> atmosphere1a -> pminmax2b -> caluu1) are recorded and
// run fused when their consumer is called, in the order of main.f90:807-1147.
// A call out of that order is an error (raishin_last_error), not a fallback.
#include <cstring>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/raishin_ref.h"
#include "rh_model.hpp"

using namespace rh;
using namespace rhm;

namespace {

enum { MK_STATE = 9, MK_WW = 27 };  // planes per mirror kind

struct Mirror {
  double* dev = nullptr;
  int planes = 0;
  bool dev_valid = false;   // the device copy holds the array's current value
  bool host_valid = true;   // the host copy does
};

// What the current RK stage has recorded so far.
struct Pending {
  int phase = 0;  // 0 idle, 1 rec2, 2 calcha, 3 hll, 4 ct, 5 recov
  const double* uri = nullptr;
  bool tenr = false, sf = false;
  int rk = 0;  // 1 rk2fst, 2 rk2snd, 3 rk3snd, 4 rk3trd
  double* rk_out = nullptr;
  const double *rk_a = nullptr, *rk_b = nullptr;
  const double* ww = nullptr;
  int adsf = 0;
  double dt = 0.0;
  bool bnd = false, inf = false, atm = false, pmm = false;
  const double* urio = nullptr;
};

struct Ctx {
  raishin_pram p{};
  bool pram_set = false;
  raishin_model* m = nullptr;
  const double *x1 = nullptr, *dx1b = nullptr, *dx2b = nullptr, *dx3b = nullptr;
  const double *gcv[4] = {nullptr, nullptr, nullptr, nullptr}, *gcn[4] = {nullptr, nullptr, nullptr, nullptr};
  const double *dtg[4] = {nullptr, nullptr, nullptr, nullptr}, *christ = nullptr;
  std::unordered_map<const void*, Mirror> mir;
  double* ak = nullptr;       // device dt/dx1b, dt/dx2b, dt/dx3b (imax + jmax + kmax)
  double* staging = nullptr;  // AoS staging for transfers (27 N)
  Pending pd;
};

Ctx g;

bool need_model(const char* who) {
  if (!g.m) {
    set_error("%s: raishin_geometry has not been called", who);
    return false;
  }
  return true;
}

bool check_bounds(const char* who, const int* is1, const int* ie1, const int* js1, const int* je1, const int* ks1,
                  const int* ke1) {
  if (*is1 != 1 || *js1 != 1 || *ks1 != 1 || *ie1 != g.p.imax || *je1 != g.p.jmax || *ke1 != g.p.kmax) {
    set_error("%s: one rank spans the grid (is1=1..ie1=imax, ...), got %d..%d %d..%d %d..%d", who, *is1, *ie1, *js1,
              *je1, *ks1, *ke1);
    return false;
  }
  return true;
}

bool check_geom(const char* who, const double* got, const double* want, const char* what) {
  if (got != want) {
    set_error("%s: %s is not the array given to raishin_geometry", who, what);
    return false;
  }
  return true;
}

// AoS host layout a(planes, i, j, k) <-> SoA device planes.  The flux array
// ww(3, 9, i, j, k) maps to plane d*9 + n.
bool to_device(Mirror& mr, const double* host) {
  const long long N = g.m->G.N;
  HIPCHK(hipMemcpyAsync(g.staging, host, (size_t)mr.planes * N * 8, hipMemcpyHostToDevice, g.m->st));
  if (mr.planes == MK_STATE) HIPCHK(launch_aos_to_soa(g.staging, mr.dev, N, g.m->st));
  else HIPCHK(launch_ww_to_planes(g.staging, mr.dev, N, g.m->st));
  mr.dev_valid = true;
  return true;
}

bool to_host(Mirror& mr, double* host) {
  const long long N = g.m->G.N;
  if (mr.planes == MK_STATE) HIPCHK(launch_soa_to_aos(mr.dev, g.staging, N, g.m->st));
  else HIPCHK(launch_planes_to_ww(mr.dev, g.staging, N, g.m->st));
  HIPCHK(hipMemcpyAsync(host, g.staging, (size_t)mr.planes * N * 8, hipMemcpyDeviceToHost, g.m->st));
  HIPCHK(hipStreamSynchronize(g.m->st));
  mr.host_valid = true;
  return true;
}

Mirror* mirror(const void* host, int planes) {
  auto it = g.mir.find(host);
  if (it != g.mir.end()) {
    if (it->second.planes != planes) {
      set_error("array at %p used with %d and %d planes", host, it->second.planes, planes);
      return nullptr;
    }
    return &it->second;
  }
  Mirror mr;
  mr.planes = planes;
  if (!dalloc(&mr.dev, (size_t)planes * g.m->G.N)) return nullptr;
  return &(g.mir[host] = mr);
}

// Device copy of an input array, uploaded when the host copy is the current one.
const double* input(const double* host, int planes) {
  Mirror* mr = mirror(host, planes);
  if (!mr) return nullptr;
  if (!mr->dev_valid && !to_device(*mr, host)) return nullptr;
  return mr->dev;
}

// Device copy of an output array; the host copy goes stale.
double* output(double* host, int planes) {
  Mirror* mr = mirror(host, planes);
  if (!mr) return nullptr;
  mr->dev_valid = true;
  mr->host_valid = false;
  return mr->dev;
}

void reset_stage() { g.pd = Pending{}; }

bool same_pram(const raishin_pram& a, const raishin_pram& b) {
  const int ia[] = {a.imax, a.jmax, a.kmax, a.irec, a.irkt, a.ict, a.ihll, a.icha, a.iwvec, a.ieos, a.model, a.metric, a.iter};
  const int ib[] = {b.imax, b.jmax, b.kmax, b.irec, b.irkt, b.ict, b.ihll, b.icha, b.iwvec, b.ieos, b.model, b.metric, b.iter};
  const double da[] = {a.cfl, a.gam, a.akm, a.dmin, a.pmin, a.pmax}, db[] = {b.cfl, b.gam, b.akm, b.dmin, b.pmin, b.pmax};
  return std::memcmp(ia, ib, sizeof ia) == 0 && std::memcmp(da, db, sizeof da) == 0 &&
         std::memcmp(a.iboux1in, b.iboux1in, sizeof a.iboux1in) == 0 &&
         std::memcmp(a.iboux1ot, b.iboux1ot, sizeof a.iboux1ot) == 0 &&
         std::memcmp(a.iboux2in, b.iboux2in, sizeof a.iboux2in) == 0 &&
         std::memcmp(a.iboux2ot, b.iboux2ot, sizeof a.iboux2ot) == 0 &&
         std::memcmp(a.iboux3in, b.iboux3in, sizeof a.iboux3in) == 0 &&
         std::memcmp(a.iboux3ot, b.iboux3ot, sizeof a.iboux3ot) == 0;
}

bool fail_order(const char* who, const char* expected) {
  set_error("%s: out of main.f90 order (expected %s); the stage is discarded", who, expected);
  reset_stage();
  return false;
}

}  // namespace

extern "C" {

void raishin_init(const raishin_pram* p) {
  if (g.m && !same_pram(g.p, *p)) {
    set_error("raishin_init: pram differs from the one the geometry was built with; call raishin_finalize first");
    return;
  }
  if (!check_pram(*p)) return;
  g.p = *p;
  g.pram_set = true;
}

void raishin_geometry(const double* x1, const double* x2, const double* x3, const double* dx1b, const double* dx2b,
                      const double* dx3b, const double* gcov, const double* gcon, const double* gcovi,
                      const double* gconi, const double* gcovj, const double* gconj, const double* gcovk,
                      const double* gconk, const double* detg, const double* detgi, const double* detgj,
                      const double* detgk, const double* christ) {
  if (!g.pram_set) { set_error("raishin_geometry: raishin_init has not been called"); return; }
  if (g.m) { set_error("raishin_geometry: geometry already set; call raishin_finalize first"); return; }
  raishin_model* m = raishin_model_create(&g.p, x1, x2, x3, dx1b, dx2b, dx3b, gcov, gcon, gcovi, gconi, gcovj, gconj,
                                          gcovk, gconk, detg, detgi, detgj, detgk, christ);
  if (!m) return;
  g.m = m;
  const long long N = m->G.N;
  if (!dalloc(&g.ak, (size_t)g.p.imax + g.p.jmax + g.p.kmax) || !dalloc(&g.staging, (size_t)MK_WW * N)) return;
  g.x1 = x1; g.dx1b = dx1b; g.dx2b = dx2b; g.dx3b = dx3b; g.christ = christ;
  const double* gv[4] = {gcov, gcovi, gcovj, gcovk};
  const double* gn[4] = {gcon, gconi, gconj, gconk};
  const double* dg[4] = {detg, detgi, detgj, detgk};
  for (int a = 0; a < 4; a++) { g.gcv[a] = gv[a]; g.gcn[a] = gn[a]; g.dtg[a] = dg[a]; }
  reset_stage();
}

void raishin_finalize(void) {
  if (g.m) (void)hipStreamSynchronize(g.m->st);
  for (auto& kv : g.mir) (void)hipFree(kv.second.dev);
  g.mir.clear();
  if (g.ak) (void)hipFree(g.ak);
  if (g.staging) (void)hipFree(g.staging);
  if (g.m) raishin_model_destroy(g.m);
  g = Ctx{};
}

void raishin_push(const void* host) {
  auto it = g.mir.find(host);
  if (it == g.mir.end()) return;  // not mirrored yet: the next use uploads it
  it->second.dev_valid = false;
  it->second.host_valid = true;
}

void raishin_pull(void* host) {
  auto it = g.mir.find(host);
  if (it == g.mir.end() || it->second.host_valid) return;
  if (g.pd.phase == 5 && host == g.pd.uri) {
    set_error("raishin_pull: uri is between recov and caluu1 of a stage");
    return;
  }
  (void)to_host(it->second, static_cast<double*>(host));
}

void raishin_sync(void) {
  if (g.m && hipStreamSynchronize(g.m->st) != hipSuccess) set_error("raishin_sync failed");
}

void raishin_get_counters(raishin_counters* c) {
  if (need_model("raishin_get_counters")) raishin_model_counters(g.m, c);
}

void raishin_reset_counters(void) {
  if (need_model("raishin_reset_counters")) raishin_model_reset_counters(g.m);
}

// func.f90:2 ident(uu,un): un = uu.
void raishin_ident(const double* uu, double* un, const int* is1, const int* ie1, const int* js1, const int* je1,
                   const int* ks1, const int* ke1) {
  if (!need_model("ident") || !check_bounds("ident", is1, ie1, js1, je1, ks1, ke1)) return;
  const double* s = input(uu, MK_STATE);
  if (!s) return;
  double* d = output(un, MK_STATE);
  if (!d) return;
  if (hipMemcpyAsync(d, s, (size_t)MK_STATE * g.m->G.N * 8, hipMemcpyDeviceToDevice, g.m->st) != hipSuccess)
    set_error("ident: device copy failed");
}

// cdtcfl.f90:2-3 (+ the MIN of main.f90:761 for one rank).
void raishin_cdtcfl4(const double* uri, const double* gcov, const double* gcon, double* dtcfl, const double* dx1,
                     const double* dx2, const double* dx3, double* dt1, double* dt2, double* dt3, const int* nm1,
                     const int* is1, const int* ie1, const int* js1, const int* je1, const int* ks1, const int* ke1) {
  const char* who = "cdtcfl4";
  if (!need_model(who) || !check_bounds(who, is1, ie1, js1, je1, ks1, ke1)) return;
  if (*nm1 != g.m->G.nm1) { set_error("cdtcfl4: nm1=%d, the configuration has %d", *nm1, g.m->G.nm1); return; }
  if (!check_geom(who, gcov, g.gcv[0], "gcov") || !check_geom(who, gcon, g.gcn[0], "gcon") ||
      !check_geom(who, dx1, g.dx1b, "dx1") || !check_geom(who, dx2, g.dx2b, "dx2") ||
      !check_geom(who, dx3, g.dx3b, "dx3"))
    return;
  const double* W = input(uri, MK_STATE);
  if (!W) return;
  (void)cdtcfl4_on(g.m, W, dtcfl, dt1, dt2, dt3);
}

// rec2 (rec.f90:2-8): recorded; run by hll.
void raishin_rec2(const double* uri, const double* x1, const double* x2, const double* x3, const double* x1a,
                  const double* x2a, const double* x3a, const double* dx1, const double* dx2, const double* dx3,
                  const double* dx1b, const double* dx2b, const double* dx3b, const double* x1b, const double* x2b,
                  const double* x3b, const double* x1ab, const double* x2ab, const double* x3ab, double* uriir,
                  double* urijr, double* urikr, double* uriil, double* urijl, double* urikl, double* uuir,
                  double* uujr, double* uukr, double* uuil, double* uujl, double* uukl, double* wwir, double* wwjr,
                  double* wwkr, double* wwil, double* wwjl, double* wwkl, const double* gcovi, const double* gcovj,
                  const double* gcovk, const double* gconi, const double* gconj, const double* gconk,
                  const double* detgi, const double* detgj, const double* detgk, const int* nm0, const int* is1,
                  const int* ie1, const int* js1, const int* je1, const int* ks1, const int* ke1) {
  const char* who = "rec2";
  if (!need_model(who) || !check_bounds(who, is1, ie1, js1, je1, ks1, ke1)) return;
  if (*nm0 != g.m->G.nm0) { set_error("rec2: nm0=%d, the configuration has %d", *nm0, g.m->G.nm0); return; }
  if (!check_geom(who, x1, g.x1, "x1") || !check_geom(who, dx1b, g.dx1b, "dx1b") ||
      !check_geom(who, dx2b, g.dx2b, "dx2b") || !check_geom(who, dx3b, g.dx3b, "dx3b") ||
      !check_geom(who, gcovi, g.gcv[1], "gcovi") || !check_geom(who, gcovj, g.gcv[2], "gcovj") ||
      !check_geom(who, gcovk, g.gcv[3], "gcovk") || !check_geom(who, gconi, g.gcn[1], "gconi") ||
      !check_geom(who, gconj, g.gcn[2], "gconj") || !check_geom(who, gconk, g.gcn[3], "gconk") ||
      !check_geom(who, detgi, g.dtg[1], "detgi") || !check_geom(who, detgj, g.dtg[2], "detgj") ||
      !check_geom(who, detgk, g.dtg[3], "detgk"))
    return;
  (void)x2; (void)x3; (void)x1a; (void)x2a; (void)x3a; (void)dx1; (void)dx2; (void)dx3; (void)x1b; (void)x2b;
  (void)x3b; (void)x1ab; (void)x2ab; (void)x3ab; (void)uriir; (void)urijr; (void)urikr; (void)uriil; (void)urijl;
  (void)urikl; (void)uuir; (void)uujr; (void)uukr; (void)uuil; (void)uujl; (void)uukl; (void)wwir; (void)wwjr;
  (void)wwkr; (void)wwil; (void)wwjl; (void)wwkl;
  if (g.pd.phase != 0) { fail_order(who, "no stage in progress"); return; }
  g.pd.phase = 1;
  g.pd.uri = uri;
}

// calcha (calcha.f90:2-5): recorded; run by hll.
void raishin_calcha(const double* uriir, const double* urijr, const double* urikr, const double* uriil,
                    const double* urijl, const double* urikl, const double* gcovi, const double* gcovj,
                    const double* gcovk, const double* gconi, const double* gconj, const double* gconk, double* cmaxi,
                    double* cmini, double* cmaxj, double* cminj, double* cmaxk, double* cmink, const int* nm0,
                    const int* is1, const int* ie1, const int* js1, const int* je1, const int* ks1, const int* ke1) {
  const char* who = "calcha";
  if (!need_model(who) || !check_bounds(who, is1, ie1, js1, je1, ks1, ke1)) return;
  (void)uriir; (void)urijr; (void)urikr; (void)uriil; (void)urijl; (void)urikl; (void)gcovi; (void)gcovj;
  (void)gcovk; (void)gconi; (void)gconj; (void)gconk; (void)cmaxi; (void)cmini; (void)cmaxj; (void)cminj;
  (void)cmaxk; (void)cmink; (void)nm0;
  if (g.pd.phase != 1) { fail_order(who, "rec2 first"); return; }
  g.pd.phase = 2;
}

// hll (hll.f90:2-5): runs the fused direction sweeps (rec2 + calcha + hlle) into ww.
void raishin_hll(double* ww, const double* uuir, const double* uujr, const double* uukr, const double* uuil,
                 const double* uujl, const double* uukl, const double* wwir, const double* wwjr, const double* wwkr,
                 const double* wwil, const double* wwjl, const double* wwkl, const double* cmaxi, const double* cmini,
                 const double* cmaxj, const double* cminj, const double* cmaxk, const double* cmink, const int* nm0,
                 const int* is1, const int* ie1, const int* js1, const int* je1, const int* ks1, const int* ke1) {
  const char* who = "hll";
  if (!need_model(who) || !check_bounds(who, is1, ie1, js1, je1, ks1, ke1)) return;
  (void)uuir; (void)uujr; (void)uukr; (void)uuil; (void)uujl; (void)uukl; (void)wwir; (void)wwjr; (void)wwkr;
  (void)wwil; (void)wwjl; (void)wwkl; (void)cmaxi; (void)cmini; (void)cmaxj; (void)cminj; (void)cmaxk; (void)cmink;
  (void)nm0;
  if (g.pd.phase != 2) { fail_order(who, "rec2, calcha first"); return; }
  const double* W = input(g.pd.uri, MK_STATE);
  if (!W) return;
  // hlle keeps ww where the wave speeds are NaN: the device copy must hold ww's current value
  if (!input(ww, MK_WW)) return;
  double* F = output(ww, MK_WW);
  if (!stage_sweeps(g.m, W, F)) return;
  g.pd.phase = 3;
  g.pd.ww = ww;
}

// ct (ct.f90:4): flux-CT in place on ww.  wwo is not used by the flux-CT branch.
void raishin_ct(double* ww, double* wwo, const double* uri, const int* nm0, const int* nm1, const int* is1,
                const int* ie1, const int* js1, const int* je1, const int* ks1, const int* ke1) {
  const char* who = "ct";
  if (!need_model(who) || !check_bounds(who, is1, ie1, js1, je1, ks1, ke1)) return;
  (void)wwo; (void)uri; (void)nm0; (void)nm1;
  if (g.pd.phase != 3 || ww != g.pd.ww) { fail_order(who, "hll on the same ww first"); return; }
  double* F = output(ww, MK_WW);
  if (!stage_ct(g.m, F)) return;
  g.pd.phase = 4;
}

// caltenr (calsf.f90:6) and calsf (calsf.f90:107): recorded; fused into the update.
void raishin_caltenr(const double* uri, double* tenr, const double* gcov, const double* gcon, const int* is1,
                     const int* ie1, const int* js1, const int* je1, const int* ks1, const int* ke1) {
  const char* who = "caltenr";
  if (!need_model(who) || !check_bounds(who, is1, ie1, js1, je1, ks1, ke1)) return;
  (void)tenr;
  if (!check_geom(who, gcov, g.gcv[0], "gcov") || !check_geom(who, gcon, g.gcn[0], "gcon")) return;
  if (g.pd.phase != 4 || uri != g.pd.uri) { fail_order(who, "ct first, on the stage's uri"); return; }
  g.pd.tenr = true;
}

void raishin_calsf(double* sf, const double* tenr, const double* christ, const double* detg, const int* is1,
                   const int* ie1, const int* js1, const int* je1, const int* ks1, const int* ke1) {
  const char* who = "calsf";
  if (!need_model(who) || !check_bounds(who, is1, ie1, js1, je1, ks1, ke1)) return;
  (void)sf; (void)tenr;
  if (!check_geom(who, christ, g.christ, "christ") || !check_geom(who, detg, g.dtg[0], "detg")) return;
  if (g.pd.phase != 4 || !g.pd.tenr) { fail_order(who, "caltenr first"); return; }
  g.pd.sf = true;
}

}  // extern "C"

namespace {

bool record_rk(const char* who, int kind, double* out, const double* ww, const double* a, const double* b,
               const int* nm1, const double* ak1, const double* ak2, const double* ak3) {
  if (*nm1 != g.m->G.nm1) { set_error("%s: nm1=%d, the configuration has %d", who, *nm1, g.m->G.nm1); return false; }
  if (g.pd.phase != 4 || ww != g.pd.ww) return fail_order(who, "ct on the same ww first");
  if (g.pd.rk) return fail_order(who, "one RK update per stage");
  const int irkt = g.p.irkt;
  const bool ok = kind == 1 || (kind == 2 && irkt == 2) || ((kind == 3 || kind == 4) && irkt == 3);
  if (!ok) return fail_order(who, "the RK routine of irkt");
  // akap*b = dt/dx*b, computed by the caller (main.f90:765-779)
  const int n1 = g.p.imax, n2 = g.p.jmax, n3 = g.p.kmax;
  std::vector<double> ak((size_t)n1 + n2 + n3, 0.0);
  for (int q = 0; q < n1 - 1; q++) ak[q] = ak1[q];
  for (int q = 0; q < n2 - 1; q++) ak[n1 + q] = ak2[q];
  for (int q = 0; q < n3 - 1; q++) ak[n1 + n2 + q] = ak3[q];
  HIPCHK(hipMemcpyAsync(g.ak, ak.data(), ak.size() * 8, hipMemcpyHostToDevice, g.m->st));
  HIPCHK(hipStreamSynchronize(g.m->st));  // the host vector goes out of scope
  g.pd.rk = kind;
  g.pd.rk_out = out;
  g.pd.rk_a = a;
  g.pd.rk_b = b;
  return true;
}

bool record_adsf(const char* who, int kind, const double* uu, const double* dt, const int* nm1) {
  if (*nm1 != g.m->G.nm1) { set_error("%s: nm1=%d, the configuration has %d", who, *nm1, g.m->G.nm1); return false; }
  if (g.pd.rk != kind || uu != g.pd.rk_out || !g.pd.sf) return fail_order(who, "calsf and the matching RK update first");
  g.pd.adsf = kind;
  g.pd.dt = *dt;
  return true;
}

}  // namespace

extern "C" {

// rkt.f90:2 rk2fst(uh,ww,uu,...): uh = uu - akap*dF.
void raishin_rk2fst(double* uh, const double* ww, const double* uu, const int* nm1, const double* akap1b,
                    const double* akap2b, const double* akap3b, const int* is1, const int* ie1, const int* js1,
                    const int* je1, const int* ks1, const int* ke1) {
  if (!need_model("rk2fst") || !check_bounds("rk2fst", is1, ie1, js1, je1, ks1, ke1)) return;
  (void)record_rk("rk2fst", 1, uh, ww, uu, nullptr, nm1, akap1b, akap2b, akap3b);
}
// rkt.f90:45 rk2snd(uu,ww,uo,uh,...)
void raishin_rk2snd(double* uu, const double* ww, const double* uo, const double* uh, const int* nm1,
                    const double* akap1b, const double* akap2b, const double* akap3b, const int* is1, const int* ie1,
                    const int* js1, const int* je1, const int* ks1, const int* ke1) {
  if (!need_model("rk2snd") || !check_bounds("rk2snd", is1, ie1, js1, je1, ks1, ke1)) return;
  (void)record_rk("rk2snd", 2, uu, ww, uo, uh, nm1, akap1b, akap2b, akap3b);
}
// rkt.f90:91 rk3snd(uu,ww,uo,uh,...)
void raishin_rk3snd(double* uu, const double* ww, const double* uo, const double* uh, const int* nm1,
                    const double* akap1b, const double* akap2b, const double* akap3b, const int* is1, const int* ie1,
                    const int* js1, const int* je1, const int* ks1, const int* ke1) {
  if (!need_model("rk3snd") || !check_bounds("rk3snd", is1, ie1, js1, je1, ks1, ke1)) return;
  (void)record_rk("rk3snd", 3, uu, ww, uo, uh, nm1, akap1b, akap2b, akap3b);
}
// rkt.f90:137 rk3trd(uu,ww,uo,us,...)
void raishin_rk3trd(double* uu, const double* ww, const double* uo, const double* us, const int* nm1,
                    const double* akap1b, const double* akap2b, const double* akap3b, const int* is1, const int* ie1,
                    const int* js1, const int* je1, const int* ks1, const int* ke1) {
  if (!need_model("rk3trd") || !check_bounds("rk3trd", is1, ie1, js1, je1, ks1, ke1)) return;
  (void)record_rk("rk3trd", 4, uu, ww, uo, us, nm1, akap1b, akap2b, akap3b);
}
// rkt.f90:183-320 rk2adsff / rk2adsfs / rk3adsfs / rk3adsft(uu,sf,dt,nm1,...)
void raishin_rk2adsff(double* uu, const double* sf, const double* dt, const int* nm1, const int* is1, const int* ie1,
                      const int* js1, const int* je1, const int* ks1, const int* ke1) {
  (void)sf;
  if (!need_model("rk2adsff") || !check_bounds("rk2adsff", is1, ie1, js1, je1, ks1, ke1)) return;
  (void)record_adsf("rk2adsff", 1, uu, dt, nm1);
}
void raishin_rk2adsfs(double* uu, const double* sf, const double* dt, const int* nm1, const int* is1, const int* ie1,
                      const int* js1, const int* je1, const int* ks1, const int* ke1) {
  (void)sf;
  if (!need_model("rk2adsfs") || !check_bounds("rk2adsfs", is1, ie1, js1, je1, ks1, ke1)) return;
  (void)record_adsf("rk2adsfs", 2, uu, dt, nm1);
}
void raishin_rk3adsfs(double* uu, const double* sf, const double* dt, const int* nm1, const int* is1, const int* ie1,
                      const int* js1, const int* je1, const int* ks1, const int* ke1) {
  (void)sf;
  if (!need_model("rk3adsfs") || !check_bounds("rk3adsfs", is1, ie1, js1, je1, ks1, ke1)) return;
  (void)record_adsf("rk3adsfs", 3, uu, dt, nm1);
}
void raishin_rk3adsft(double* uu, const double* sf, const double* dt, const int* nm1, const int* is1, const int* ie1,
                      const int* js1, const int* je1, const int* ks1, const int* ke1) {
  (void)sf;
  if (!need_model("rk3adsft") || !check_bounds("rk3adsft", is1, ie1, js1, je1, ks1, ke1)) return;
  (void)record_adsf("rk3adsft", 4, uu, dt, nm1);
}

// recov.f90:2-3 recov(uu,uri,gcov,gcon,detg,x1,x2,x3,nm1,...): runs the fused
// update (RK + sources + recov1a) and the in-order fix-up.
void raishin_recov(double* uu, double* uri, const double* gcov, const double* gcon, const double* detg,
                   const double* x1, const double* x2, const double* x3, const int* nm1, const int* is1,
                   const int* ie1, const int* js1, const int* je1, const int* ks1, const int* ke1) {
  const char* who = "recov";
  if (!need_model(who) || !check_bounds(who, is1, ie1, js1, je1, ks1, ke1)) return;
  (void)x2; (void)x3; (void)nm1;
  if (!check_geom(who, gcov, g.gcv[0], "gcov") || !check_geom(who, gcon, g.gcn[0], "gcon") ||
      !check_geom(who, detg, g.dtg[0], "detg") || !check_geom(who, x1, g.x1, "x1"))
    return;
  if (g.pd.phase != 4 || !g.pd.adsf || uu != g.pd.rk_out || uri != g.pd.uri) {
    fail_order(who, "the stage's RK update and source add first, on its uri");
    return;
  }
  const double* W = input(uri, MK_STATE);
  const double* Ua = input(g.pd.rk_a, MK_STATE);
  const double* Ub = g.pd.rk_b ? input(g.pd.rk_b, MK_STATE) : nullptr;
  const double* F = input(g.pd.ww, MK_WW);
  double* Uout = output(uu, MK_STATE);
  if (!W || !Ua || (g.pd.rk_b && !Ub) || !F || !Uout) return;
  HostDt hd;
  hd.ak1 = g.ak;
  hd.ak2 = g.ak + g.p.imax;
  hd.ak3 = g.ak + g.p.imax + g.p.jmax;
  hd.dt = g.pd.dt;
  if (!stage_update(g.m, g.pd.rk, W, Ua, Ub, Uout, F, &hd)) return;
  g.pd.phase = 5;
}

// bnd.f90:4-5 bnd4(uu,uo,x1,nm1,...,myranki,myrankj,myrankk,icputable): recorded.
void raishin_bnd4(double* uri, const double* urio, const double* x1, const int* nm1, const int* is1, const int* ie1,
                  const int* js1, const int* je1, const int* ks1, const int* ke1, const int* myranki,
                  const int* myrankj, const int* myrankk, const int* icputable) {
  const char* who = "bnd4";
  if (!need_model(who) || !check_bounds(who, is1, ie1, js1, je1, ks1, ke1)) return;
  (void)x1; (void)nm1; (void)myranki; (void)myrankj; (void)myrankk; (void)icputable;
  if (g.pd.phase != 5 || uri != g.pd.uri || g.pd.bnd) { fail_order(who, "recov on the same uri first"); return; }
  g.pd.bnd = true;
  g.pd.urio = urio;
}

// bnd.f90:1636 infchk(uri,gcov,gcon,nm1,...,myranki): recorded.
void raishin_infchk(double* uri, const double* gcov, const double* gcon, const int* nm1, const int* is1,
                    const int* ie1, const int* js1, const int* je1, const int* ks1, const int* ke1,
                    const int* myranki) {
  const char* who = "infchk";
  if (!need_model(who) || !check_bounds(who, is1, ie1, js1, je1, ks1, ke1)) return;
  (void)gcov; (void)gcon; (void)nm1; (void)myranki;
  if (g.pd.phase != 5 || !g.pd.bnd || uri != g.pd.uri || g.pd.inf) { fail_order(who, "bnd4 first"); return; }
  g.pd.inf = true;
}

// correction.f90:226 atmosphere1a(uu,uri,gcov,gcon,detg,x1,...): recorded.
void raishin_atmosphere1a(double* uu, double* uri, const double* gcov, const double* gcon, const double* detg,
                          const double* x1, const int* is1, const int* ie1, const int* js1, const int* je1,
                          const int* ks1, const int* ke1) {
  const char* who = "atmosphere1a";
  if (!need_model(who) || !check_bounds(who, is1, ie1, js1, je1, ks1, ke1)) return;
  (void)uu; (void)gcov; (void)gcon; (void)detg; (void)x1;
  if (g.pd.phase != 5 || !g.pd.bnd || uri != g.pd.uri || g.pd.atm) { fail_order(who, "bnd4 first"); return; }
  g.pd.atm = true;
}

// correction.f90:744 pminmax2b(uu,uri,gcov,gcon,detg,x1,...): recorded.
void raishin_pminmax2b(double* uu, double* uri, const double* gcov, const double* gcon, const double* detg,
                       const double* x1, const int* is1, const int* ie1, const int* js1, const int* je1,
                       const int* ks1, const int* ke1) {
  const char* who = "pminmax2b";
  if (!need_model(who) || !check_bounds(who, is1, ie1, js1, je1, ks1, ke1)) return;
  (void)uu; (void)gcov; (void)gcon; (void)detg; (void)x1;
  if (g.pd.phase != 5 || !g.pd.bnd || uri != g.pd.uri || g.pd.pmm) { fail_order(who, "bnd4 first"); return; }
  g.pd.pmm = true;
}

// caluu.f90:2 caluu1(uri,uu,gcov,gcon,detg,x1,...).  At the end of a stage it
// runs the fused ghost fill + floors + conserved variables; outside a stage
// (the initial call of main.f90:530) it is caluu1 alone.
void raishin_caluu1(double* uri, double* uu, const double* gcov, const double* gcon, const double* detg,
                    const double* x1, const int* is1, const int* ie1, const int* js1, const int* je1, const int* ks1,
                    const int* ke1) {
  const char* who = "caluu1";
  if (!need_model(who) || !check_bounds(who, is1, ie1, js1, je1, ks1, ke1)) return;
  if (!check_geom(who, gcov, g.gcv[0], "gcov") || !check_geom(who, gcon, g.gcn[0], "gcon") ||
      !check_geom(who, detg, g.dtg[0], "detg") || !check_geom(who, x1, g.x1, "x1"))
    return;
  raishin_model* m = g.m;
  if (g.pd.phase == 0) {
    const double* W = input(uri, MK_STATE);
    double* U = output(uu, MK_STATE);
    if (!W || !U) return;
    if (launch_caluu1(m->G, W, m->met[0], g.p.gam, U, m->st) != hipSuccess) set_error("caluu1: launch failed");
    return;
  }
  const bool want_inf = g.p.model >= 7 && g.p.model <= 20, want_atm = g.p.model >= 7 && g.p.model <= 15;
  if (g.pd.phase != 5 || uri != g.pd.uri || !g.pd.bnd || !g.pd.pmm || g.pd.inf != want_inf ||
      g.pd.atm != want_atm) {
    fail_order(who, "bnd4, infchk (model 7..20), atmosphere1a (7..15), pminmax2b first");
    return;
  }
  Mirror* mw = mirror(uri, MK_STATE);
  const double* Wo = m->need_wo ? input(g.pd.urio, MK_STATE) : mw->dev;
  double* U = output(uu, MK_STATE);
  if (!Wo || !U) return;
  if (!stage_post(m, mw->dev, Wo, m->Wb, U)) return;
  std::swap(mw->dev, m->Wb);  // the new primitives become uri's device copy
  mw->dev_valid = true;
  mw->host_valid = false;
  reset_stage();
}

}  // extern "C"
