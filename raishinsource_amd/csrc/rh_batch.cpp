// Batches of resident models: the same step of many models (of any size,
// sharing irec and irkt) made as one launch per kernel of the step.
//
// A plan records each member's launches for two consecutive steps (the
// launch helpers append records instead of launching while a recorder is
// set), checks that the members make the same sequence of kernels, and merges
// the records position by position into batched launches: one device table
// of BatchItem per launch, each member owning a contiguous range of the
// grid's workgroups.  Members whose variant differs at a position (FP64 and
// mixed sweeps, symmetric and general Christoffel tables) get one launch per
// variant.  Each step swaps buffer pointers on the host; the plan keeps the
// members' pointers at the start of each of its two steps and sets them as a
// model-by-model run_step would, so the members stay usable one by one.
#include <cstring>
#include <vector>

#include "rh_model.hpp"

using namespace rh;
using namespace rhm;

namespace {

struct Ptrs {
  double *W, *Wb, *Ucur, *Unext;
};

Ptrs ptrs_of(const raishin_model* m) { return Ptrs{m->W, m->Wb, m->Ucur, m->Unext}; }
void set_ptrs(raishin_model* m, const Ptrs& p) {
  m->W = p.W; m->Wb = p.Wb; m->Ucur = p.Ucur; m->Unext = p.Unext;
}
bool same_ptrs(const Ptrs& a, const Ptrs& b) {
  return a.W == b.W && a.Wb == b.Wb && a.Ucur == b.Ucur && a.Unext == b.Unext;
}

struct Op {
  int kind, variant, threads;
  long long item0;  // first entry in the device table
  int nitems;
  long long blocks;
};

}  // namespace

struct raishin_batch {
  std::vector<raishin_model*> ms;
  hipStream_t st = nullptr;
  hipEvent_t ev = nullptr;
  // plan
  bool planned = false;
  std::vector<Op> half[2];        // the launches of a step starting at parity p
  std::vector<Ptrs> at[2];        // each member's pointers at the start of half p
  std::vector<bool> at_mixed;
  std::vector<unsigned> at_gen;   // each member's geometry generation
  std::vector<BatchItem> host_items;
  BatchItem* items = nullptr;
  size_t items_cap = 0;
  int parity = 0;
  // time gather
  const double** tptr = nullptr;  // device: each member's tdt
  double* tdev = nullptr;
  double* thost = nullptr;        // pinned
  int tcap = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
};

namespace {

void drop_graph(raishin_batch* b) {
  if (b->gexec) { (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
  if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
}

bool plan_valid(const raishin_batch* b) {
  if (!b->planned) return false;
  for (size_t q = 0; q < b->ms.size(); q++)
    if (!same_ptrs(ptrs_of(b->ms[q]), b->at[b->parity][q]) || b->ms[q]->mixed != b->at_mixed[q] ||
        b->ms[q]->geo_gen != b->at_gen[q])
      return false;
  return true;
}

// Record two steps of every member and merge them into batched launches.
bool make_plan(raishin_batch* b) {
  drop_graph(b);
  b->planned = false;
  const size_t n = b->ms.size();
  std::vector<std::vector<LaunchRec>> rec[2];
  rec[0].resize(n);
  rec[1].resize(n);
  b->at[0].resize(n);
  b->at[1].resize(n);
  b->at_mixed.resize(n);
  b->at_gen.resize(n);
  for (size_t q = 0; q < n; q++) {
    raishin_model* m = b->ms[q];
    const unsigned timing = m->timing;
    m->timing = 0;
    bool ok = true;
    for (int h = 0; h < 2 && ok; h++) {
      b->at[h][q] = ptrs_of(m);
      set_recorder(&rec[h][q]);
      ok = run_step(m);
      set_recorder(nullptr);
    }
    m->timing = timing;
    if (!ok) return false;
    if (!same_ptrs(ptrs_of(m), b->at[0][q])) { set_error("raishin_batch: a member's buffers did not return after two steps"); return false; }
    b->at_mixed[q] = m->mixed;
    b->at_gen[q] = m->geo_gen;
  }
  b->host_items.clear();
  for (int h = 0; h < 2; h++) {
    b->half[h].clear();
    const size_t len = rec[h][0].size();
    for (size_t q = 1; q < n; q++)
      if (rec[h][q].size() != len) {
        set_error("raishin_batch: members make different step sequences (irec, irkt, the boundary types and the PPM sweep form (split or fused) must agree)");
        return false;
      }
    for (size_t pos = 0; pos < len; pos++) {
      const int kind = rec[h][0][pos].kind;
      for (size_t q = 1; q < n; q++)
        if (rec[h][q][pos].kind != kind) {
          set_error("raishin_batch: members make different step sequences (irec, irkt, the boundary types and the PPM sweep form (split or fused) must agree)");
          return false;
        }
      // one launch per (variant, workgroup size) present at this position,
      // in order of first appearance
      std::vector<std::pair<int, int>> variants;
      for (size_t q = 0; q < n; q++) {
        const std::pair<int, int> v(rec[h][q][pos].variant, rec[h][q][pos].threads);
        bool seen = false;
        for (const auto& x : variants) seen = seen || x == v;
        if (!seen) variants.push_back(v);
      }
      const bool per_model = one_group_per_model(kind);
      for (const auto& v : variants) {
        Op op{kind, v.first, v.second, (long long)b->host_items.size(), 0, 0};
        for (size_t q = 0; q < n; q++) {
          const LaunchRec& r = rec[h][q][pos];
          if (r.variant != v.first || r.threads != v.second || r.nb <= 0) continue;
          BatchItem it;
          std::memset(&it, 0, sizeof it);
          it.G = r.G;
          it.A = r.A;
          it.off = per_model ? (long long)op.nitems : op.blocks;
          b->host_items.push_back(it);
          op.nitems++;
          op.blocks += per_model ? 1 : r.nb;
        }
        if (op.nitems > 0) b->half[h].push_back(op);
      }
    }
  }
  // the device table; earlier launches may still read the old one
  const size_t need = b->host_items.size();
  if (need > b->items_cap) {
    HIPCHK(hipStreamSynchronize(b->st));
    if (b->items) HIPCHK(hipFree(b->items));
    b->items = nullptr;
    b->items_cap = 0;
    HIPCHK(hipMalloc((void**)&b->items, need * sizeof(BatchItem)));
    b->items_cap = need;
  } else {
    HIPCHK(hipStreamSynchronize(b->st));
  }
  HIPCHK(hipMemcpyAsync(b->items, b->host_items.data(), need * sizeof(BatchItem), hipMemcpyHostToDevice, b->st));
  HIPCHK(hipStreamSynchronize(b->st));
  // time gather table
  if ((int)n > b->tcap) {
    if (b->tptr) HIPCHK(hipFree((void*)b->tptr));
    if (b->tdev) HIPCHK(hipFree(b->tdev));
    if (b->thost) HIPCHK(hipHostFree(b->thost));
    b->tptr = nullptr; b->tdev = nullptr; b->thost = nullptr; b->tcap = 0;
    HIPCHK(hipMalloc((void**)&b->tptr, n * sizeof(double*)));
    HIPCHK(hipMalloc((void**)&b->tdev, n * sizeof(double)));
    HIPCHK(hipHostMalloc((void**)&b->thost, n * sizeof(double), hipHostMallocDefault));
    b->tcap = (int)n;
  }
  std::vector<const double*> tp(n);
  for (size_t q = 0; q < n; q++) tp[q] = b->ms[q]->tdt;
  if (n) {
    HIPCHK(hipMemcpyAsync((void*)b->tptr, tp.data(), n * sizeof(double*), hipMemcpyHostToDevice, b->st));
    HIPCHK(hipStreamSynchronize(b->st));
  }
  b->parity = 0;
  b->planned = true;
  return true;
}

// the batch stream waits for work already queued on the members' streams
bool join_members(raishin_batch* b) {
  for (raishin_model* m : b->ms) {
    HIPCHK(hipEventRecord(b->ev, m->st));
    HIPCHK(hipStreamWaitEvent(b->st, b->ev, 0));
  }
  return true;
}

// later work on the members' streams waits for the batch
bool release_members(raishin_batch* b) {
  HIPCHK(hipEventRecord(b->ev, b->st));
  for (raishin_model* m : b->ms) HIPCHK(hipStreamWaitEvent(m->st, b->ev, 0));
  return true;
}

bool run_half(raishin_batch* b) {
  const int h = b->parity;
  for (const Op& op : b->half[h])
    HIPCHK(launch_batched(op.kind, op.variant, op.threads, b->items + op.item0, op.nitems, op.blocks, b->st));
  b->parity ^= 1;
  for (size_t q = 0; q < b->ms.size(); q++) set_ptrs(b->ms[q], b->at[b->parity][q]);
  return true;
}

bool ready(raishin_batch* b) {
  if (b->ms.empty()) return true;
  if (plan_valid(b)) return true;
  return make_plan(b);
}

}  // namespace

extern "C" {

raishin_batch* raishin_batch_create(void) {
  raishin_batch* b = new raishin_batch();
  if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev, hipEventDisableTiming) != hipSuccess) {
    set_error("raishin_batch_create: stream or event creation failed");
    if (b->st) (void)hipStreamDestroy(b->st);
    delete b;
    return nullptr;
  }
  return b;
}

void raishin_batch_destroy(raishin_batch* b) {
  if (!b) return;
  (void)hipStreamSynchronize(b->st);
  drop_graph(b);
  if (b->items) (void)hipFree(b->items);
  if (b->tptr) (void)hipFree((void*)b->tptr);
  if (b->tdev) (void)hipFree(b->tdev);
  if (b->thost) (void)hipHostFree(b->thost);
  (void)hipEventDestroy(b->ev);
  (void)hipStreamDestroy(b->st);
  delete b;
}

int raishin_batch_set(raishin_batch* b, raishin_model* const* models, int n) {
  if (!b || n < 0 || (n > 0 && !models)) { set_error("raishin_batch_set: bad arguments"); return -1; }
  for (int q = 0; q < n; q++)
    if (!models[q]) { set_error("raishin_batch_set: null member"); return -1; }
  for (int q = 0; q < n; q++)
    for (int r = q + 1; r < n; r++)
      if (models[q] == models[r]) { set_error("raishin_batch_set: a model appears twice"); return -1; }
  if (hipStreamSynchronize(b->st) != hipSuccess) { set_error("raishin_batch_set: synchronisation failed"); return -1; }
  drop_graph(b);
  b->ms.assign(models, models + n);
  b->planned = false;
  return ready(b) ? 0 : -1;
}

int raishin_batch_step(raishin_batch* b, int nsteps) {
  if (!b || nsteps < 0) { set_error("raishin_batch_step: bad arguments"); return -1; }
  if (b->ms.empty() || nsteps == 0) return 0;
  if (!ready(b) || !join_members(b)) return -1;
  for (int s = 0; s < nsteps; s++)
    if (!run_half(b)) return -1;
  return release_members(b) ? 0 : -1;
}

int raishin_batch_step_graph(raishin_batch* b, int npairs) {
  if (!b || npairs < 0) { set_error("raishin_batch_step_graph: bad arguments"); return -1; }
  if (b->ms.empty() || npairs == 0) return 0;
  if (!ready(b) || !join_members(b)) return -1;
  auto go = [&]() -> bool {
    // from the plan's second step: one step to its start, whole pairs from
    // the graph, one step at the end (2*npairs steps in all)
    const bool odd = b->parity == 1;
    if (odd) {
      if (!run_half(b)) return false;
      npairs -= 1;
    }
    if (!b->gexec) {
      HIPCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
      bool ok = true;
      for (int h = 0; h < 2 && ok; h++)
        for (const Op& op : b->half[h])
          if (launch_batched(op.kind, op.variant, op.threads, b->items + op.item0, op.nitems, op.blocks, b->st) != hipSuccess) {
            ok = false;
            break;
          }
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(b->st, &g);
      if (!ok || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        set_error("raishin_batch_step_graph: capture failed");
        return false;
      }
      if (hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(g);
        set_error("raishin_batch_step_graph: instantiation failed");
        return false;
      }
      b->graph = g;
    }
    for (int q = 0; q < npairs; q++) HIPCHK(hipGraphLaunch(b->gexec, b->st));
    return odd ? run_half(b) : true;
  };
  if (!go()) return -1;
  return release_members(b) ? 0 : -1;
}

int raishin_batch_times(raishin_batch* b, double* times) {
  if (!b || (!times && !b->ms.empty())) { set_error("raishin_batch_times: bad arguments"); return -1; }
  const int n = (int)b->ms.size();
  if (n == 0) return 0;
  if (!ready(b) || !join_members(b)) return -1;
  auto go = [&]() -> bool {
    HIPCHK(launch_gather_first(b->tptr, b->tdev, n, b->st));
    HIPCHK(hipMemcpyAsync(b->thost, b->tdev, n * sizeof(double), hipMemcpyDeviceToHost, b->st));
    HIPCHK(hipStreamSynchronize(b->st));
    std::memcpy(times, b->thost, n * sizeof(double));
    return true;
  };
  return go() ? 0 : -1;
}

int raishin_batch_sync(raishin_batch* b) {
  if (!b) return -1;
  if (hipStreamSynchronize(b->st) != hipSuccess) { set_error("raishin_batch_sync: synchronisation failed"); return -1; }
  return 0;
}

}  // extern "C"
