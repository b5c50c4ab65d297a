// The CFL reduction (cdtcfl4 + MIN) of a resident model with no allocation
// and no host synchronisation: the per-workgroup partial results live in a
// fixed device array of this file instead of buffers the model allocated.
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "rh_model.hpp"

namespace {

// Per-workgroup partials of k_cfl: 4 positions (long long) then 4 minima
// (double) per workgroup, for at most NOALLOC_CFL_BLOCKS workgroups.
constexpr long long NOALLOC_CFL_BLOCKS = 131072;
constexpr size_t SCRATCH_BYTES = (size_t)NOALLOC_CFL_BLOCKS * 4 * (sizeof(long long) + sizeof(double));
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

// The array's address on each device.  On the host `scratch` only names the
// symbol; the device address comes from hipGetSymbolAddress, once per
// device, with that device current.
constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void* g_ptr[MAX_DEVICES] = {};

void* scratch_on(int dev) {
  if (dev < 0 || dev >= MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
    g_ptr[dev] = p;
  }
  return g_ptr[dev];
}

}  // namespace

extern "C" int raishin_model_cdtcfl4_noalloc(raishin_model* m, double* out) {
  using namespace rh;
  if (!m || !out) { rhm::set_error("raishin_model_cdtcfl4_noalloc: null argument"); return -1; }
  if (reinterpret_cast<uintptr_t>(out) % alignof(double)) {
    rhm::set_error("raishin_model_cdtcfl4_noalloc: out is not aligned to 8 bytes");
    return -1;
  }
  const long long nb = cfl_blocks(m->G);
  if (nb > NOALLOC_CFL_BLOCKS) {
    rhm::set_error("raishin_model_cdtcfl4_noalloc: the model needs %lld CFL workgroups (%lld cells), over the "
                   "fixed scratch's %lld; use raishin_model_cdtcfl4",
                   nb, nb * 256, NOALLOC_CFL_BLOCKS);
    return -1;
  }
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev != m->dev) {
    rhm::set_error("raishin_model_cdtcfl4_noalloc: the model lives on device %d, device %d is current", m->dev, dev);
    return -1;
  }
  unsigned char* base = static_cast<unsigned char*>(scratch_on(dev));
  if (!base) { rhm::set_error("raishin_model_cdtcfl4_noalloc: hipGetSymbolAddress failed"); return -1; }
  // every workgroup writes its partials before the finishing workgroup
  // reads them, so nothing needs clearing between calls
  CflArgs C{};
  C.W = m->W; C.metC = m->met[0]; C.dx1b = m->dxb[0]; C.dx2b = m->dxb[1]; C.dx3b = m->dxb[2];
  C.dtbits = reinterpret_cast<unsigned long long*>(out);
  C.blast = reinterpret_cast<long long*>(base);
  C.bmin = reinterpret_cast<double*>(base + (size_t)NOALLOC_CFL_BLOCKS * 4 * sizeof(long long));
  C.gam = m->p.gam;
  if (launch_cfl(m->G, C, m->st) != hipSuccess) {
    rhm::set_error("raishin_model_cdtcfl4_noalloc: launch failed");
    return -1;
  }
  return 0;
}
