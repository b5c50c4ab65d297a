// The reference's restart and analysis files (restar.f90) written and read
// on the host.  Fortran unformatted sequential records: a 4-byte
// little-endian byte count, the payload, the same count again.  The cell
// records are formatted in parallel into a bounded staging buffer and
// written in file order, so a 256x128x128 restart (1.3 GB) streams through
// about 80 MB of host memory.
#include <omp.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/raishin_io.h"

namespace rhm {
void set_error(const char* fmt, ...);  // rh_model.cpp: the sticky error
// rh_capi.cpp: raishin_model_get_state / set_state reporting this call's outcome
bool model_get_state(raishin_model* m, double* uri, double* uu, double* time);
bool model_set_state(raishin_model* m, const double* uri, const double* uu, double time);
}
using rhm::set_error;

namespace {

constexpr int NV = 9;
constexpr long long CHUNK_BYTES = 80LL << 20;

// One record into o: count, payload, count.
inline char* put_rec(char* o, const void* payload, uint32_t n) {
  std::memcpy(o, &n, 4);
  std::memcpy(o + 4, payload, n);
  std::memcpy(o + 4 + n, &n, 4);
  return o + 8 + n;
}

// One record out of p, checking both counts against n.
inline bool get_rec(const char* p, void* payload, uint32_t n) {
  uint32_t a, b;
  std::memcpy(&a, p, 4);
  std::memcpy(&b, p + 4 + n, 4);
  if (a != n || b != n) return false;
  std::memcpy(payload, p + 4, n);
  return true;
}

struct OutFile {
  FILE* f = nullptr;
  std::string path;
  explicit OutFile(const std::string& p) : path(p) { f = std::fopen(p.c_str(), "wb"); }
  ~OutFile() {
    if (f) std::fclose(f);
  }
  bool put(const char* p, size_t n) {
    if (std::fwrite(p, 1, n, f) != n) {
      set_error("write %s: short write", path.c_str());
      return false;
    }
    return true;
  }
  bool close() {
    const int r = std::fclose(f);
    f = nullptr;
    if (r != 0) set_error("write %s: close failed", path.c_str());
    return r == 0;
  }
};

struct InFile {
  FILE* f = nullptr;
  std::string path;
  explicit InFile(const std::string& p) : path(p) { f = std::fopen(p.c_str(), "rb"); }
  ~InFile() {
    if (f) std::fclose(f);
  }
  long long size() {
    std::fseek(f, 0, SEEK_END);
    const long long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    return n;
  }
  bool get(char* p, size_t n) { return std::fread(p, 1, n, f) == n; }
};

bool extents(const int* is1, const int* ie1, const int* js1, const int* je1, const int* ks1, const int* ke1,
             long long& ni, long long& nj, long long& nk) {
  ni = (long long)*ie1 - *is1 + 1;
  nj = (long long)*je1 - *js1 + 1;
  nk = (long long)*ke1 - *ks1 + 1;
  if (ni < 1 || nj < 1 || nk < 1) {
    set_error("empty index range (%d:%d, %d:%d, %d:%d)", *is1, *ie1, *js1, *je1, *ks1, *ke1);
    return false;
  }
  return true;
}

// main.f90:339-343 (unit 9, restartRRR.outdat) and restar.f90:182 (structrRRRR-NNNN.outdat)
std::string restart_name(const char* path, int myrank) {
  if (path) return path;
  char b[64];
  std::snprintf(b, sizeof b, "restart%03d.outdat", myrank);
  return b;
}
std::string output_name(const char* path, int myrank, int nd) {
  if (path) return path;
  char b[64];
  std::snprintf(b, sizeof b, "structr%04d-%04d.outdat", myrank, nd);
  return b;
}

constexpr long long CELL_REC = 4 * (8 + NV * 8);  // store1: four records of nv reals per cell
constexpr long long HEAD_REC = 3 * 8 + 8 + 8 + 8;  // (myrank, nd), (it0, ih0), (time)
constexpr long long OUT_REC = 8 + 13 * 8;          // output1: one record of 13 reals per cell

bool write_restart(const std::string& path, const double* uu, const double* u0, const double* uri,
                   const double* uri0, int it0, int ih0, double time, int nd, int myrank, long long ncell) {
  OutFile out(path);
  if (!out.f) {
    set_error("open %s for writing failed", path.c_str());
    return false;
  }
  char head[HEAD_REC];
  char* o = head;
  const int32_t a[2] = {myrank, nd}, b[2] = {it0, ih0};
  o = put_rec(o, a, 8);
  o = put_rec(o, b, 8);
  o = put_rec(o, &time, 8);
  if (!out.put(head, HEAD_REC)) return false;
  const long long per = CHUNK_BYTES / CELL_REC;
  std::vector<char> buf((size_t)(std::min(per, ncell) * CELL_REC));
  const double* src[4] = {uu, u0, uri, uri0};
  for (long long c0 = 0; c0 < ncell; c0 += per) {
    const long long n = std::min(per, ncell - c0);
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < n; c++) {
      char* q = buf.data() + c * CELL_REC;
      for (int s = 0; s < 4; s++) q = put_rec(q, src[s] + (c0 + c) * NV, NV * 8);
    }
    if (!out.put(buf.data(), (size_t)(n * CELL_REC))) return false;
  }
  return out.close();
}

bool read_restart(const std::string& path, double* uu, double* u0, double* uri, double* uri0, int* it0, int* ih0,
                  double* time, int* nd, int myrank, long long ncell) {
  InFile in(path);
  if (!in.f) {
    set_error("open %s for reading failed", path.c_str());
    return false;
  }
  const long long want = HEAD_REC + ncell * CELL_REC;
  const long long have = in.size();
  if (have != want) {
    set_error("%s: %lld bytes, a restart file of %lld cells has %lld", path.c_str(), have, ncell, want);
    return false;
  }
  char head[HEAD_REC];
  int32_t a[2], b[2];
  double t;
  if (!in.get(head, HEAD_REC) || !get_rec(head, a, 8) || !get_rec(head + 16, b, 8) || !get_rec(head + 32, &t, 8)) {
    set_error("%s: header records are not (2 integers), (2 integers), (1 real)", path.c_str());
    return false;
  }
  if (a[0] != myrank) {  // restar.f90:323-326
    set_error("%s was written by rank %d, read by rank %d", path.c_str(), a[0], myrank);
    return false;
  }
  const long long per = CHUNK_BYTES / CELL_REC;
  std::vector<char> buf((size_t)(std::min(per, ncell) * CELL_REC));
  double* dst[4] = {uu, u0, uri, uri0};
  for (long long c0 = 0; c0 < ncell; c0 += per) {
    const long long n = std::min(per, ncell - c0);
    if (!in.get(buf.data(), (size_t)(n * CELL_REC))) {
      set_error("%s: read failed at cell %lld", path.c_str(), c0);
      return false;
    }
    long long bad = -1;
#pragma omp parallel for schedule(static) reduction(max : bad)
    for (long long c = 0; c < n; c++) {
      const char* q = buf.data() + c * CELL_REC;
      for (int s = 0; s < 4; s++) {
        double tmp[NV];
        if (!get_rec(q, tmp, NV * 8)) bad = std::max(bad, c0 + c);
        else if (dst[s]) std::memcpy(dst[s] + (c0 + c) * NV, tmp, NV * 8);
        q += 8 + NV * 8;
      }
    }
    if (bad >= 0) {
      set_error("%s: cell %lld: a record is not %d reals", path.c_str(), bad, NV);
      return false;
    }
  }
  if (it0) *it0 = b[0];
  if (ih0) *ih0 = b[1];
  if (nd) *nd = a[1];
  if (time) *time = t;
  return true;
}

bool write_output(const std::string& path, const double* uri, const double* detg, const double* x1,
                  const double* x2, const double* x3, double time, int myrank, int npe, long long ni, long long nj,
                  long long nk, const int lo[3], const int hi[3]) {
  OutFile out(path);
  if (!out.f) {
    set_error("open %s for writing failed", path.c_str());
    return false;
  }
  char head[8 + 16 + 8 + 24];
  char* o = head;
  char h1[16];
  const int32_t r[2] = {myrank, npe};
  std::memcpy(h1, &time, 8);
  std::memcpy(h1 + 8, r, 8);
  o = put_rec(o, h1, 16);
  const int32_t ix[6] = {lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]};
  o = put_rec(o, ix, 24);
  if (!out.put(head, sizeof head)) return false;
  // ranges as offsets into the arrays (lo/hi are in the caller's is1.. numbering)
  const long long i0 = 0, i1 = ni, j0 = 0, j1 = nj, k0 = 0, k1 = nk;
  const long long per_i = (j1 - j0) * (k1 - k0);
  const long long iper = std::max(1LL, CHUNK_BYTES / (per_i * OUT_REC));
  std::vector<char> buf((size_t)(std::min(iper, i1 - i0) * per_i * OUT_REC));
  for (long long ib = i0; ib < i1; ib += iper) {
    const long long ni_ = std::min(iper, i1 - ib);
#pragma omp parallel for collapse(2) schedule(static)
    for (long long ii = 0; ii < ni_; ii++)
      for (long long j = j0; j < j1; j++) {
        const long long i = ib + ii;
        char* q = buf.data() + ((ii * (j1 - j0)) + (j - j0)) * (k1 - k0) * OUT_REC;
        for (long long k = k0; k < k1; k++) {
          double rec[13];
          const long long c = i + ni * (j + nj * k);
          rec[0] = x1[i];
          rec[1] = x2[j];
          rec[2] = x3[k];
          for (int n = 0; n < NV; n++) rec[3 + n] = uri[c * NV + n];
          rec[12] = detg[c];
          q = put_rec(q, rec, 13 * 8);
        }
      }
    if (!out.put(buf.data(), (size_t)(ni_ * per_i * OUT_REC))) return false;
  }
  return out.close();
}

}  // namespace

extern "C" {

void raishin_store1(const char* path, const double* uu, const double* u0, const double* uri, const double* uri0,
                    const int* it0, const int* ih0, const double* time, const int* nd, const int* npe,
                    const int* myrank, const int* is1, const int* ie1, const int* js1, const int* je1,
                    const int* ks1, const int* ke1) {
  (void)npe;
  if (!uu || !u0 || !uri || !uri0 || !it0 || !ih0 || !time || !nd || !myrank || !is1 || !ie1 || !js1 || !je1 ||
      !ks1 || !ke1) {
    set_error("raishin_store1: null argument");
    return;
  }
  long long ni, nj, nk;
  if (!extents(is1, ie1, js1, je1, ks1, ke1, ni, nj, nk)) return;
  write_restart(restart_name(path, *myrank), uu, u0, uri, uri0, *it0, *ih0, *time, *nd, *myrank, ni * nj * nk);
}

void raishin_restar2(const char* path, double* uu, double* u0, double* uri, double* uri0, int* it0, int* ih0,
                     double* time, int* nd, const int* npe, const int* myrank, const int* is1, const int* ie1,
                     const int* js1, const int* je1, const int* ks1, const int* ke1) {
  (void)npe;
  if (!uu || !uri || !myrank || !is1 || !ie1 || !js1 || !je1 || !ks1 || !ke1) {
    set_error("raishin_restar2: null argument");
    return;
  }
  long long ni, nj, nk;
  if (!extents(is1, ie1, js1, je1, ks1, ke1, ni, nj, nk)) return;
  read_restart(restart_name(path, *myrank), uu, u0, uri, uri0, it0, ih0, time, nd, *myrank, ni * nj * nk);
}

void raishin_output1(const char* path, const double* uri, const double* detg, const double* x1, const double* x2,
                     const double* x3, const double* time, const int* nm1, const int* nd, const int* npe,
                     const int* myrank, const int* myranki, const int* myrankj, const int* myrankk,
                     const int* is1, const int* ie1, const int* js1, const int* je1, const int* ks1,
                     const int* ke1) {
  (void)nm1;
  if (!uri || !detg || !x1 || !x2 || !x3 || !time || !nd || !npe || !myrank || !myranki || !myrankj || !myrankk ||
      !is1 || !ie1 || !js1 || !je1 || !ks1 || !ke1) {
    set_error("raishin_output1: null argument");
    return;
  }
  long long ni, nj, nk;
  if (!extents(is1, ie1, js1, je1, ks1, ke1, ni, nj, nk)) return;
  if (*myranki != 0 || *myrankj != 0 || *myrankk != 0) {
    set_error("output1: one rank only (myranki, myrankj, myrankk = %d, %d, %d)", *myranki, *myrankj, *myrankk);
    return;
  }
  // restar.f90:124-150 with iprocs = jprocs = kprocs = 1: the whole array, ghosts included
  const int lo[3] = {*is1, *js1, *ks1}, hi[3] = {*ie1, *je1, *ke1};
  write_output(output_name(path, *myrank, *nd), uri, detg, x1, x2, x3, *time, *myrank, *npe, ni, nj, nk, lo, hi);
}

int raishin_model_store1(raishin_model* m, const char* path, const double* u0, const double* uri0, int it0,
                         int ih0, int nd, int myrank) {
  if (!m) { set_error("raishin_model_store1: null model"); return -1; }
  raishin_pram p;
  raishin_model_get_pram(m, &p);
  const long long n = (long long)p.imax * p.jmax * p.kmax;
  std::vector<double> uri(n * NV), uu(n * NV);
  double time = 0.0;
  if (!rhm::model_get_state(m, uri.data(), uu.data(), &time)) return -1;
  return write_restart(restart_name(path, myrank), uu.data(), u0 ? u0 : uu.data(), uri.data(),
                       uri0 ? uri0 : uri.data(), it0, ih0, time, nd, myrank, n)
             ? 0
             : -1;
}

int raishin_model_restar2(raishin_model* m, const char* path, double* u0, double* uri0, int* it0, int* ih0,
                          int* nd, int myrank) {
  if (!m) { set_error("raishin_model_restar2: null model"); return -1; }
  raishin_pram p;
  raishin_model_get_pram(m, &p);
  const long long n = (long long)p.imax * p.jmax * p.kmax;
  std::vector<double> uri(n * NV), uu(n * NV);
  double time = 0.0;
  if (!read_restart(restart_name(path, myrank), uu.data(), u0, uri.data(), uri0, it0, ih0, &time, nd, myrank, n))
    return -1;
  return rhm::model_set_state(m, uri.data(), uu.data(), time) ? 0 : -1;
}

int raishin_model_output1(raishin_model* m, const char* path, const double* detg, const double* x1,
                          const double* x2, const double* x3, int nd, int myrank, int npe) {
  if (!m || !detg || !x1 || !x2 || !x3) { set_error("raishin_model_output1: null argument"); return -1; }
  raishin_pram p;
  raishin_model_get_pram(m, &p);
  const long long n = (long long)p.imax * p.jmax * p.kmax;
  std::vector<double> uri(n * NV), uu(n * NV);
  double time = 0.0;
  if (!rhm::model_get_state(m, uri.data(), uu.data(), &time)) return -1;
  const int lo[3] = {1, 1, 1}, hi[3] = {p.imax, p.jmax, p.kmax};
  return write_output(output_name(path, myrank, nd), uri.data(), detg, x1, x2, x3, time, myrank, npe, p.imax,
                      p.jmax, p.kmax, lo, hi)
             ? 0
             : -1;
}

}  // extern "C"
