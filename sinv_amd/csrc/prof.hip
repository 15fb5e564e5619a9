// Launch-level profiler: while enabled, each GEMM / leaf-inverse launch is
// bracketed by a pair of hipEvents recorded on the launch's own stream, and
// its algorithmic flops and bytes are tallied per kernel kind.  Used to measure the
// dominant kernel's average duration without an external profiler.
// Process-wide state: enable it around a single-threaded region only.
#include <vector>

#include "sinv_core.hpp"

namespace sinv {

namespace {

struct Rec {
  hipEvent_t a, b;
  int kind;
  double flops;
  double bytes;
  long long dims[5];  // GEMM: M, N, K, batch, flags (opA | opB << 2 | has C0 << 4)
};

struct Last {
  int kind;
  long long dims[5];
  float ms;
};

struct Prof {
  bool on = false;
  size_t next = 0;
  std::vector<Rec> recs;
  bool overflow = false;
  std::vector<Last> last;  // per-launch records of the last completed region
};

Prof g_prof;

}  // namespace

int prof_open(hipStream_t s) {
  if (!g_prof.on) return -1;
  if (g_prof.next >= g_prof.recs.size()) {
    g_prof.overflow = true;
    return -1;
  }
  const int id = (int)g_prof.next++;
  (void)hipEventRecord(g_prof.recs[id].a, s);
  return id;
}

void prof_shape(int id, int M, int N, int K, long long batch, int flags) {
  if (id < 0) return;
  long long* d = g_prof.recs[id].dims;
  d[0] = M;
  d[1] = N;
  d[2] = K;
  d[3] = batch;
  d[4] = flags;
}

void prof_close(int id, hipStream_t s, int kind, double flops, double bytes) {
  if (id < 0) return;
  (void)hipEventRecord(g_prof.recs[id].b, s);
  g_prof.recs[id].kind = kind;
  g_prof.recs[id].flops = flops;
  g_prof.recs[id].bytes = bytes;
}

}  // namespace sinv

using namespace sinv;

extern "C" int sinv_prof_begin(int max_launches) {
  if (max_launches <= 0 || g_prof.on) {
    set_error("sinv_prof_begin: bad capacity or already active");
    return SINV_E_INVAL;
  }
  g_prof.recs.resize((size_t)max_launches);
  for (auto& r : g_prof.recs) {
    SINV_HIP_TRY(hipEventCreate(&r.a));
    SINV_HIP_TRY(hipEventCreate(&r.b));
    r.kind = -1;
    r.flops = 0.0;
    r.bytes = 0.0;
    for (long long& x : r.dims) x = 0;
  }
  g_prof.next = 0;
  g_prof.overflow = false;
  g_prof.on = true;
  return SINV_OK;
}

extern "C" int sinv_prof_end(double* ms, long long* launches, double* flops, double* bytes,
                             int nkinds) {
  if (!g_prof.on) {
    set_error("sinv_prof_end: profiler not active");
    return SINV_E_INVAL;
  }
  g_prof.on = false;
  for (int k = 0; k < nkinds; ++k) {
    ms[k] = 0.0;
    launches[k] = 0;
    flops[k] = 0.0;
    if (bytes != nullptr) bytes[k] = 0.0;
  }
  int rc = SINV_OK;
  g_prof.last.clear();
  for (size_t i = 0; i < g_prof.next; ++i) {
    Rec& r = g_prof.recs[i];
    if (hipEventSynchronize(r.b) != hipSuccess) rc = SINV_E_HIP;
    float t = 0.f;
    if (hipEventElapsedTime(&t, r.a, r.b) != hipSuccess) rc = SINV_E_HIP;
    Last l{r.kind, {r.dims[0], r.dims[1], r.dims[2], r.dims[3], r.dims[4]}, t};
    g_prof.last.push_back(l);
    if (r.kind >= 0 && r.kind < nkinds) {
      ms[r.kind] += t;
      launches[r.kind] += 1;
      flops[r.kind] += r.flops;
      if (bytes != nullptr) bytes[r.kind] += r.bytes;
    }
  }
  for (auto& r : g_prof.recs) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  g_prof.recs.clear();
  if (g_prof.overflow) {
    set_error("sinv_prof_end: more launches than capacity; totals cover the first ones");
    return SINV_E_INVAL;
  }
  if (rc != SINV_OK) set_error("sinv_prof_end: event query failed");
  return rc;
}

extern "C" int sinv_prof_records(int cap, int* kind, long long* dims, float* ms) {
  const int n = (int)g_prof.last.size();
  for (int i = 0; i < n && i < cap; ++i) {
    kind[i] = g_prof.last[i].kind;
    for (int j = 0; j < 5; ++j) dims[5 * i + j] = g_prof.last[i].dims[j];
    ms[i] = g_prof.last[i].ms;
  }
  return n;
}
