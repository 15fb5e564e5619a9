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
};

struct Prof {
  bool on = false;
  size_t next = 0;
  std::vector<Rec> recs;
  bool overflow = false;
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
  for (size_t i = 0; i < g_prof.next; ++i) {
    Rec& r = g_prof.recs[i];
    if (hipEventSynchronize(r.b) != hipSuccess) rc = SINV_E_HIP;
    float t = 0.f;
    if (hipEventElapsedTime(&t, r.a, r.b) != hipSuccess) rc = SINV_E_HIP;
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
