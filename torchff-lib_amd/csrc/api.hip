// Error reporting, ABI version and optional per-kernel event timing for the torchff gfx950 library.

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "tff_common.h"
#include "torchff_hip.h"

namespace {
thread_local char g_err[1024] = "";

struct ProfRec {
  std::string name;
  hipEvent_t a, b;
};
std::mutex g_prof_mu;
bool g_prof_on = false;
std::vector<ProfRec> g_prof;
std::vector<hipEvent_t> g_pool;

hipEvent_t pool_get() {
  if (!g_pool.empty()) {
    hipEvent_t e = g_pool.back();
    g_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}
}  // namespace

void tff_set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

void* tff_prof_begin(const char* name, hipStream_t stream) {
  if (!g_prof_on) return nullptr;
  std::lock_guard<std::mutex> lk(g_prof_mu);
  hipEvent_t a = pool_get(), b = pool_get();
  if (!a || !b) return nullptr;
  hipEventRecord(a, stream);
  g_prof.push_back({name, a, b});
  return (void*)(g_prof.size());
}

void tff_prof_end(void* token, hipStream_t stream) {
  if (!token) return;
  std::lock_guard<std::mutex> lk(g_prof_mu);
  size_t k = (size_t)token - 1;
  if (k < g_prof.size()) hipEventRecord(g_prof[k].b, stream);
}

bool tff_prof_active() {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  return g_prof_on;
}

extern "C" const char* tff_last_error(void) { return g_err; }

extern "C" int tff_abi_version(void) { return 2; }

extern "C" void tff_prof_enable(int on) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_prof_on = on != 0;
}

extern "C" int tff_prof_collect(const char* name, double* total_ms, long long* count) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  double tot = 0.0;
  long long n = 0;
  std::vector<ProfRec> keep;
  for (auto& r : g_prof) {
    if (name && r.name != name) {
      keep.push_back(r);
      continue;
    }
    float ms = 0.f;
    if (hipEventSynchronize(r.b) != hipSuccess || hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess) {
      tff_set_error("tff_prof_collect: event query failed");
      return TFF_ERR_LAUNCH;
    }
    tot += ms;
    ++n;
    g_pool.push_back(r.a);
    g_pool.push_back(r.b);
  }
  g_prof.swap(keep);
  if (total_ms) *total_ms = tot;
  if (count) *count = n;
  return TFF_OK;
}
