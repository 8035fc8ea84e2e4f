// Optional per-kernel timing with HIP events recorded on the launch stream.
// Disabled unless aa_profile_start() is called; used by the benchmark to
// time the dominant kernel inside the timed region.
#include <vector>

#include "aa_common.h"
#include "prof.h"
#include "../../include/aeroapps_amd.h"

namespace aa {
namespace {
struct Span { hipEvent_t a, b; };
struct Prof {
  bool on = false;
  std::vector<Span> spans[PROF_KINDS];
  std::vector<hipEvent_t> pool;
  size_t used = 0;
} g_prof;

hipEvent_t take_event() {
  if (g_prof.used == g_prof.pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    g_prof.pool.push_back(e);
  }
  return g_prof.pool[g_prof.used++];
}
}  // namespace

bool prof_on() { return g_prof.on; }

void prof_begin(int kind, hipStream_t st) {
  if (!g_prof.on) return;
  Span s{take_event(), nullptr};
  if (s.a) (void)hipEventRecord(s.a, st);
  g_prof.spans[kind].push_back(s);
}

void prof_end(int kind, hipStream_t st) {
  if (!g_prof.on || g_prof.spans[kind].empty()) return;
  Span& s = g_prof.spans[kind].back();
  s.b = take_event();
  if (s.b) (void)hipEventRecord(s.b, st);
}
}  // namespace aa

extern "C" {

int aa_profile_start(void) {
  for (auto& v : aa::g_prof.spans) v.clear();
  aa::g_prof.used = 0;
  aa::g_prof.on = true;
  return 0;
}

int aa_profile_read(double* ms, long long* launches, int nkinds) {
  if (!ms || !launches || nkinds <= 0) return AA_EARG;
  AA_CHECK(hipDeviceSynchronize());
  for (int k = 0; k < nkinds; ++k) {
    ms[k] = 0.0; launches[k] = 0;
    if (k >= aa::PROF_KINDS) continue;
    for (auto& s : aa::g_prof.spans[k]) {
      if (!s.a || !s.b) continue;
      float t = 0.0f;
      AA_CHECK(hipEventElapsedTime(&t, s.a, s.b));
      ms[k] += t;
      launches[k] += 1;
    }
  }
  aa::g_prof.on = false;
  return 0;
}

}  // extern "C"
