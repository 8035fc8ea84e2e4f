// Launch-shape and schedule table of the Fourier-term kernels, per
// (streams, Stokes, precision) class: the band solver's lanes per item, the
// SAB/DAB kernel's pair groups per wavefront and the work-chunk size.  The
// tuner (aeroapps_amd/tune.py) measures the candidates on the device and
// writes the table as text; the library reads tuning_gfx950.txt beside the
// shared library on first use, or the file aa_load_tuning names.  Every entry
// gives the same results as the defaults: the shapes only change which lanes
// and wavefronts do an item's work, never the operations on it.  Host code.
#include <dlfcn.h>

#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <tuple>

#include "../../include/aeroapps_amd.h"
#include "rt_work.h"

namespace aa {

namespace {

std::mutex g_mu;
std::map<std::tuple<int, int, int>, TuneEntry> g_table;
bool g_loaded = false;

std::string default_path() {
  Dl_info info;
  if (dladdr((void*)&aa_set_tuning, &info) == 0 || !info.dli_fname) return "";
  std::string p(info.dli_fname);
  const size_t k = p.rfind('/');
  return (k == std::string::npos ? std::string(".") : p.substr(0, k)) + "/tuning_gfx950.txt";
}

// entry of an aa_tuning for (N, S), or false when a field is not a shape
// this build has
bool to_entry(int N, int S, const aa_tuning& t, TuneEntry& e) {
  const int n = N * S;
  e = TuneEntry{};
  if (t.bvp_lanes != 0 && t.bvp_lanes != bvp_group(n)) {
    if (t.bvp_lanes != bvp_group_alt(n)) return false;
    e.shape_bvp = 1;
  }
  if (t.sab_pairs != 0 && t.sab_pairs != 16 * sab_groups(N, S)) {
    if (t.sab_pairs != 16 * sab_groups_alt(N, S)) return false;
    e.shape_sab = 1;
  }
  if (t.chunk_items < 0) return false;
  e.chunk_items = t.chunk_items;
  return true;
}

// Lines "nstreams nstokes precision bvp_lanes sab_pairs chunk_items", '#'
// comments.  Returns the number of entries read, -1 if the file cannot be
// read, -2 on a malformed line.
int load_locked(const std::string& path) {
  std::ifstream f(path);
  if (!f) return -1;
  std::map<std::tuple<int, int, int>, TuneEntry> t;
  std::string line;
  while (std::getline(f, line)) {
    const size_t h = line.find('#');
    if (h != std::string::npos) line.resize(h);
    std::istringstream is(line);
    int N, S, P;
    aa_tuning v{};
    if (!(is >> N)) continue;   // blank line
    if (!(is >> S >> P >> v.bvp_lanes >> v.sab_pairs >> v.chunk_items)) return -2;
    TuneEntry e;
    if (N < 1 || (S != 1 && S != 3) || (P != 0 && P != 1) || !to_entry(N, S, v, e)) return -2;
    t[std::make_tuple(N, S, P)] = e;
  }
  g_table.swap(t);
  return (int)g_table.size();
}

void ensure_loaded() {
  if (g_loaded) return;
  g_loaded = true;
  const std::string p = default_path();
  if (!p.empty()) (void)load_locked(p);   // no file: defaults
}

}  // namespace

TuneEntry tuning_for(int N, int S, int precision) {
  std::lock_guard<std::mutex> lk(g_mu);
  ensure_loaded();
  auto it = g_table.find(std::make_tuple(N, S, precision));
  return it == g_table.end() ? TuneEntry{} : it->second;
}

}  // namespace aa

extern "C" {

int aa_set_tuning(int nstreams, int nstokes, int precision, const aa_tuning* t) {
  using namespace aa;
  if (nstreams < 1 || (nstokes != 1 && nstokes != 3) || (precision != 0 && precision != 1)) return AA_EARG;
  std::lock_guard<std::mutex> lk(g_mu);
  ensure_loaded();
  const auto key = std::make_tuple(nstreams, nstokes, precision);
  if (!t) { g_table.erase(key); return 0; }
  TuneEntry e;
  if (!to_entry(nstreams, nstokes, *t, e)) return AA_EARG;
  g_table[key] = e;
  return 0;
}

int aa_get_tuning(int nstreams, int nstokes, int precision, aa_tuning* t) {
  using namespace aa;
  if (!t || nstreams < 1 || (nstokes != 1 && nstokes != 3) || (precision != 0 && precision != 1)) return AA_EARG;
  const TuneEntry e = tuning_for(nstreams, nstokes, precision);
  const int n = nstreams * nstokes;
  t->bvp_lanes = e.shape_bvp ? bvp_group_alt(n) : bvp_group(n);
  t->sab_pairs = 16 * (e.shape_sab ? sab_groups_alt(nstreams, nstokes) : sab_groups(nstreams, nstokes));
  t->chunk_items = (int)e.chunk_items;
  return 0;
}

int aa_load_tuning(const char* path) {
  using namespace aa;
  std::lock_guard<std::mutex> lk(g_mu);
  g_loaded = true;
  const int r = load_locked(path ? std::string(path) : default_path());
  return r == -1 ? AA_ENOFILE : (r == -2 ? AA_EARG : 0);
}

}  // extern "C"
