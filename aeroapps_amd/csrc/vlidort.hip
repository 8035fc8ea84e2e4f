// C-ABI of the discrete-ordinate radiance path: the drop-in replacements of
// VLIDORT_Scalar_Lambert / VLIDORT_Vector_Lambert
// (src/Components/rtms/geo_vlidort/VLIDORT_LAMBERT.F90:34-342) and the fused
// device path (Mie optics -> radiance on resident data).
//
// Host side: a grow-only device arena per GPU, the quadrature/PI tables of the
// quadrature streams (identical for every column of a call), and the chunked
// pipeline setup -> Fourier terms.
#include <algorithm>
#include <atomic>
#include <map>
#include <mutex>
#include <tuple>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "rt_work.h"
#include "rt_device.h"
#include "../../include/aeroapps_amd.h"

struct aa_mie;

namespace aa {

int rt_setup(const RTWork& W, int nitems, int* d_active, int* d_nactive, hipStream_t st);
int rt_latch(int nitems, int nch, int nobs, const double* sza, const double* vza, const double* raa,
             const double* albedo, const double* kernel_wt, const double* kparam, const double* u10m,
             const double* v10m, const double* mr, const double* bpdf, double missing, int* d_latch,
             hipStream_t st);
const std::vector<double>& brdf_table_cached(int N, int A);
const std::vector<double>& brdf_table_full_cached(int N);
#define AA_DECL(N, S) \
  int rt_fourier_##N##_##S(const RTWork& W, int nitems, int* const lists[2], int* counts, hipStream_t st);
AA_DECL(1, 1) AA_DECL(2, 1) AA_DECL(3, 1) AA_DECL(4, 1) AA_DECL(5, 1) AA_DECL(6, 1) AA_DECL(7, 1) AA_DECL(8, 1)
AA_DECL(9, 1) AA_DECL(10, 1) AA_DECL(11, 1) AA_DECL(12, 1) AA_DECL(13, 1) AA_DECL(14, 1) AA_DECL(15, 1)
AA_DECL(16, 1) AA_DECL(17, 1) AA_DECL(18, 1) AA_DECL(19, 1) AA_DECL(20, 1)
AA_DECL(1, 3) AA_DECL(2, 3) AA_DECL(3, 3) AA_DECL(4, 3) AA_DECL(5, 3) AA_DECL(6, 3)
#undef AA_DECL

// Fourier-term kernels for the call's (streams, Stokes) pair; -4 when the
// pair has no compiled instantiation.
int rt_fourier(const RTWork& W, int nitems, int* const lists[2], int* counts, hipStream_t st) {
  if (W.S == 1) {
    switch (W.N) {
      case 1: return rt_fourier_1_1(W, nitems, lists, counts, st);
      case 2: return rt_fourier_2_1(W, nitems, lists, counts, st);
      case 3: return rt_fourier_3_1(W, nitems, lists, counts, st);
      case 4: return rt_fourier_4_1(W, nitems, lists, counts, st);
      case 5: return rt_fourier_5_1(W, nitems, lists, counts, st);
      case 6: return rt_fourier_6_1(W, nitems, lists, counts, st);
      case 7: return rt_fourier_7_1(W, nitems, lists, counts, st);
      case 8: return rt_fourier_8_1(W, nitems, lists, counts, st);
      case 9: return rt_fourier_9_1(W, nitems, lists, counts, st);
      case 10: return rt_fourier_10_1(W, nitems, lists, counts, st);
      case 11: return rt_fourier_11_1(W, nitems, lists, counts, st);
      case 12: return rt_fourier_12_1(W, nitems, lists, counts, st);
      case 13: return rt_fourier_13_1(W, nitems, lists, counts, st);
      case 14: return rt_fourier_14_1(W, nitems, lists, counts, st);
      case 15: return rt_fourier_15_1(W, nitems, lists, counts, st);
      case 16: return rt_fourier_16_1(W, nitems, lists, counts, st);
      case 17: return rt_fourier_17_1(W, nitems, lists, counts, st);
      case 18: return rt_fourier_18_1(W, nitems, lists, counts, st);
      case 19: return rt_fourier_19_1(W, nitems, lists, counts, st);
      case 20: return rt_fourier_20_1(W, nitems, lists, counts, st);
      default: return -4;
    }
  }
  switch (W.N) {
    case 1: return rt_fourier_1_3(W, nitems, lists, counts, st);
    case 2: return rt_fourier_2_3(W, nitems, lists, counts, st);
    case 3: return rt_fourier_3_3(W, nitems, lists, counts, st);
    case 4: return rt_fourier_4_3(W, nitems, lists, counts, st);
    case 5: return rt_fourier_5_3(W, nitems, lists, counts, st);
    case 6: return rt_fourier_6_3(W, nitems, lists, counts, st);
    default: return -4;
  }
}
int mie_upload(aa_mie* t);
bool mie_resident(aa_mie* t);
int launch_get_aop(const aa_mie* t, int km, int nobs, int nch, int nq, const int* d_tracer_idx,
                   const int* d_chan_idx, const float* d_qm, const float* d_rh, int nmom, int npol, int ntr_eff,
                   int normalize, float* d_tau, float* d_ssa, float* d_g, float* d_pmom, hipStream_t st);

namespace {

// ---------------------------------------------------------------- device arenas
// One grow-only arena per (device, stream, purpose), so calls enqueued on
// different streams never share work memory and can overlap on the GPU.
struct Arena {
  char* base = nullptr;
  size_t cap = 0;
};
std::mutex g_arena_mu;
std::map<std::tuple<int, hipStream_t, int>, Arena> g_arenas;

int arena_reserve(hipStream_t st, int purpose, size_t bytes, char** out) {
  int dev = 0;
  AA_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_arena_mu);
  Arena& a = g_arenas[std::make_tuple(dev, st, purpose)];
  if (a.cap < bytes) {
    if (a.base) {
      AA_CHECK(hipStreamSynchronize(st));
      AA_CHECK(hipFree(a.base));
    }
    a.base = nullptr; a.cap = 0;
    size_t want = bytes + bytes / 8;
    AA_CHECK(hipMalloc((void**)&a.base, want));
    a.cap = want;
  }
  *out = a.base;
  return 0;
}
enum ArenaPurpose { AR_WORK = 0, AR_OPTICS = 1, AR_CALL = 2, AR_HOST = 3 };

// Budgets of the work and optics arenas of later calls (0: AA_WORK_BYTES /
// AA_OPTICS_BYTES, else 24 GiB / 4 GiB, which leaves room for several
// streams with an arena each).
std::atomic<size_t> g_work_bytes{0}, g_optics_bytes{0};
size_t env_bytes(const char* name, size_t dflt);
size_t work_budget() {
  const size_t v = g_work_bytes.load();
  return v ? v : env_bytes("AA_WORK_BYTES", (size_t)24 << 30);
}
size_t optics_budget() {
  const size_t v = g_optics_bytes.load();
  return v ? v : env_bytes("AA_OPTICS_BYTES", (size_t)4 << 30);
}

// Sequential carving of one arena block (a first pass with base == nullptr
// measures the block).
struct Carver {
  char* base;
  size_t off = 0;
  explicit Carver(char* b) : base(b) {}
  template <class T>
  T* take(size_t count) {
    off = (off + 255) & ~(size_t)255;
    T* p = base ? (T*)(base + off) : nullptr;
    off += count * sizeof(T);
    return p;
  }
};

size_t env_bytes(const char* name, size_t dflt) {
  const char* s = std::getenv(name);
  if (!s || !*s) return dflt;
  return (size_t)std::strtoull(s, nullptr, 10);
}

// Quadrature streams and their PI matrices for every Fourier term
// (GETQUAD2 REG/vlidort_aux.f90:960; PIMATRIX_SETUP REG/vlidort_miscsetups.f90:2273).
template <int S, int LM>
void quad_table_t(int N, int nmoments, std::vector<double>& q) {
  QOff QO(N, S);
  q.assign(QO.size(), 0.0);
  std::vector<double> qs(N), qw(N);
  if (N == 1) { qs[0] = std::sqrt(1.0 / 3.0); qw[0] = 1.0; }
  else getquad2(0.0, 1.0, N, qs.data(), qw.data());
  for (int i = 0; i < N; ++i) { q[QO.qs(i)] = qs[i]; q[QO.qw(i)] = qw[i]; }
  for (int M = 0; M < LM; ++M) {
    if (M > nmoments) break;
    for (int i = 0; i < N; ++i) {
      double pi[LM][S][S];
      pimatrix_x<S, LM>(M, nmoments, qs[i], pi);
      for (int l = 0; l < LM; ++l)
        for (int a = 0; a < S; ++a)
          for (int b = 0; b < S; ++b) q[QO.xq(M, l, i, a, b)] = pi[l][a][b];
    }
  }
}

void quad_table(int N, int S, int nmoments, std::vector<double>& q) {
  if (S == 1) {
    switch (N) {
      case 1: quad_table_t<1, 2>(N, nmoments, q); break;
      case 2: quad_table_t<1, 4>(N, nmoments, q); break;
      case 3: quad_table_t<1, 6>(N, nmoments, q); break;
      case 4: quad_table_t<1, 8>(N, nmoments, q); break;
      case 5: quad_table_t<1, 10>(N, nmoments, q); break;
      case 6: quad_table_t<1, 12>(N, nmoments, q); break;
      case 7: quad_table_t<1, 14>(N, nmoments, q); break;
      case 8: quad_table_t<1, 16>(N, nmoments, q); break;
      case 9: quad_table_t<1, 18>(N, nmoments, q); break;
      case 10: quad_table_t<1, 20>(N, nmoments, q); break;
      case 11: quad_table_t<1, 22>(N, nmoments, q); break;
      case 12: quad_table_t<1, 24>(N, nmoments, q); break;
      case 13: quad_table_t<1, 26>(N, nmoments, q); break;
      case 14: quad_table_t<1, 28>(N, nmoments, q); break;
      case 15: quad_table_t<1, 30>(N, nmoments, q); break;
      case 16: quad_table_t<1, 32>(N, nmoments, q); break;
      case 17: quad_table_t<1, 34>(N, nmoments, q); break;
      case 18: quad_table_t<1, 36>(N, nmoments, q); break;
      case 19: quad_table_t<1, 38>(N, nmoments, q); break;
      default: quad_table_t<1, 40>(N, nmoments, q); break;
    }
    static_assert(MAX_STREAMS == 20, "one quadrature-table case per scalar stream count");
  } else {
    switch (N) {
      case 1: quad_table_t<3, 2>(N, nmoments, q); break;
      case 2: quad_table_t<3, 4>(N, nmoments, q); break;
      case 3: quad_table_t<3, 6>(N, nmoments, q); break;
      case 4: quad_table_t<3, 8>(N, nmoments, q); break;
      case 5: quad_table_t<3, 10>(N, nmoments, q); break;
      default: quad_table_t<3, 12>(N, nmoments, q); break;
    }
    static_assert(MAX_STREAMS_VECTOR == 6, "one quadrature-table case per vector stream count");
  }
}

int max_streams(int S) { return S == 1 ? MAX_STREAMS : MAX_STREAMS_VECTOR; }

// Quadrature tables are built once per (N, S, nmoments) and kept for the
// process lifetime (the host copy source of every asynchronous upload).
const std::vector<double>& quad_table_cached(int N, int S, int nmoments) {
  static std::mutex mu;
  static std::map<std::tuple<int, int, int>, std::vector<double>> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto key = std::make_tuple(N, S, nmoments);
  auto it = cache.find(key);
  if (it == cache.end()) {
    std::vector<double> q;
    quad_table(N, S, nmoments, q);
    it = cache.emplace(key, std::move(q)).first;
  }
  return it->second;
}

// Contraction weights of the mixed-precision SAB/DAB GEMM for every Fourier
// term: SAB(r,c) = sum_k Y[r*n+c][k] og[k], DAB(r,c) = sum_k Y[n^2+r*n+c][k] og[k]
// with k = (l,o,o3) (restates the SAB/DAB sums of REG/vlidort_solutions.f90:
// 439-475 with the quadrature PI matrices folded in).
const std::vector<double>& ytab64_cached(int N, int S, int nmoments) {
  static std::mutex mu;
  static std::map<std::tuple<int, int, int>, std::vector<double>> cache;
  const std::vector<double>& q = quad_table_cached(N, S, nmoments);
  std::lock_guard<std::mutex> lk(mu);
  auto key = std::make_tuple(N, S, nmoments);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  const int n = N * S, LM = 2 * N;
  QOff QO(N, S);
  YOff YO(N, S);
  std::vector<double> y((size_t)YO.size(LM), 0.0);
  const double DM[4] = {1.0, 1.0, -1.0, -1.0};
  for (int M = 0; M < LM && M <= nmoments; ++M)
    for (int i = 0; i < N; ++i)
      for (int o1 = 0; o1 < S; ++o1)
        for (int j = 0; j < N; ++j)
          for (int o2 = 0; o2 < S; ++o2) {
            const int r = i * S + o1, c = j * S + o2;
            const double fac = (-1.0 / q[QO.qs(i)]) * (0.5 * q[QO.qw(j)]);
            for (int l = M; l < LM && l <= nmoments; ++l) {
              const double ps = (((l - M) & 1) == 0) ? 1.0 : -1.0;
              for (int o = 0; o < S; ++o)
                for (int o3 = 0; o3 < S; ++o3) {
                  const int k = (l * S + o) * S + o3;
                  const double xx = q[QO.xq(M, l, i, o1, o)] * q[QO.xq(M, l, j, o3, o2)];
                  y[YO.at(M, r * n + c, k)] = fac * xx * (1.0 + DM[o3] * ps);
                  y[YO.at(M, n * n + r * n + c, k)] = fac * xx * (1.0 - DM[o3] * ps);
                }
            }
          }
  return cache.emplace(key, std::move(y)).first->second;
}

// The same weights rounded to FP32 for the mixed path.
const std::vector<float>& ytab_cached(int N, int S, int nmoments) {
  static std::mutex mu;
  static std::map<std::tuple<int, int, int>, std::vector<float>> cache;
  const std::vector<double>& y = ytab64_cached(N, S, nmoments);
  std::lock_guard<std::mutex> lk(mu);
  auto key = std::make_tuple(N, S, nmoments);
  auto it = cache.find(key);
  if (it == cache.end()) it = cache.emplace(key, std::vector<float>(y.begin(), y.end())).first;
  return it->second;
}

// Library streams (default priority) used to run stream-count buckets of one
// call concurrently; created once per (device, caller stream, bucket slot), so
// calls on different caller streams never share them.
hipStream_t bucket_stream(hipStream_t caller, int k) {
  static std::mutex mu;
  static std::map<std::tuple<int, hipStream_t, int>, hipStream_t> pool;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(mu);
  auto key = std::make_tuple(dev, caller, k);
  auto it = pool.find(key);
  if (it != pool.end()) return it->second;
  hipStream_t st = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return nullptr;
  pool[key] = st;
  return st;
}

// Per-item work of one chunk, carved from the arena.
size_t carve_work(Carver& C, RTWork& W, int cap) {
  const int n = W.N * W.S, LM = 2 * W.N, km = W.km;
  W.lstride = LOff(W.N, W.S).size();
  W.bstride = BOff(W.N, n, W.S).size();
  W.brdf_stride = W.bpdf ? BPOff(W.N).size() : ((W.kernel_wt || W.u10m) ? BROff(W.N).size() : 0);
  W.cap = cap;
  W.status = C.take<int>(cap);
  W.L_it = C.take<int>(cap);
  W.beam_cutoff = C.take<int>(cap);
  W.testconv = C.take<int>(cap);
  W.fsaved = C.take<int>(cap);
  W.nfour = C.take<int>(cap);
  W.loff = C.take<int>(cap);
  W.lown = C.take<int>((size_t)cap * km);
  W.csza = C.take<double>(cap);
  W.usec = C.take<double>(cap);
  W.tsb = C.take<double>(cap);
  W.ss = C.take<double>((size_t)cap * 3);
  W.db = C.take<double>((size_t)cap * 3);
  W.stokes = C.take<double>((size_t)cap * 3);
  W.deltau = C.take<double>((size_t)cap * km);
  W.fac1 = C.take<double>((size_t)cap * km);
  W.ext = C.take<double>((size_t)cap * km);
  W.exact = C.take<double>((size_t)cap * km * 3);
  W.og = C.take<double>((size_t)cap * km * LM * W.S * W.S);
  W.lsc = C.take<unsigned>((size_t)cap * km);
  W.tr = C.take<double>((size_t)cap * km * TR_NFIELDS);
  W.lev = C.take<double>((size_t)cap * (km + 1) * 4);
  W.lostrans = C.take<double>((size_t)cap * (km + 1));
  W.lbuf = C.take<double>((size_t)cap * km * W.lstride);
  W.bbuf = C.take<double>((size_t)cap * W.bstride);
  W.brdf = (W.kernel_wt || W.u10m || W.bpdf) ? C.take<double>((size_t)cap * W.brdf_stride) : nullptr;
  W.wsa_cache = W.kernel_wt ? C.take<double>(4) : nullptr;
  W.sab32 = (W.precision == 1) ? C.take<float>((size_t)cap * km * 2 * n * n) : nullptr;
  W.sab64 = (W.precision == 0) ? C.take<double>((size_t)cap * km * 2 * n * n) : nullptr;
  return C.off;
}

size_t bytes_per_item(const RTWork& W0) {
  RTWork W = W0;
  Carver a(nullptr), b(nullptr);
  size_t s1 = carve_work(a, W, 1), s2 = carve_work(b, W, 2);
  return s2 - s1 + 256 * 24;
}

// Items of one work chunk of W's shape on the current GPU: the work budget
// (capped at 60% of the free device memory) over the bytes of one item.
int work_cap_items(const RTWork& W, size_t* cap) {
  size_t freeb = 0, totalb = 0;
  AA_CHECK(hipMemGetInfo(&freeb, &totalb));
  const size_t budget = std::min(work_budget(), (size_t)(0.6 * (double)freeb));
  *cap = std::max<size_t>(1, budget / bytes_per_item(W));
  const size_t envcap = env_bytes("AA_CHUNK_ITEMS", 0);
  if (envcap) *cap = std::min(*cap, envcap);
  const TuneEntry tu = tuning_for(W.N, W.S, W.precision);
  if (tu.chunk_items > 0) *cap = std::min(*cap, (size_t)tu.chunk_items);
  return 0;
}

// First item of the SZA==0 nadir latch for the current call, in the call
// arena of the stream.
int compute_latch(RTWork& W, int* lp, hipStream_t st) {
  static const int big = INT_MAX;
  AA_CHECK(hipMemcpyAsync(lp, &big, sizeof(int), hipMemcpyHostToDevice, st));
  const int nitems = W.nobs * W.nch;
  if (nitems > 0 && rt_latch(nitems, W.nch, W.nobs, W.sza, W.vza, W.raa, W.albedo, W.kernel_wt, W.kparam, W.u10m,
                             W.v10m, W.mr, W.bpdf, W.missing, lp, st))
    return AA_EHIP;
  W.latch = lp;
  return 0;
}

// Runs the radiance pipeline over items [g0, g1) (item g = obs*nch + ch), or
// over d_items[g0..g1) when d_items (device) is given.  W carries the inputs
// and outputs (device pointers) and the latch.  Work memory comes from the
// stream's work arena, or, when fixed_base is given, from that caller-owned
// device block of fixed_bytes (chunks are then sized to fit it).
int run_items(RTWork W, int g0, int g1, hipStream_t st, const int* d_items = nullptr, char* fixed_base = nullptr,
              size_t fixed_bytes = 0) {
  const int nitems = g1 - g0;
  if (nitems <= 0) return 0;
  {
    const TuneEntry tu = tuning_for(W.N, W.S, W.precision);
    W.shape_bvp = tu.shape_bvp;
    W.shape_sab = tu.shape_sab;
  }
  const std::vector<double>& qtab = quad_table_cached(W.N, W.S, W.nmoments);
  static const std::vector<float> no_ytab;
  const std::vector<float>& ytab = (W.precision == 1) ? ytab_cached(W.N, W.S, W.nmoments) : no_ytab;
  static const std::vector<double> no_ytab64;
  const std::vector<double>& ytab64 = (W.precision == 0) ? ytab64_cached(W.N, W.S, W.nmoments) : no_ytab64;
  static const std::vector<double> no_btab;
  const std::vector<double>& btab =
      W.bpdf ? brdf_table_full_cached(W.N)
             : (W.kernel_wt ? brdf_table_cached(W.N, BR_NAZ) : (W.u10m ? brdf_table_cached(W.N, BR_NAZ_CM) : no_btab));
  double* dbt = nullptr;
  // chunk capacity from the work budget, or from the fixed block: its bytes
  // less the tables, over the bytes of an item (work record + two list
  // entries), checked by carving below
  size_t capi = 1;
  const size_t tab_bytes = qtab.size() * sizeof(double) + std::max<size_t>(1, ytab.size()) * sizeof(float) +
                           std::max<size_t>(1, ytab64.size()) * sizeof(double) +
                           std::max<size_t>(1, btab.size()) * sizeof(double) + sizeof(int) * (W.nmoments + 2) +
                           16 * 256;
  if (fixed_base) {
    const size_t per = bytes_per_item(W) + 2 * LBINS * sizeof(int);
    if (fixed_bytes <= tab_bytes + per) return AA_EDIMS;
    capi = (fixed_bytes - tab_bytes) / per;
  } else if (work_cap_items(W, &capi)) {
    return AA_EHIP;
  }
  // chunks of equal size: a small remainder chunk would run every Fourier
  // term on a nearly idle GPU
  const size_t nchunk = ((size_t)nitems + capi - 1) / capi;
  const int cap = (int)(((size_t)nitems + nchunk - 1) / nchunk);
  // arena: work + quad table + active lists + counters
  float* dy = nullptr;
  double* dy64 = nullptr;
  auto carve_all = [&](char* base, RTWork& w, double** dq, int** l0, int** l1, int** cnt) {
    Carver C(base);
    carve_work(C, w, cap);
    *dq = C.take<double>(qtab.size());
    dy = C.take<float>(std::max<size_t>(1, ytab.size()));
    dy64 = C.take<double>(std::max<size_t>(1, ytab64.size()));
    dbt = C.take<double>(std::max<size_t>(1, btab.size()));
    *l0 = C.take<int>((size_t)LBINS * cap);
    *l1 = C.take<int>((size_t)LBINS * cap);
    // per Fourier term: items of each layer-count list, then (item, layer) slots
    *cnt = C.take<int>((LBINS + 1) * (W.nmoments + 2));
    return C.off;
  };
  double* dq; int *l0, *l1, *cnt;
  RTWork tmp = W;
  size_t total = carve_all(nullptr, tmp, &dq, &l0, &l1, &cnt);
  char* base = fixed_base;
  if (fixed_base) {
    if (total > fixed_bytes) return AA_EDIMS;
  } else if (arena_reserve(st, AR_WORK, total, &base)) {
    return AA_EHIP;
  }
  carve_all(base, W, &dq, &l0, &l1, &cnt);
  AA_CHECK(hipMemcpyAsync(dq, qtab.data(), qtab.size() * sizeof(double), hipMemcpyHostToDevice, st));
  W.quad = dq;
  if (!ytab.empty()) {
    AA_CHECK(hipMemcpyAsync(dy, ytab.data(), ytab.size() * sizeof(float), hipMemcpyHostToDevice, st));
    W.ytab = dy;
  }
  if (!ytab64.empty()) {
    AA_CHECK(hipMemcpyAsync(dy64, ytab64.data(), ytab64.size() * sizeof(double), hipMemcpyHostToDevice, st));
    W.ytab64 = dy64;
  }
  if (!btab.empty()) {
    AA_CHECK(hipMemcpyAsync(dbt, btab.data(), btab.size() * sizeof(double), hipMemcpyHostToDevice, st));
    W.brdf_tab = dbt;
  }
  int* lists[2] = {l0, l1};
  W.item_list = d_items;
  W.lcnt = cnt + LBINS * (W.nmoments + 2);
  for (int item0 = g0; item0 < g1; item0 += cap) {
    const int ni = std::min(cap, g1 - item0);
    W.item0 = item0;
    AA_CHECK(hipMemsetAsync(cnt, 0, sizeof(int) * (LBINS + 1) * (W.nmoments + 2), st));
    if (rt_setup(W, ni, lists[0], cnt, st)) return AA_EHIP;
    int e = rt_fourier(W, ni, lists, cnt, st);
    if (e) return e == -4 ? AA_EDIMS : AA_EHIP;
  }
  return 0;
}

}  // namespace

// Device memory of the host-pointer entry points: one grow-only block per
// device, which the synchronous host-pointer calls use one at a time
// (host_call_mutex) instead of allocating and freeing per call.
std::mutex& host_call_mutex() {
  static std::mutex m;
  return m;
}
int host_call_block(size_t bytes, char** out) { return arena_reserve(nullptr, AR_HOST, bytes, out); }

namespace {

int lambert_host(bool vector, int km, int nch, int nobs, int nstreams, int plane_parallel, int nmom, int npol,
                 const double* rot, const double* depol, const double* alpha, const double* tau, const double* ssa,
                 const double* pmom, const double* he, const double* albedo, const double* sza, const double* raa,
                 const double* vza, const double* flux_factor, double missing, int precision, double* radiance,
                 double* reflectance, double* q, double* u, int* nfourier, int* rc,
                 const double* kernel_wt = nullptr, const double* kparam = nullptr, double* br = nullptr,
                 const double* u10m = nullptr, const double* v10m = nullptr, const double* mr = nullptr,
                 bool stokes1 = false, const int* host_latch = nullptr, const double* bpdf = nullptr,
                 double* br_q = nullptr, double* br_u = nullptr) {
  if (!rc) return AA_EARG;
  *rc = 0;
  const bool cx = u10m != nullptr;
  const bool pol = bpdf != nullptr;                // polarised land surface (BPDF kernel last)
  if (cx && (!v10m || !mr || !br || kernel_wt || stokes1 || pol)) return AA_EARG;
  if (pol && (!vector || stokes1 || !br || !br_q || !br_u)) return AA_EARG;
  const bool modis = kernel_wt != nullptr || cx;   // a BRDF surface instead of the albedo
  const bool surf_brdf = modis || pol;             // the radiance runs on a BRDF record
  if (km <= 0 || nch <= 0 || nobs < 0 || nmom <= 0 || npol <= 0) return AA_EARG;
  if (!rot || !depol || !tau || !ssa || !pmom || !he || (!albedo && !modis) || !sza || !raa || !vza || !radiance ||
      !reflectance)
    return AA_EARG;
  if (kernel_wt && ((vector && !pol) || !kparam || !br)) return AA_EARG;
  // VLIDORT_Init of the BRDF drivers (V283/VLIDORT_Mod.F90:369) rejects nMom
  // beyond the input maximum for the whole call
  if (surf_brdf && nmom > MAX_MOMENTS_IN) { *rc = 5; return 0; }
  // (the vector Cox-Munk and BPDF drivers set no flux factor: VLIDORT_Init's 1.0)
  if (vector && (!q || !u || (!flux_factor && !cx && !pol) || npol < 6)) return AA_EARG;
  if (precision != 0 && precision != 1) return AA_EARG;
  // stokes1: the vector driver run with NSTOKES = 1 (its DO_2OS_CORRECTION
  // option): Stokes-1 arithmetic on the vector driver's inputs (gas
  // absorption, flux factor); Q and U are 0
  const int S = (vector && !stokes1) ? 3 : 1;
  if (nstreams < 1 || nstreams > max_streams(S)) { *rc = 2; return AA_EDIMS; }
  if (km > MAX_LAYERS) { *rc = 3; return AA_EDIMS; }
  if (nobs == 0) return 0;
  RTWork W;
  W.km = km; W.nch = nch; W.nobs = nobs; W.nmom = nmom; W.npol = npol;
  W.N = nstreams; W.S = S; W.nmoments = std::min(2 * nstreams - 1, nmom);
  W.plane_parallel = plane_parallel ? 1 : 0;
  W.missing = missing;
  W.precision = precision;
  const size_t nl = (size_t)km * nobs * nch;
  const size_t no = (size_t)nobs * nch;
  std::lock_guard<std::mutex> hl(host_call_mutex());
  // device copies of the inputs and the outputs, carved from the host-call block
  struct Dev {
    double *rot, *depol, *alpha, *tau, *ssa, *pmom, *he, *albedo, *kw, *kp, *u10, *v10, *mr, *br, *sza, *raa, *vza,
        *flux, *rad, *refl, *q, *u, *bp, *brq, *bru;
    int* nf;
  } d{};
  auto plan = [&](Carver& C) {
    d.rot = C.take<double>(nl);
    d.depol = C.take<double>(nch);
    d.alpha = (vector && alpha) ? C.take<double>(nl) : nullptr;
    d.tau = C.take<double>(nl);
    d.ssa = C.take<double>(nl);
    d.pmom = C.take<double>(nl * nmom * npol);
    d.he = C.take<double>((size_t)(km + 1) * nobs);
    d.albedo = modis ? nullptr : C.take<double>(no);
    d.kw = kernel_wt ? C.take<double>(no * 3) : nullptr;
    d.kp = kernel_wt ? C.take<double>(no * 2) : nullptr;
    d.u10 = cx ? C.take<double>(nobs) : nullptr;
    d.v10 = cx ? C.take<double>(nobs) : nullptr;
    d.mr = cx ? C.take<double>(nch) : nullptr;
    d.br = surf_brdf ? C.take<double>(no) : nullptr;
    d.bp = pol ? C.take<double>(no * 3) : nullptr;
    d.brq = pol ? C.take<double>(no) : nullptr;
    d.bru = pol ? C.take<double>(no) : nullptr;
    d.sza = C.take<double>(nobs);
    d.raa = C.take<double>(nobs);
    d.vza = C.take<double>(nobs);
    d.flux = (vector && flux_factor) ? C.take<double>(no) : nullptr;
    d.rad = C.take<double>(no);
    d.refl = C.take<double>(no);
    d.q = S == 3 ? C.take<double>(no) : nullptr;
    d.u = S == 3 ? C.take<double>(no) : nullptr;
    d.nf = C.take<int>(no);
  };
  {
    Carver c0(nullptr);
    plan(c0);
    char* base = nullptr;
    if (host_call_block(c0.off, &base)) return AA_EHIP;
    Carver c1(base);
    plan(c1);
  }
  auto up = [](double* dst, const double* src, size_t n) {
    return (n && src) ? hipMemcpy(dst, src, n * sizeof(double), hipMemcpyHostToDevice) : hipSuccess;
  };
  if (up(d.rot, rot, nl) || up(d.depol, depol, nch) || (d.alpha && up(d.alpha, alpha, nl)) || up(d.tau, tau, nl) ||
      up(d.ssa, ssa, nl) || up(d.pmom, pmom, nl * nmom * npol) || up(d.he, he, (size_t)(km + 1) * nobs) ||
      (d.albedo && up(d.albedo, albedo, no)) || (d.kw && up(d.kw, kernel_wt, no * 3)) ||
      (d.kp && up(d.kp, kparam, no * 2)) || (d.u10 && up(d.u10, u10m, nobs)) || (d.v10 && up(d.v10, v10m, nobs)) ||
      (d.mr && up(d.mr, mr, nch)) || up(d.sza, sza, nobs) || up(d.raa, raa, nobs) || up(d.vza, vza, nobs) ||
      (d.flux && up(d.flux, flux_factor, no)) || (d.bp && up(d.bp, bpdf, no * 3)))
    return AA_EHIP;
  W.rot = d.rot; W.depol = d.depol; W.alpha = d.alpha;
  W.tau = d.tau; W.ssa = d.ssa; W.pmom = d.pmom;
  W.optics_f32 = 0; W.opt_obs0 = 0; W.opt_nobs = nobs;
  W.he = d.he; W.albedo = d.albedo;
  W.kernel_wt = d.kw; W.kparam = d.kp;
  W.u10m = d.u10; W.v10m = d.v10; W.mr = d.mr;
  W.br_out = d.br;
  W.bpdf = d.bp; W.brq_out = d.brq; W.bru_out = d.bru;
  W.sza = d.sza; W.raa = d.raa; W.vza = d.vza;
  W.flux = d.flux;
  W.radiance = d.rad; W.reflectance = d.refl; W.qout = d.q; W.uout = d.u; W.nfourier = d.nf;
  int* lp = nullptr;
  {
    char* cb = nullptr;
    if (arena_reserve(0, AR_CALL, 256, &cb)) return AA_EHIP;
    lp = (int*)cb;
  }
  int e = 0;
  if (host_latch) {
    AA_CHECK(hipMemcpy(lp, host_latch, sizeof(int), hipMemcpyHostToDevice));
    W.latch = lp;
  } else {
    e = compute_latch(W, lp, 0);
  }
  if (!e) e = run_items(W, 0, nobs * nch, 0);
  if (e) return e;
  AA_CHECK(hipDeviceSynchronize());
  AA_CHECK(hipMemcpy(radiance, W.radiance, no * sizeof(double), hipMemcpyDeviceToHost));
  AA_CHECK(hipMemcpy(reflectance, W.reflectance, no * sizeof(double), hipMemcpyDeviceToHost));
  if (S == 3) {
    AA_CHECK(hipMemcpy(q, W.qout, no * sizeof(double), hipMemcpyDeviceToHost));
    AA_CHECK(hipMemcpy(u, W.uout, no * sizeof(double), hipMemcpyDeviceToHost));
  } else if (vector) {
    // NSTOKES = 1: the reference returns STOKES(..,2..3,..) = 0, MISSING where
    // the element is MISSING
    for (size_t i = 0; i < no; ++i) {
      const bool miss = std::fabs(radiance[i] / missing - 1.0) < 0.001;
      q[i] = u[i] = miss ? missing : 0.0;
    }
  }
  if (nfourier) AA_CHECK(hipMemcpy(nfourier, W.nfourier, no * sizeof(int), hipMemcpyDeviceToHost));
  if (surf_brdf) AA_CHECK(hipMemcpy(br, W.br_out, no * sizeof(double), hipMemcpyDeviceToHost));
  if (pol) {
    AA_CHECK(hipMemcpy(br_q, W.brq_out, no * sizeof(double), hipMemcpyDeviceToHost));
    AA_CHECK(hipMemcpy(br_u, W.bru_out, no * sizeof(double), hipMemcpyDeviceToHost));
  }
  return 0;
}

// VLIDORT_Vector_Lambert_MultiGeom (GEO/VLIDORT_LAMBERT.F90:346-520 over
// VLIDORT_Run_Vector_MultiGeom, V283/VLIDORT_ScatMod.F90:621-972): ngeom
// (sza, vza, raa) triplets per column in one VLIDORT_MASTER call.  Every
// geometry has its own beam solution, convergence test and Fourier count, so a
// geometry's result equals that of a single-geometry run of the same column,
// except for what the driver shares across geometries:
//   * the missing-geometry and missing-albedo checks look at geometry 1 only
//     (IS_MISSING(solar_zenith(j,1)), albedo(j,i,1)) and the surface albedo is
//     albedo(j,i,1) for every geometry (VLIDORT_SurfaceLamb_MultiGeom);
//   * the overhead-sun switch to nadir first-order mode fires when any
//     geometry of the column has SZA == 0 and stays on for the rest of the call;
//   * a failed MASTER call makes every geometry of that (column, channel)
//     MISSING.
// The columns are expanded to nobs*ngeom single-geometry observations
// (observation o = j*ngeom + g) and run through the single-geometry path
// with the latch item computed here.
int lambert_multigeom_host(int km, int nch, int nobs, int ngeom, int nstreams, int plane_parallel, int nmom,
                           int npol, const double* rot, const double* depol, const double* alpha, const double* tau,
                           const double* ssa, const double* pmom, const double* he, const double* albedo,
                           const double* sza, const double* raa, const double* vza, const double* flux_factor,
                           double missing, int precision, bool stokes1, double* radiance, double* reflectance,
                           double* q, double* u, int* nfourier, int* rc) {
  if (!rc) return AA_EARG;
  *rc = 0;
  if (km <= 0 || nch <= 0 || nobs < 0 || ngeom <= 0 || nmom <= 0 || npol < 6) return AA_EARG;
  if (!rot || !depol || !tau || !ssa || !pmom || !he || !albedo || !sza || !raa || !vza || !flux_factor ||
      !radiance || !reflectance || !q || !u)
    return AA_EARG;
  // VLIDORT_Init rejects more beams than MAX_SZANGLES (V283/VLIDORT_Mod.F90:370)
  if (ngeom > MAX_OBSGEOMS) { *rc = 6; return 0; }
  if (nobs == 0) return 0;
  const int no = nobs * ngeom;
  const size_t L = (size_t)km;
  auto miss = [&](double x) { return std::fabs(x / missing - 1.0) < 0.001; };
  std::vector<double> e_rot((size_t)km * no * nch), e_alpha(alpha ? e_rot.size() : 0);
  std::vector<double> e_tau(L * nch * no), e_ssa(e_tau.size()), e_pmom(e_tau.size() * nmom * npol);
  std::vector<double> e_he((L + 1) * no), e_alb((size_t)no * nch), e_sza(no), e_raa(no), e_vza(no);
  std::vector<double> e_ff((size_t)nch * no);
  int latch = INT_MAX;
  for (int j = 0; j < nobs; ++j) {
    const bool gmiss = miss(sza[j]) || miss(vza[j]) || miss(raa[j]);
    bool any_zero = false;
    for (int g = 0; g < ngeom; ++g) any_zero = any_zero || sza[j + (size_t)g * nobs] == 0.0;
    for (int g = 0; g < ngeom; ++g) {
      const int o = j * ngeom + g;
      const size_t gj = j + (size_t)g * nobs;
      e_sza[o] = gmiss ? missing : sza[gj];
      e_vza[o] = gmiss ? missing : vza[gj];
      e_raa[o] = gmiss ? missing : raa[gj];
      std::memcpy(&e_he[(size_t)o * (L + 1)], &he[(size_t)j * (L + 1)], (L + 1) * sizeof(double));
      for (int c = 0; c < nch; ++c) {
        e_alb[o + (size_t)c * no] = albedo[j + (size_t)c * nobs];            // albedo(j,c,1)
        e_ff[c + (size_t)o * nch] = flux_factor[c + (size_t)j * nch];
        std::memcpy(&e_rot[((size_t)c * no + o) * L], &rot[((size_t)c * nobs + j) * L], L * sizeof(double));
        if (alpha)
          std::memcpy(&e_alpha[((size_t)c * no + o) * L], &alpha[((size_t)c * nobs + j) * L], L * sizeof(double));
        std::memcpy(&e_tau[((size_t)o * nch + c) * L], &tau[((size_t)j * nch + c) * L], L * sizeof(double));
        std::memcpy(&e_ssa[((size_t)o * nch + c) * L], &ssa[((size_t)j * nch + c) * L], L * sizeof(double));
        for (int pm = 0; pm < nmom * npol; ++pm)
          std::memcpy(&e_pmom[((size_t)pm * no * nch + (size_t)o * nch + c) * L],
                      &pmom[((size_t)pm * nobs * nch + (size_t)j * nch + c) * L], L * sizeof(double));
      }
    }
    // the latch fires on the first (column, channel) that reaches the run
    if (latch == INT_MAX && any_zero && !gmiss)
      for (int c = 0; c < nch; ++c)
        if (!miss(albedo[j + (size_t)c * nobs])) { latch = (j * ngeom) * nch + c; break; }
  }
  const size_t nout = (size_t)no * nch;
  std::vector<double> r(nout), f(nout), qq(nout), uu(nout);
  std::vector<int> nf(nout);
  int e = lambert_host(true, km, nch, no, nstreams, plane_parallel, nmom, npol, e_rot.data(), depol,
                       alpha ? e_alpha.data() : nullptr, e_tau.data(), e_ssa.data(), e_pmom.data(), e_he.data(),
                       e_alb.data(), e_sza.data(), e_raa.data(), e_vza.data(), e_ff.data(), missing, precision,
                       r.data(), f.data(), qq.data(), uu.data(), nf.data(), rc, nullptr, nullptr, nullptr, nullptr,
                       nullptr, nullptr, stokes1, &latch);
  if (e) return e;
  // outputs (nobs, nch, ngeom); one failed geometry fails the (column, channel)
  for (int j = 0; j < nobs; ++j)
    for (int c = 0; c < nch; ++c) {
      bool failed = false;
      for (int g = 0; g < ngeom; ++g) failed = failed || miss(r[(size_t)c * no + j * ngeom + g]);
      for (int g = 0; g < ngeom; ++g) {
        const size_t src = (size_t)c * no + j * ngeom + g;
        const size_t dst = j + (size_t)c * nobs + (size_t)g * nobs * nch;
        radiance[dst] = failed ? missing : r[src];
        reflectance[dst] = failed ? missing : f[src];
        q[dst] = failed ? missing : qq[src];
        u[dst] = failed ? missing : uu[src];
        if (nfourier) nfourier[dst] = failed ? 0 : nf[src];
      }
    }
  return 0;
}

}  // namespace
}  // namespace aa

extern "C" {

int aa_vlidort_scalar_lambert(int km, int nch, int nobs, const double* channels, int nstreams, int plane_parallel,
                              int nmom, int npol, const double* rot, const double* depol, const double* tau,
                              const double* ssa, const double* g, const double* pmom, const double* pe,
                              const double* he, const double* te, const double* albedo,
                              const double* solar_zenith, const double* relat_azymuth, const double* sensor_zenith,
                              double missing, int verbose, int precision, double* radiance, double* reflectance,
                              int* nfourier, int* rc) {
  (void)channels; (void)g; (void)pe; (void)te; (void)verbose;
  return aa::lambert_host(false, km, nch, nobs, nstreams, plane_parallel, nmom, npol, rot, depol, nullptr, tau, ssa,
                          pmom, he, albedo, solar_zenith, relat_azymuth, sensor_zenith, nullptr, missing, precision,
                          radiance, reflectance, nullptr, nullptr, nfourier, rc);
}

int aa_vlidort_vector_lambert(int km, int nch, int nobs, const double* channels, int nstreams, int plane_parallel,
                              int nmom, int npol, const double* rot, const double* depol, const double* alpha,
                              const double* tau, const double* ssa, const double* pmom, const double* pe,
                              const double* he, const double* te, const double* albedo,
                              const double* solar_zenith, const double* relat_azymuth, const double* sensor_zenith,
                              const double* flux_factor, double missing, int verbose, int precision,
                              double* radiance, double* reflectance, double* q, double* u, int* nfourier, int* rc) {
  return aa_vlidort_vector_lambert_opt(km, nch, nobs, channels, nstreams, plane_parallel, nmom, npol, rot, depol,
                                       alpha, tau, ssa, pmom, pe, he, te, albedo, solar_zenith, relat_azymuth,
                                       sensor_zenith, flux_factor, missing, verbose, precision, 0, 0, radiance,
                                       reflectance, q, u, nfourier, rc);
}

int aa_vlidort_vector_lambert_opt(int km, int nch, int nobs, const double* channels, int nstreams,
                                  int plane_parallel, int nmom, int npol, const double* rot, const double* depol,
                                  const double* alpha, const double* tau, const double* ssa, const double* pmom,
                                  const double* pe, const double* he, const double* te, const double* albedo,
                                  const double* solar_zenith, const double* relat_azymuth,
                                  const double* sensor_zenith, const double* flux_factor, double missing,
                                  int verbose, int precision, int do_2os_correction, int do_boa, double* radiance,
                                  double* reflectance, double* q, double* u, int* nfourier, int* rc) {
  (void)channels; (void)pe; (void)te; (void)verbose;
  if (do_boa) return AA_EUNSUP;
  return aa::lambert_host(true, km, nch, nobs, nstreams, plane_parallel, nmom, npol, rot, depol, alpha, tau, ssa,
                          pmom, he, albedo, solar_zenith, relat_azymuth, sensor_zenith, flux_factor, missing,
                          precision, radiance, reflectance, q, u, nfourier, rc, nullptr, nullptr, nullptr, nullptr,
                          nullptr, nullptr, do_2os_correction != 0);
}

int aa_vlidort_vector_lambert_multigeom(int km, int nch, int nobs, int ngeom, const double* channels, int nstreams,
                                        int plane_parallel, int nmom, int npol, const double* rot,
                                        const double* depol, const double* alpha, const double* tau,
                                        const double* ssa, const double* pmom, const double* pe, const double* he,
                                        const double* te, const double* albedo, const double* solar_zenith,
                                        const double* relat_azymuth, const double* sensor_zenith,
                                        const double* flux_factor, double missing, int verbose, int precision,
                                        int do_2os_correction, int do_boa, double* radiance, double* reflectance,
                                        double* q, double* u, int* nfourier, int* rc) {
  (void)channels; (void)pe; (void)te; (void)verbose;
  if (do_boa) return AA_EUNSUP;
  return aa::lambert_multigeom_host(km, nch, nobs, ngeom, nstreams, plane_parallel, nmom, npol, rot, depol, alpha,
                                    tau, ssa, pmom, he, albedo, solar_zenith, relat_azymuth, sensor_zenith,
                                    flux_factor, missing, precision, do_2os_correction != 0, radiance, reflectance,
                                    q, u, nfourier, rc);
}

int aa_vlidort_scalar_cx(int km, int nch, int nobs, const double* channels, int nstreams, int plane_parallel,
                         int nmom, int npol, const double* rot, const double* depol, const double* tau,
                         const double* ssa, const double* g, const double* pmom, const double* pe, const double* he,
                         const double* te, const double* u10m, const double* v10m, const double* mr,
                         const double* solar_zenith, const double* relat_azymuth, const double* sensor_zenith,
                         double missing, int verbose, int precision, double* radiance, double* reflectance,
                         double* brdf, int* nfourier, int* rc) {
  (void)channels; (void)g; (void)pe; (void)te; (void)verbose;
  if (!u10m) return AA_EARG;
  return aa::lambert_host(false, km, nch, nobs, nstreams, plane_parallel, nmom, npol, rot, depol, nullptr, tau, ssa,
                          pmom, he, nullptr, solar_zenith, relat_azymuth, sensor_zenith, nullptr, missing, precision,
                          radiance, reflectance, nullptr, nullptr, nfourier, rc, nullptr, nullptr, brdf, u10m, v10m,
                          mr);
}

int aa_vlidort_vector_cx(int km, int nch, int nobs, const double* channels, int nstreams, int plane_parallel,
                         int nmom, int npol, const double* rot, const double* depol, const double* tau,
                         const double* ssa, const double* pmom, const double* pe, const double* he, const double* te,
                         const double* u10m, const double* v10m, const double* mr, const double* solar_zenith,
                         const double* relat_azymuth, const double* sensor_zenith, double missing, int verbose,
                         int precision, double* radiance, double* reflectance, double* q, double* u, double* brdf,
                         double* brdf_q, double* brdf_u, int* nfourier, int* rc) {
  (void)channels; (void)pe; (void)te; (void)verbose;
  if (!u10m || !brdf_q || !brdf_u) return AA_EARG;
  const int e = aa::lambert_host(true, km, nch, nobs, nstreams, plane_parallel, nmom, npol, rot, depol, nullptr,
                                 tau, ssa, pmom, he, nullptr, solar_zenith, relat_azymuth, sensor_zenith, nullptr,
                                 missing, precision, radiance, reflectance, q, u, nfourier, rc, nullptr, nullptr,
                                 brdf, u10m, v10m, mr);
  if (e || !rc || *rc) return e;
  // BRDF_Q = BRDF_U = 0 where the driver ran the surface (GEO/VLIDORT_BRDF_CX.F90:724-725)
  const size_t no = (size_t)nobs * nch;
  for (size_t i = 0; i < no; ++i) {
    const bool miss = std::fabs(brdf[i] / missing - 1.0) < 0.001;
    brdf_q[i] = brdf_u[i] = miss ? missing : 0.0;
  }
  return 0;
}

int aa_vlidort_vector_lambert_bpdf(int km, int nch, int nobs, const double* channels, int nstreams,
                                   int plane_parallel, int nmom, int npol, const double* rot, const double* depol,
                                   const double* tau, const double* ssa, const double* pmom, const double* pe,
                                   const double* he, const double* te, const double* albedo, const double* bpdf_param,
                                   const double* solar_zenith, const double* relat_azymuth,
                                   const double* sensor_zenith, double missing, int verbose, int precision,
                                   double* radiance, double* reflectance, double* br, double* q, double* u,
                                   double* br_q, double* br_u, int* nfourier, int* rc) {
  (void)channels; (void)pe; (void)te; (void)verbose;
  if (!albedo || !bpdf_param) return AA_EARG;
  return aa::lambert_host(true, km, nch, nobs, nstreams, plane_parallel, nmom, npol, rot, depol, nullptr, tau, ssa,
                          pmom, he, albedo, solar_zenith, relat_azymuth, sensor_zenith, nullptr, missing, precision,
                          radiance, reflectance, q, u, nfourier, rc, nullptr, nullptr, br, nullptr, nullptr, nullptr,
                          false, nullptr, bpdf_param, br_q, br_u);
}

int aa_vlidort_vector_landmodis_bpdf(int km, int nch, int nobs, const double* channels, int nstreams,
                                     int plane_parallel, int nmom, int npol, const double* rot, const double* depol,
                                     const double* tau, const double* ssa, const double* pmom, const double* pe,
                                     const double* he, const double* te, const double* kernel_wt,
                                     const double* rtls_param, const double* bpdf_param, const double* solar_zenith,
                                     const double* relat_azymuth, const double* sensor_zenith, double missing,
                                     int verbose, int precision, double* radiance, double* reflectance, double* br,
                                     double* q, double* u, double* br_q, double* br_u, int* nfourier, int* rc) {
  (void)channels; (void)pe; (void)te; (void)verbose;
  if (!kernel_wt || !rtls_param || !bpdf_param) return AA_EARG;
  return aa::lambert_host(true, km, nch, nobs, nstreams, plane_parallel, nmom, npol, rot, depol, nullptr, tau, ssa,
                          pmom, he, nullptr, solar_zenith, relat_azymuth, sensor_zenith, nullptr, missing, precision,
                          radiance, reflectance, q, u, nfourier, rc, kernel_wt, rtls_param, br, nullptr, nullptr,
                          nullptr, false, nullptr, bpdf_param, br_q, br_u);
}

int aa_vlidort_scalar_landmodis(int km, int nch, int nobs, const double* channels, int nstreams, int plane_parallel,
                                int nmom, int npol, const double* rot, const double* depol, const double* tau,
                                const double* ssa, const double* g, const double* pmom, const double* pe,
                                const double* he, const double* te, const double* kernel_wt, const double* param,
                                const double* solar_zenith, const double* relat_azymuth, const double* sensor_zenith,
                                double missing, int verbose, int precision, double* radiance, double* reflectance,
                                double* br, int* nfourier, int* rc) {
  (void)channels; (void)g; (void)pe; (void)te; (void)verbose;
  if (!kernel_wt) return AA_EARG;
  return aa::lambert_host(false, km, nch, nobs, nstreams, plane_parallel, nmom, npol, rot, depol, nullptr, tau, ssa,
                          pmom, he, nullptr, solar_zenith, relat_azymuth, sensor_zenith, nullptr, missing, precision,
                          radiance, reflectance, nullptr, nullptr, nfourier, rc, kernel_wt, param, br);
}

}  // extern "C"

namespace aa {
namespace {

// Fused device path shared by the Lambertian and the MODIS land-surface entry
// points (d_kernel_wt set: BRDF surface, d_albedo unused).
// Optics blocks of a one-stream-count call on two library streams (see the
// caller).  Each stream has its own optics block and work arena; block b + 1
// waits for block b's mid-loop event, block b + 2 follows block b on its
// stream.  The caller stream waits for both streams before this returns.
constexpr int BLOCK_MID_TERM = 4;
int blocks_two_streams(aa_mie* t, const RTWork& W, hipStream_t st, int nblk, size_t ob, int nq,
                       const int* tracer_idx, const int* chan_idx, const float* d_qm, const float* d_rh,
                       size_t obytes) {
  const int km = W.km, nch = W.nch, nobs = W.nobs;
  hipStream_t ss[2] = {bucket_stream(st, 0), bucket_stream(st, 1)};
  if (!ss[0] || !ss[1]) return AA_EHIP;
  float* opt[2];
  for (int k = 0; k < 2; ++k) {
    char* base = nullptr;
    if (arena_reserve(ss[k], AR_OPTICS, obytes, &base)) return AA_EHIP;
    opt[k] = (float*)base;
  }
  const size_t n3 = ob * km * nch;
  hipEvent_t ready = nullptr, mid[2] = {nullptr, nullptr}, done[2] = {nullptr, nullptr};
  int rcode = 0;
  auto mk = [&](hipEvent_t* e) {
    if (!rcode && hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess) rcode = AA_EHIP;
  };
  mk(&ready); mk(&mid[0]); mk(&mid[1]); mk(&done[0]); mk(&done[1]);
  if (!rcode && hipEventRecord(ready, st) != hipSuccess) rcode = AA_EHIP;
  for (int k = 0; k < 2 && !rcode; ++k)
    if (hipStreamWaitEvent(ss[k], ready, 0) != hipSuccess) rcode = AA_EHIP;
  for (int b = 0; b < nblk && !rcode; ++b) {
    const int k = b & 1;
    hipStream_t s = ss[k];
    if (b > 0 && hipStreamWaitEvent(s, mid[k ^ 1], 0) != hipSuccess) { rcode = AA_EHIP; break; }
    const int o0 = (int)(b * ob), no = std::min((int)ob, nobs - o0);
    float* dt = opt[k];
    float* ds = dt + n3;
    float* dg = ds + n3;
    float* dp = dg + n3;
    rcode = launch_get_aop(t, km, no, nch, nq, tracer_idx, chan_idx, d_qm + (size_t)o0 * km * nq,
                           d_rh + (size_t)o0 * km, W.nmom, W.npol, nq, 1, dt, ds, dg, dp, s);
    if (rcode) break;
    RTWork Wc = W;
    Wc.tau = dt; Wc.ssa = ds; Wc.pmom = dp;
    Wc.opt_obs0 = o0; Wc.opt_nobs = no;
    Wc.mid_event = mid[k]; Wc.mid_term = BLOCK_MID_TERM;
    rcode = run_items(Wc, o0 * nch, (o0 + no) * nch, s);
  }
  // join both streams to the caller stream, also after an error
  for (int k = 0; k < 2; ++k)
    if (done[k] && hipEventRecord(done[k], ss[k]) == hipSuccess) (void)hipStreamWaitEvent(st, done[k], 0);
  for (hipEvent_t e : {ready, mid[0], mid[1], done[0], done[1]}) if (e) (void)hipEventDestroy(e);
  return rcode;
}

int column_radiance(aa_mie* t, int km, int nobs, int nch, int nq, const int* tracer_idx, const int* chan_idx,
                    const float* d_qm, const float* d_rh, int nmom, int npol, int nstreams, int vector, int precision,
                    const int* d_nlayers, const int* d_nstreams_col, const double* d_rot, const double* d_depol,
                    const double* d_he, const double* d_albedo, const double* d_kernel_wt, const double* d_param,
                    double* d_br, const double* d_sza, const double* d_raa, const double* d_vza,
                    const double* d_flux_factor, double missing, double* d_radiance, double* d_reflectance,
                    double* d_q, double* d_u, int* d_nfourier, void* stream) {
  const bool modis = d_kernel_wt != nullptr;
  if (!t || km <= 0 || km > MAX_LAYERS || nobs < 0 || nch <= 0 || nq <= 0 || nmom <= 0 || npol <= 0) return AA_EARG;
  if (!tracer_idx || !chan_idx || !d_qm || !d_rh || !d_rot || !d_depol || !d_he || (!d_albedo && !modis) || !d_sza ||
      !d_raa || !d_vza || !d_radiance || !d_reflectance || !d_nfourier)
    return AA_EARG;
  if (modis && (vector || !d_param || !d_br)) return AA_EARG;
  if (precision != 0 && precision != 1) return AA_EARG;
  const int S = vector ? 3 : 1;
  if (vector && (!d_q || !d_u || npol < 6)) return AA_EARG;
  if (!d_nstreams_col && (nstreams < 1 || nstreams > max_streams(S))) return AA_EDIMS;
  if (nobs == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  if (mie_upload(t)) return AA_EHIP;
  // stream counts per column -> buckets (items grouped per optics block, then per N)
  std::vector<int> nstr(nobs, nstreams);
  if (d_nstreams_col) {
    AA_CHECK(hipMemcpyAsync(nstr.data(), d_nstreams_col, sizeof(int) * nobs, hipMemcpyDeviceToHost, st));
    AA_CHECK(hipStreamSynchronize(st));
    for (int v : nstr) if (v < 1 || v > max_streams(S)) return AA_EDIMS;
  }
  RTWork W;
  W.km = km; W.nch = nch; W.nobs = nobs; W.nmom = nmom; W.npol = npol;
  W.N = nstreams; W.S = S; W.nmoments = std::min(2 * nstreams - 1, nmom);
  W.missing = missing;
  W.rot = d_rot; W.depol = d_depol; W.alpha = nullptr;
  W.he = d_he; W.albedo = modis ? nullptr : d_albedo; W.sza = d_sza; W.raa = d_raa; W.vza = d_vza;
  W.kernel_wt = d_kernel_wt; W.kparam = d_param; W.br_out = d_br;
  W.flux = vector ? d_flux_factor : nullptr;
  W.nlayers = d_nlayers;
  W.radiance = d_radiance; W.reflectance = d_reflectance; W.qout = d_q; W.uout = d_u; W.nfourier = d_nfourier;
  W.optics_f32 = 1;
  W.precision = precision;
  // optics blocks split evenly; with one stream count they also fit one work
  // chunk, so that each block's items run as one chunk (per-column stream
  // counts split a block into buckets, each chunked on its own)
  const size_t per_obs = (size_t)km * nch * (3 + (size_t)nmom * npol) * sizeof(float);
  size_t ob = std::max<size_t>(1, optics_budget() / per_obs);
  // one stream count and AA_BLOCK_STREAMS=2: blocks run on two library
  // streams (below), each with half the optics and work budgets
  const bool two = !d_nstreams_col && env_bytes("AA_BLOCK_STREAMS", 1) >= 2;
  if (two) ob = std::max<size_t>(1, ob / 2);
  if (!d_nstreams_col) {
    size_t wcap = 1;
    if (work_cap_items(W, &wcap)) return AA_EHIP;
    if (two) wcap = std::max<size_t>(1, wcap / 2);
    ob = std::min<size_t>(ob, std::max<size_t>(1, wcap / (size_t)nch));
  }
  ob = std::min<size_t>(ob, (size_t)nobs);
  int nblk = (int)((nobs + ob - 1) / ob);
  ob = ((size_t)nobs + nblk - 1) / nblk;
  nblk = (int)((nobs + ob - 1) / ob);
  // host item lists: for block b and stream count N, items [off, off+cnt)
  struct Bucket { int blk, N, off, cnt; };
  std::vector<Bucket> buckets;
  std::vector<int> hitems;
  if (d_nstreams_col) {
    hitems.reserve((size_t)nobs * nch);
    for (int b = 0; b < nblk; ++b) {
      const int o0 = (int)(b * ob), o1 = std::min(nobs, (int)((b + 1) * ob));
      for (int N = 1; N <= max_streams(S); ++N) {
        const int off = (int)hitems.size();
        for (int o = o0; o < o1; ++o)
          if (nstr[o] == N) for (int c = 0; c < nch; ++c) hitems.push_back(o * nch + c);
        const int cnt = (int)hitems.size() - off;
        if (cnt) buckets.push_back({b, N, off, cnt});
      }
    }
  }
  // call arena: latch + item list; optics arena: float optics of one block
  char* cb = nullptr;
  const size_t call_bytes = 256 + sizeof(int) * std::max<size_t>(1, hitems.size());
  if (arena_reserve(st, AR_CALL, call_bytes, &cb)) return AA_EHIP;
  int* lp = (int*)cb;
  int* d_items = (int*)(cb + 256);
  if (!hitems.empty())
    AA_CHECK(hipMemcpyAsync(d_items, hitems.data(), sizeof(int) * hitems.size(), hipMemcpyHostToDevice, st));
  const size_t n3 = ob * km * nch;
  const size_t obytes = (3 * n3 + n3 * (size_t)nmom * npol) * sizeof(float) + 1024;
  int rcode = compute_latch(W, lp, st);
  // With AA_BLOCK_STREAMS=2 (one stream count), consecutive optics blocks
  // alternate between two library streams, each with its own optics block and
  // work arena (half the budgets each), and block b + 1 starts once block b
  // has enqueued its Fourier term BLOCK_MID_TERM, so that its first, full
  // terms fill the GPU while block b's last terms, with few items left,
  // finish: 2% more columns/s on 500,000 columns x 4 channels, but the
  // kernels of the two streams overlap, so per-kernel times no longer add up
  // to the step.  By default the blocks run on the caller stream one after
  // the other.
  if (!rcode && two && nblk > 1) {
    rcode = blocks_two_streams(t, W, st, nblk, ob, nq, tracer_idx, chan_idx, d_qm, d_rh, obytes);
    const hipError_t se = hipStreamSynchronize(st);
    return rcode ? rcode : (se != hipSuccess ? AA_EHIP : 0);
  }
  char* obase = nullptr;
  if (!rcode && arena_reserve(st, AR_OPTICS, obytes, &obase)) rcode = AA_EHIP;
  float* dt = (float*)obase;
  float* ds = dt + n3;
  float* dg = ds + n3;
  float* dp = dg + n3;
  W.tau = dt; W.ssa = ds; W.pmom = dp;
  for (int b = 0; b < nblk && !rcode; ++b) {
    const int o0 = (int)(b * ob), no = std::min((int)ob, nobs - o0);
    rcode = launch_get_aop(t, km, no, nch, nq, tracer_idx, chan_idx, d_qm + (size_t)o0 * km * nq,
                           d_rh + (size_t)o0 * km, nmom, npol, nq, 1, dt, ds, dg, dp, st);
    if (rcode) break;
    RTWork Wc = W;
    Wc.opt_obs0 = o0; Wc.opt_nobs = no;
    if (!d_nstreams_col) {
      rcode = run_items(Wc, o0 * nch, (o0 + no) * nch, st);
      continue;
    }
    // stream-count buckets of this block run concurrently on library streams
    hipEvent_t ready;
    AA_CHECK(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
    AA_CHECK(hipEventRecord(ready, st));
    std::vector<hipEvent_t> done;
    int k = 0;
    for (const Bucket& bk : buckets) {
      if (bk.blk != b) continue;
      hipStream_t bs = bucket_stream(st, k++);
      if (!bs) { rcode = AA_EHIP; break; }
      AA_CHECK(hipStreamWaitEvent(bs, ready, 0));
      RTWork Wb = Wc;
      Wb.N = bk.N; Wb.nmoments = std::min(2 * bk.N - 1, nmom);
      rcode = run_items(Wb, bk.off, bk.off + bk.cnt, bs, d_items);
      hipEvent_t e;
      AA_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      AA_CHECK(hipEventRecord(e, bs));
      done.push_back(e);
      if (rcode) break;
    }
    for (hipEvent_t e : done) { AA_CHECK(hipStreamWaitEvent(st, e, 0)); (void)hipEventDestroy(e); }
    (void)hipEventDestroy(ready);
  }
  AA_CHECK(hipStreamSynchronize(st));
  return rcode;
}

// ------------------------------------------------------------ fixed scratch
// Work memory of aa_column_radiance_dev_noalloc: a file-scope device array,
// one per GPU (each GPU the library runs on holds it until the process ends).
// Host code must not use the array's name as an address: its device address
// is looked up once per device (hipGetSymbolAddress with that device
// current) and only that pointer is passed to kernels and copies.
constexpr size_t SCRATCH_BYTES = (size_t)512 << 20;
constexpr int SCRATCH_MAX_DEVICES = 64;
__device__ __align__(256) char scratch[SCRATCH_BYTES];
std::mutex g_scratch_mu;
char* g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

int scratch_ptr(char** out) {
  int dev = 0;
  AA_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return AA_EARG;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return AA_EHIP;
    g_scratch_ptr[dev] = (char*)p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

// aa_column_radiance_dev with one stream count for all columns, its work
// memory carved from the fixed scratch: [call block | optics block | work].
// Columns go through in optics blocks of ob columns whose items fit the
// work part in one chunk; blocks run one after another on the stream.
int column_radiance_noalloc(aa_mie* t, int km, int nobs, int nch, int nq, const int* tracer_idx,
                            const int* chan_idx, const float* d_qm, const float* d_rh, int nmom, int npol,
                            int nstreams, int vector, int precision, const int* d_nlayers, const double* d_rot,
                            const double* d_depol, const double* d_he, const double* d_albedo, const double* d_sza,
                            const double* d_raa, const double* d_vza, const double* d_flux_factor, double missing,
                            double* d_radiance, double* d_reflectance, double* d_q, double* d_u, int* d_nfourier,
                            void* stream) {
  if (!t || km <= 0 || km > MAX_LAYERS || nobs < 0 || nch <= 0 || nq <= 0 || nmom <= 0 || npol <= 0) return AA_EARG;
  if (!tracer_idx || !chan_idx || !d_qm || !d_rh || !d_rot || !d_depol || !d_he || !d_albedo || !d_sza || !d_raa ||
      !d_vza || !d_radiance || !d_reflectance || !d_nfourier)
    return AA_EARG;
  if (precision != 0 && precision != 1) return AA_EARG;
  const int S = vector ? 3 : 1;
  if (vector && (!d_q || !d_u || npol < 6)) return AA_EARG;
  if (nstreams < 1 || nstreams > max_streams(S)) return AA_EDIMS;
  // uploading the tables would allocate: they must already be resident
  if (!mie_resident(t)) return AA_ENOTRES;
  if (nobs == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  char* base = nullptr;
  if (int e = scratch_ptr(&base)) return e;
  RTWork W;
  W.km = km; W.nch = nch; W.nobs = nobs; W.nmom = nmom; W.npol = npol;
  W.N = nstreams; W.S = S; W.nmoments = std::min(2 * nstreams - 1, nmom);
  W.missing = missing;
  W.rot = d_rot; W.depol = d_depol; W.alpha = nullptr;
  W.he = d_he; W.albedo = d_albedo; W.sza = d_sza; W.raa = d_raa; W.vza = d_vza;
  W.flux = vector ? d_flux_factor : nullptr;
  W.nlayers = d_nlayers;
  W.radiance = d_radiance; W.reflectance = d_reflectance; W.qout = d_q; W.uout = d_u; W.nfourier = d_nfourier;
  W.optics_f32 = 1;
  W.precision = precision;
  // block size: ob columns' float optics + their items' work in one chunk
  const size_t call_bytes = 256;
  const size_t per_obs = (size_t)km * nch * (3 + (size_t)nmom * npol) * sizeof(float);
  const size_t per_item = bytes_per_item(W) + 2 * sizeof(int);
  const size_t tables = (size_t)8 << 20;   // quadrature / contraction tables and counters (<= 4.2 MB)
  if (SCRATCH_BYTES <= call_bytes + tables + per_obs + (size_t)nch * per_item + 1024) return AA_EDIMS;
  size_t ob = (SCRATCH_BYTES - call_bytes - tables - 1024) / (per_obs + (size_t)nch * per_item);
  ob = std::min<size_t>(ob, (size_t)nobs);
  const int nblk = (int)((nobs + ob - 1) / ob);
  ob = ((size_t)nobs + nblk - 1) / nblk;
  int* lp = (int*)base;
  const size_t n3 = ob * km * nch;
  float* dt = (float*)(base + call_bytes);
  float* ds = dt + n3;
  float* dg = ds + n3;
  float* dp = dg + n3;
  const size_t obytes = ((3 * n3 + n3 * (size_t)nmom * npol) * sizeof(float) + 255) & ~(size_t)255;
  char* wbase = base + call_bytes + obytes;
  const size_t wbytes = SCRATCH_BYTES - call_bytes - obytes;
  W.tau = dt; W.ssa = ds; W.pmom = dp;
  int rcode = compute_latch(W, lp, st);
  for (int b = 0; b < nblk && !rcode; ++b) {
    const int o0 = (int)(b * ob), no = std::min((int)ob, nobs - o0);
    rcode = launch_get_aop(t, km, no, nch, nq, tracer_idx, chan_idx, d_qm + (size_t)o0 * km * nq,
                           d_rh + (size_t)o0 * km, nmom, npol, nq, 1, dt, ds, dg, dp, st);
    if (rcode) break;
    RTWork Wc = W;
    Wc.opt_obs0 = o0; Wc.opt_nobs = no;
    rcode = run_items(Wc, o0 * nch, (o0 + no) * nch, st, nullptr, wbase, wbytes);
  }
  return rcode;
}

}  // namespace
}  // namespace aa

extern "C" {

int aa_set_arena_bytes(size_t work_bytes, size_t optics_bytes) {
  aa::g_work_bytes.store(work_bytes);
  aa::g_optics_bytes.store(optics_bytes);
  return 0;
}

int aa_release_arenas(void) {
  int dev = 0;
  AA_CHECK(hipGetDevice(&dev));
  AA_CHECK(hipDeviceSynchronize());
  std::lock_guard<std::mutex> lk(aa::g_arena_mu);
  for (auto it = aa::g_arenas.begin(); it != aa::g_arenas.end();) {
    if (std::get<0>(it->first) == dev) {
      if (it->second.base) AA_CHECK(hipFree(it->second.base));
      it = aa::g_arenas.erase(it);
    } else {
      ++it;
    }
  }
  return 0;
}

int aa_column_radiance_dev(aa_mie* t, int km, int nobs, int nch, int nq, const int* tracer_idx, const int* chan_idx,
                           const float* d_qm, const float* d_rh, int nmom, int npol, int nstreams, int vector,
                           int precision, const int* d_nlayers, const int* d_nstreams_col, const double* d_rot,
                           const double* d_depol, const double* d_pe, const double* d_he, const double* d_te,
                           const double* d_albedo, const double* d_sza, const double* d_raa, const double* d_vza,
                           const double* d_flux_factor, double missing, double* d_radiance, double* d_reflectance,
                           double* d_q, double* d_u, int* d_nfourier, void* stream) {
  (void)d_pe; (void)d_te;
  if (!d_albedo) return AA_EARG;
  return aa::column_radiance(t, km, nobs, nch, nq, tracer_idx, chan_idx, d_qm, d_rh, nmom, npol, nstreams, vector,
                             precision, d_nlayers, d_nstreams_col, d_rot, d_depol, d_he, d_albedo, nullptr, nullptr,
                             nullptr, d_sza, d_raa, d_vza, d_flux_factor, missing, d_radiance, d_reflectance, d_q,
                             d_u, d_nfourier, stream);
}

int aa_column_radiance_modis_dev(aa_mie* t, int km, int nobs, int nch, int nq, const int* tracer_idx,
                                 const int* chan_idx, const float* d_qm, const float* d_rh, int nmom, int nstreams,
                                 int precision, const int* d_nlayers, const int* d_nstreams_col, const double* d_rot,
                                 const double* d_depol, const double* d_pe, const double* d_he, const double* d_te,
                                 const double* d_kernel_wt, const double* d_param, const double* d_sza,
                                 const double* d_raa, const double* d_vza, double missing, double* d_radiance,
                                 double* d_reflectance, double* d_br, int* d_nfourier, void* stream) {
  (void)d_pe; (void)d_te;
  if (!d_kernel_wt) return AA_EARG;
  return aa::column_radiance(t, km, nobs, nch, nq, tracer_idx, chan_idx, d_qm, d_rh, nmom, 1, nstreams, 0, precision,
                             d_nlayers, d_nstreams_col, d_rot, d_depol, d_he, nullptr, d_kernel_wt, d_param, d_br,
                             d_sza, d_raa, d_vza, nullptr, missing, d_radiance, d_reflectance, nullptr, nullptr,
                             d_nfourier, stream);
}

int aa_column_radiance_dev_noalloc(aa_mie* t, int km, int nobs, int nch, int nq, const int* tracer_idx,
                                   const int* chan_idx, const float* d_qm, const float* d_rh, int nmom, int npol,
                                   int nstreams, int vector, int precision, const int* d_nlayers,
                                   const double* d_rot, const double* d_depol, const double* d_pe,
                                   const double* d_he, const double* d_te, const double* d_albedo,
                                   const double* d_sza, const double* d_raa, const double* d_vza,
                                   const double* d_flux_factor, double missing, double* d_radiance,
                                   double* d_reflectance, double* d_q, double* d_u, int* d_nfourier, void* stream) {
  (void)d_pe; (void)d_te;
  return aa::column_radiance_noalloc(t, km, nobs, nch, nq, tracer_idx, chan_idx, d_qm, d_rh, nmom, npol, nstreams,
                                     vector, precision, d_nlayers, d_rot, d_depol, d_he, d_albedo, d_sza, d_raa,
                                     d_vza, d_flux_factor, missing, d_radiance, d_reflectance, d_q, d_u, d_nfourier,
                                     stream);
}

const char* aa_version(void) { return "aeroapps_amd 0.1 (gfx950)"; }

}  // extern "C"
