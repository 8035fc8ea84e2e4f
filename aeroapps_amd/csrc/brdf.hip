// BRDF supplement of the MODIS land surface on the device: for every item
// (column, channel) the Fourier components M < 2N of the Lambertian +
// Ross-thick + Li-sparse BRDF that the discrete-ordinate solver reads
// (BRDF_F, BRDF_F_0, USER_BRDF_F, USER_BRDF_F_0), the exact direct-bounce
// BRDF and the white-sky-albedo check.
//
// Reference (VB = src/Components/rtms/RTSI/VLIDORT2p8p3/vsup/vbrdf,
// V283 = src/Components/rtms/RTSI/VLIDORT2p8p3):
//   set-up               V283/VLIDORT_SurfaceMod.F90:1376-1596 (VLIDORT_LANDMODIS_SingleGeom)
//   VBRDF_MAINMASTER     VB/vbrdf_sup_masters.f90:1559-3293
//   VBRDF_MAKER          VB/vbrdf_sup_routines.f90:153-687
//   SCALING_FOURIER_ZERO VB/vbrdf_sup_routines.f90:1962-2069 (white-sky albedo)
//   VBRDF_FOURIER        VB/vbrdf_sup_routines.f90:2073-2491
//   quadratures          VB/vbrdf_sup_aux.f90:180-262, 356-461
//
// Mapping: one wavefront per item.  The white-sky albedo of the Li-sparse
// kernel (24 x 24 scaling streams x 25 azimuths) runs one lane per stream
// pair; the Fourier components run one lane per (incident, reflected)
// direction pair, each lane evaluating both kernels at the 25 azimuths once
// and projecting them on every Fourier term.  Sums keep the reference's
// order (DOT_PRODUCT from zero, kernels accumulated Lambertian, Ross-thick,
// Li-sparse); the file is compiled without contraction.
#include <map>
#include <mutex>
#include <vector>

#include "rt_work.h"
#include "rt_device.h"
#include "brdf_dev.h"
#include "../../include/aeroapps_amd.h"

namespace aa {

namespace {

// Table of one N (BrTab layout): azimuth quadrature, Fourier cosine factors,
// white-sky quadrature, quadrature streams, and the white-sky albedo of
// Ross-thick (parameter free, so evaluated once here with the reference's sums).
std::vector<double> build_brdf_table(int N, int A) {
  const BrTab T(N, A);
  std::vector<double> t((size_t)T.size(), 0.0);
  std::vector<double> xb(A), ab(A);
  getquad2(0.0, 1.0, A, xb.data(), ab.data());
  for (int k = 0; k < A; ++k) {
    const double x = PIE * xb[k];
    t[T.x(k)] = x;
    t[T.cx(k)] = cos(x);
    t[T.sx(k)] = sin(x);
    t[T.a(k)] = ab[k];
  }
  for (int M = 0; M < 2 * N; ++M)
    for (int k = 0; k < A; ++k)
      t[T.caz(M, k)] = (M == 0) ? ab[k] : ab[k] * cos((double)M * t[T.x(k)]);
  double sq[BR_NSC], sw[BR_NSC];
  getquad2(0.0, 1.0, BR_NSC, sq, sw);
  for (int i = 0; i < BR_NSC; ++i) {
    t[T.sq(i)] = sq[i];
    t[T.ssin(i)] = sqrt(1.0 - sq[i] * sq[i]);
    t[T.ssw(i)] = sq[i] * sw[i];
  }
  std::vector<double> qs(N), qw(N);
  getquad2(0.0, 1.0, N, qs.data(), qw.data());
  for (int i = 0; i < N; ++i) {
    t[T.qs(i)] = qs[i];
    t[T.qsin(i)] = sqrt(1.0 - qs[i] * qs[i]);
  }
  double help = 0.0;
  for (int i = 0; i < BR_NSC; ++i) {
    double s = 0.0;
    for (int j = 0; j < BR_NSC; ++j) {
      double f = 0.0;
      for (int k = 0; k < A; ++k)
        f = f + ross_thick(t[T.sq(j)], t[T.ssin(j)], t[T.sq(i)], t[T.ssin(i)], t[T.cx(k)]) * t[T.a(k)];
      s = s + f * t[T.ssw(j)];
    }
    help = help + s * t[T.ssw(i)];
  }
  t[T.wsa_ross()] = help * 4.0 / PIE;
  return t;
}

// White-sky albedo of Li-sparse for (h/b, b/r) (SCALING_FOURIER_ZERO, half
// range, then the WSA sums of VBRDF_MAINMASTER), computed by the 64 lanes of
// one wavefront; every lane returns the value.
template <int N>
__device__ double li_wsa(const double* __restrict__ tab, double hb, double br) {
#pragma clang fp contract(off)
  constexpr BrTab T(N);
  __shared__ double sang[BR_NSC][4];        // Li-sparse transforms of the scaling streams
  __shared__ double sbf[BR_NSC * BR_NSC];
  __shared__ double rsum[BR_NSC];
  const int t = threadIdx.x;
  if (t < BR_NSC) {
    const LiAng a = li_angle(br, tab[T.sq(t)], tab[T.ssin(t)]);
    sang[t][0] = a.t; sang[t][1] = a.t2; sang[t][2] = a.x; sang[t][3] = a.sx;
  }
  __syncthreads();
  auto sa = [&](int i) { LiAng a; a.t = sang[i][0]; a.t2 = sang[i][1]; a.x = sang[i][2]; a.sx = sang[i][3]; return a; };
  for (int p = t; p < BR_NSC * BR_NSC; p += 64) {
    const int i = p / BR_NSC, j = p - i * BR_NSC;     // incident j, reflected i
    const LiAng in = sa(j), rf = sa(i);
    double f = 0.0;
#pragma unroll 1
    for (int k = 0; k < BR_NAZ; ++k) f = f + li_sparse(hb, in, rf, tab[T.cx(k)], tab[T.sx(k)]) * tab[T.a(k)];
    sbf[p] = f;
  }
  __syncthreads();
  if (t < BR_NSC) {
    double s = 0.0;
    for (int j = 0; j < BR_NSC; ++j) s = s + sbf[t * BR_NSC + j] * tab[T.ssw(j)];
    rsum[t] = s;
  }
  __syncthreads();
  double help = 0.0;
  for (int i = 0; i < BR_NSC; ++i) help = help + rsum[i] * tab[T.ssw(i)];
  return help * 4.0 / PIE;
}

// One item's supplement.  kw = (fiso, fgeo, fvol), kp = (h/b, b/r).  Lane 0
// writes the direct bounce to *db_out and, when the Fourier components were
// computed, whether the white-sky-albedo check dropped Li-sparse to
// *wsa_fail.  The Fourier components are skipped when `skip` is set or, with
// `stop_negative`, for a negative direct bounce (the MODIS driver does not
// run VLIDORT then).
// wcache (may be null) holds (h/b, b/r, white-sky albedo) of one parameter
// pair; items with bit-identical parameters take the albedo from it.
template <int N>
__device__ void brdf_item(const double* __restrict__ tab, const double* kw, const double* kp, double sza, double vza,
                          double raa, double* __restrict__ rec, double* db_out, int* wsa_fail, bool skip,
                          bool stop_negative, const double* wcache = nullptr) {
#pragma clang fp contract(off)
  constexpr BrTab T(N);
  constexpr BROff RO(N);
  constexpr int NT = N * N + 2 * N + 1;     // direction pairs of the Fourier components
  __shared__ double dang[N + 2][4];         // Li-sparse transforms of the quadrature streams, sun, view
  const int t = threadIdx.x;
  const double fiso = kw[0], fgeo = kw[1], fvol = kw[2], hb = kp[0], br = kp[1];
  const double db = modis_db(fiso, fgeo, fvol, hb, br, sza, vza, raa);
  if (t == 0) {
    *db_out = db;
    rec[RO.db()] = db;
  }
  if (skip || (stop_negative && db < 0.0)) return;   // uniform over the wavefront

  // ---- white-sky albedo of Li-sparse
  const bool cached = wcache && hb == wcache[0] && br == wcache[1];   // uniform over the wavefront
  const double wsa_li = cached ? wcache[2] : li_wsa<N>(tab, hb, br);
  const double mu0 = cos(sza * DEG_TO_RAD), s0 = sqrt(1.0 - mu0 * mu0);
  const double muv = cos(vza * DEG_TO_RAD), sv = sqrt(1.0 - muv * muv);
  if (t < N + 2) {
    const double x = (t < N) ? tab[T.qs(t)] : (t == N ? mu0 : muv);
    const double sx = (t < N) ? tab[T.qsin(t)] : (t == N ? s0 : sv);
    const LiAng a = li_angle(br, x, sx);
    dang[t][0] = a.t; dang[t][1] = a.t2; dang[t][2] = a.x; dang[t][3] = a.sx;
  }
  __syncthreads();
  auto da = [&](int i) { LiAng a; a.t = dang[i][0]; a.t2 = dang[i][1]; a.x = dang[i][2]; a.sx = dang[i][3]; return a; };
  double total = fiso * 1.0;
  total = total + fvol * tab[T.wsa_ross()];
  total = total + fgeo * wsa_li;
  const bool with_li = !(total <= 0.0 || total > 1.0);
  if (t == 0) *wsa_fail = with_li ? 0 : 1;

  // ---- Fourier components (VBRDF_FOURIER, half range: HELP = DELFAC)
  for (int p = t; p < NT; p += 64) {
    double xi, sxi, xj, sxj;
    LiAng in, rf;
    int kind, a0;
    if (p < N * N) {
      const int i = p / N, j = p - i * N;
      xi = tab[T.qs(j)]; sxi = tab[T.qsin(j)]; xj = tab[T.qs(i)]; sxj = tab[T.qsin(i)];
      in = da(j); rf = da(i); kind = 0; a0 = p;
    } else if (p < N * N + N) {
      const int i = p - N * N;
      xi = mu0; sxi = s0; xj = tab[T.qs(i)]; sxj = tab[T.qsin(i)];
      in = da(N); rf = da(i); kind = 1; a0 = i;
    } else if (p < N * N + 2 * N) {
      const int j = p - N * N - N;
      xi = tab[T.qs(j)]; sxi = tab[T.qsin(j)]; xj = muv; sxj = sv;
      in = da(j); rf = da(N + 1); kind = 2; a0 = j;
    } else {
      xi = mu0; sxi = s0; xj = muv; sxj = sv;
      in = da(N); rf = da(N + 1); kind = 3; a0 = 0;
    }
    // per-M sums over the azimuths in the reference's order (k outer keeps
    // every Fourier term's DOT_PRODUCT sequential in k)
    double dr[2 * N], dl[2 * N];
#pragma unroll
    for (int M = 0; M < 2 * N; ++M) { dr[M] = 0.0; dl[M] = 0.0; }
#pragma unroll 1
    for (int k = 0; k < BR_NAZ; ++k) {
      const double kr = ross_thick(xi, sxi, xj, sxj, tab[T.cx(k)]);
      const double kl = li_sparse(hb, in, rf, tab[T.cx(k)], tab[T.sx(k)]);
#pragma unroll
      for (int M = 0; M < 2 * N; ++M) {
        const double c = tab[T.caz(M, k)];
        dr[M] = dr[M] + kr * c;
        dl[M] = dl[M] + kl * c;
      }
    }
#pragma unroll
    for (int M = 0; M < 2 * N; ++M) {
      const double delfac = (M == 0) ? 1.0 : 2.0;
      double v = (M == 0) ? fiso * 1.0 : 0.0;
      v = v + fvol * (delfac * dr[M]);
      if (with_li) v = v + fgeo * (delfac * dl[M]);
      const int o = kind == 0 ? RO.f(M, a0 / N, a0 - (a0 / N) * N)
                  : kind == 1 ? RO.f0(M, a0) : kind == 2 ? RO.uf(M, a0) : RO.uf0(M);
      rec[o] = v;
    }
  }
}

// One item's Cox-Munk supplement (VLIDORT_CoxMunk_SingleGeom: one kernel of
// factor 1, 50 half-range azimuths, shadowing; the white-sky check drops the
// kernel's Fourier terms, leaving them zero).  Same mapping as brdf_item; the
// shadowing terms are computed once per direction.
template <int N>
__device__ void brdf_item_cx(const double* __restrict__ tab, double sigma2, double m2, double sza, double vza,
                             double raa, double* __restrict__ rec, double* db_out, int* wsa_fail, bool skip) {
#pragma clang fp contract(off)
  constexpr int A = BR_NAZ_CM;
  constexpr BrTab T(N, A);
  constexpr BROff RO(N);
  constexpr int NT = N * N + 2 * N + 1;
  __shared__ double ssh[BR_NSC];          // shadowing terms of the scaling streams
  __shared__ double dsh[N + 2];           // ... of the quadrature streams, sun, view
  __shared__ double sbf[BR_NSC * BR_NSC];
  __shared__ double rsum[BR_NSC];
  const int t = threadIdx.x;
  const double db = cm_db(sigma2, m2, sza, vza, raa);
  if (t == 0) {
    *db_out = db;
    rec[RO.db()] = db;
  }
  if (skip) return;   // uniform over the wavefront
  if (t < BR_NSC) ssh[t] = cm_shadow(sigma2, tab[T.sq(t)]);
  const double mu0 = cos(sza * DEG_TO_RAD), s0 = sqrt(1.0 - mu0 * mu0);
  const double muv = cos(vza * DEG_TO_RAD), sv = sqrt(1.0 - muv * muv);
  if (t < N + 2) dsh[t] = cm_shadow(sigma2, (t < N) ? tab[T.qs(t)] : (t == N ? mu0 : muv));
  __syncthreads();
  for (int p = t; p < BR_NSC * BR_NSC; p += 64) {
    const int i = p / BR_NSC, j = p - i * BR_NSC;     // incident j, reflected i
    double f = 0.0;
#pragma unroll 1
    for (int k = 0; k < A; ++k)
      f = f + cox_munk(sigma2, m2, tab[T.sq(j)], tab[T.ssin(j)], tab[T.sq(i)], tab[T.ssin(i)], tab[T.cx(k)], ssh[j],
                       ssh[i]) * tab[T.a(k)];
    sbf[p] = f;
  }
  __syncthreads();
  if (t < BR_NSC) {
    double s = 0.0;
    for (int j = 0; j < BR_NSC; ++j) s = s + sbf[t * BR_NSC + j] * tab[T.ssw(j)];
    rsum[t] = s;
  }
  __syncthreads();
  double help = 0.0;
  for (int i = 0; i < BR_NSC; ++i) help = help + rsum[i] * tab[T.ssw(i)];
  const double total = 0.0 + 1.0 * (help * 4.0 / PIE);
  const bool with_k = !(total <= 0.0 || total > 1.0);
  if (t == 0) *wsa_fail = with_k ? 0 : 1;
  for (int p = t; p < NT; p += 64) {
    double xi, sxi, xj, sxj, shi, shr;
    int o0;
    if (p < N * N) {
      const int i = p / N, j = p - i * N;
      xi = tab[T.qs(j)]; sxi = tab[T.qsin(j)]; xj = tab[T.qs(i)]; sxj = tab[T.qsin(i)];
      shi = dsh[j]; shr = dsh[i]; o0 = RO.f(0, i, j);
    } else if (p < N * N + N) {
      const int i = p - N * N;
      xi = mu0; sxi = s0; xj = tab[T.qs(i)]; sxj = tab[T.qsin(i)];
      shi = dsh[N]; shr = dsh[i]; o0 = RO.f0(0, i);
    } else if (p < N * N + 2 * N) {
      const int j = p - N * N - N;
      xi = tab[T.qs(j)]; sxi = tab[T.qsin(j)]; xj = muv; sxj = sv;
      shi = dsh[j]; shr = dsh[N + 1]; o0 = RO.uf(0, j);
    } else {
      xi = mu0; sxi = s0; xj = muv; sxj = sv;
      shi = dsh[N]; shr = dsh[N + 1]; o0 = RO.uf0(0);
    }
    // stride between Fourier terms of this output in the record
    const int ms = (p < N * N) ? N * N : ((p < N * N + 2 * N) ? N : 1);
    double dc[2 * N];
#pragma unroll
    for (int M = 0; M < 2 * N; ++M) dc[M] = 0.0;
#pragma unroll 1
    for (int k = 0; k < A; ++k) {
      const double kv = cox_munk(sigma2, m2, xi, sxi, xj, sxj, tab[T.cx(k)], shi, shr);
#pragma unroll
      for (int M = 0; M < 2 * N; ++M) dc[M] = dc[M] + kv * tab[T.caz(M, k)];
    }
#pragma unroll
    for (int M = 0; M < 2 * N; ++M) {
      const double delfac = (M == 0) ? 1.0 : 2.0;
      rec[o0 + M * ms] = with_k ? 0.0 + 1.0 * (delfac * dc[M]) : 0.0;
    }
  }
}

template <int N, bool CX>
__global__ __launch_bounds__(64, 2) void k_brdf(RTWork W, int nitems) {
  const int it = blockIdx.x;
  if (it >= nitems) return;
  const int g = global_item(W, it);
  const int obs = g / W.nch, ch = g - obs * W.nch;
  const size_t oi = (size_t)ch * W.nobs + obs;
  const int st = W.status[it];
  if (st == 1) {   // missing geometry: the driver skips the column
    if (threadIdx.x == 0 && W.br_out) W.br_out[oi] = W.missing;
    return;
  }
  const size_t kq = (size_t)obs * W.nch + ch;
  __shared__ double dbv;
  __shared__ int wf;
  if constexpr (!CX) {
    brdf_item<N>(W.brdf_tab, W.kernel_wt + kq * 3, W.kparam + kq * 2, W.sza[obs], W.vza[obs], W.raa[obs],
                 W.brdf + (size_t)it * W.brdf_stride, &dbv, &wf, st != 0, true, W.wsa_cache);
  } else {
    const double m = W.mr[ch];
    brdf_item_cx<N>(W.brdf_tab, cm_sigma2(W.u10m[obs], W.v10m[obs]), m * m, W.sza[obs], W.vza[obs], W.raa[obs],
                    W.brdf + (size_t)it * W.brdf_stride, &dbv, &wf, st != 0);
  }
  __syncthreads();
  if constexpr (CX) {   // the Cox-Munk drivers keep any sign of the direct bounce
    if (threadIdx.x == 0 && W.br_out) W.br_out[oi] = dbv;
    return;
  }
  if (threadIdx.x == 0) {
    // negative direct bounce: the driver writes -500 and skips VLIDORT (status 6)
    if (W.br_out) W.br_out[oi] = (dbv < 0.0) ? -500.0 : dbv;
    if (dbv < 0.0) W.status[it] = 6;
  }
}

// White-sky albedo of the first item's Li-sparse parameters into W.wsa_cache
// (one wavefront), reused by k_brdf for every item with the same parameters.
template <int N>
__global__ __launch_bounds__(64) void k_li_wsa_cache(RTWork W) {
  const int g = global_item(W, 0);
  const int obs = g / W.nch, ch = g - obs * W.nch;
  const double* kp = W.kparam + ((size_t)obs * W.nch + ch) * 2;
  const double hb = kp[0], br = kp[1];
  const double w = li_wsa<N>(W.brdf_tab, hb, br);
  if (threadIdx.x == 0) { W.wsa_cache[0] = hb; W.wsa_cache[1] = br; W.wsa_cache[2] = w; }
}

// Supplement alone for explicit geometries (aa_vbrdf_modis): item g uses
// geometry g and its own kernel weights; the record goes to rec + g * stride.
template <int N>
__global__ __launch_bounds__(64) void k_brdf_only(int n, const double* tab, const double* kw, const double* kp,
                                                  const double* sza, const double* vza, const double* raa,
                                                  double* rec, int stride, double* db, int* ok) {
  const int g = blockIdx.x;
  if (g >= n) return;
  __shared__ double dbv;
  __shared__ int wf;
  brdf_item<N>(tab, kw + (size_t)g * 3, kp + (size_t)g * 2, sza[g], vza[g], raa[g], rec + (size_t)g * stride, &dbv,
               &wf, false, false);
  __syncthreads();
  // the supplement's status: 4 (serious) when the white-sky check failed
  if (threadIdx.x == 0) { db[g] = dbv; ok[g] = wf ? 4 : 0; }
}

// Cox-Munk supplement alone for explicit geometries (aa_vbrdf_coxmunk).
template <int N>
__global__ __launch_bounds__(64) void k_brdf_cx_only(int n, const double* tab, const double* u10m,
                                                     const double* v10m, const double* mr, const double* sza,
                                                     const double* vza, const double* raa, double* rec, int stride,
                                                     double* db, int* ok) {
  const int g = blockIdx.x;
  if (g >= n) return;
  __shared__ double dbv;
  __shared__ int wf;
  brdf_item_cx<N>(tab, cm_sigma2(u10m[g], v10m[g]), mr[g] * mr[g], sza[g], vza[g], raa[g], rec + (size_t)g * stride,
                  &dbv, &wf, false);
  __syncthreads();
  if (threadIdx.x == 0) { db[g] = dbv; ok[g] = wf ? 4 : 0; }
}

}  // namespace

const std::vector<double>& brdf_table_cached(int N, int A) {
  static std::mutex mu;
  static std::map<std::pair<int, int>, std::vector<double>> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto key = std::make_pair(N, A);
  auto it = cache.find(key);
  if (it == cache.end()) it = cache.emplace(key, build_brdf_table(N, A)).first;
  return it->second;
}

int rt_brdf(const RTWork& W, int nitems, hipStream_t st) {
  // (the table in W.brdf_tab was built for W.N and the surface's azimuth count)
  if (nitems <= 0) return 0;
  switch (W.N) {
#define AA_BR(NN)                                                                                    \
  case NN:                                                                                           \
    if (W.kernel_wt && W.wsa_cache) hipLaunchKernelGGL(k_li_wsa_cache<NN>, dim3(1), dim3(64), 0, st, W); \
    if (W.kernel_wt) hipLaunchKernelGGL((k_brdf<NN, false>), dim3(nitems), dim3(64), 0, st, W, nitems);  \
    else hipLaunchKernelGGL((k_brdf<NN, true>), dim3(nitems), dim3(64), 0, st, W, nitems);             \
    break;
    AA_BR(1) AA_BR(2) AA_BR(3) AA_BR(4) AA_BR(5) AA_BR(6) AA_BR(7) AA_BR(8) AA_BR(9) AA_BR(10) AA_BR(11) AA_BR(12) AA_BR(13) AA_BR(14) AA_BR(15) AA_BR(16) AA_BR(17) AA_BR(18) AA_BR(19) AA_BR(20)
#undef AA_BR
    default: return -4;
  }
  AA_CHECK(hipGetLastError());
  return 0;
}

// Device memory of the synchronous host-pointer entry points (vlidort.hip):
// one grow-only block per device, used under host_call_mutex.
std::mutex& host_call_mutex();
int host_call_block(size_t bytes, char** out);

}  // namespace aa

namespace {
// Supplement alone over ngeom geometries: upload, one wavefront per geometry,
// outputs in Fortran order.  cm: Cox-Munk (p1 = u10m, p2 = v10m, p3 = mr, one
// value per geometry) else MODIS (p1 = kernel_wt(3,g), p2 = param(2,g)).
int vbrdf_alone(bool cm, int nstreams, int ngeom, const double* sza, const double* vza, const double* raa,
                const double* p1, const double* p2, const double* p3, double* brdf_f, double* brdf_f0,
                double* user_brdf_f, double* user_brdf_f0, double* dbounce, int* status) {
  using namespace aa;
  if (nstreams < 1 || nstreams > MAX_STREAMS || ngeom < 0) return AA_EARG;
  if (!sza || !vza || !raa || !p1 || !p2 || (cm && !p3) || !brdf_f || !brdf_f0 || !user_brdf_f || !user_brdf_f0 ||
      !dbounce || !status)
    return AA_EARG;
  if (ngeom == 0) return 0;
  const int N = nstreams;
  const BROff RO(N);
  const std::vector<double>& tab = brdf_table_cached(N, cm ? BR_NAZ_CM : BR_NAZ);
  const double* kernel_wt = p1;
  const double* param = p2;
  const size_t ng = (size_t)ngeom;
  const size_t n1 = cm ? ng : 3 * ng, n2 = 2 * ng;   // Cox-Munk: d_kw = u10m, d_kp = (v10m, mr)
  std::lock_guard<std::mutex> hl(host_call_mutex());
  // device copies carved from the host-call block, 256-byte aligned
  const size_t sz[7] = {tab.size() * sizeof(double), 3 * ng * sizeof(double), n1 * sizeof(double),
                        n2 * sizeof(double), ng * RO.size() * sizeof(double), ng * sizeof(double), ng * sizeof(int)};
  size_t off[7], tot = 0;
  for (int i = 0; i < 7; ++i) { off[i] = tot; tot += (sz[i] + 255) & ~(size_t)255; }
  char* base = nullptr;
  if (host_call_block(tot, &base)) return AA_EHIP;
  double *d_tab = (double*)(base + off[0]), *d_g = (double*)(base + off[1]), *d_kw = (double*)(base + off[2]);
  double *d_kp = (double*)(base + off[3]), *d_rec = (double*)(base + off[4]), *d_db = (double*)(base + off[5]);
  int* d_ok = (int*)(base + off[6]);
  int rc = 0;
  auto fail = [&]() { rc = AA_EHIP; };
  if (!rc) {
    bool ok = hipMemcpy(d_tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(d_g, sza, ng * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(d_g + ng, vza, ng * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(d_g + 2 * ng, raa, ng * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(d_kw, kernel_wt, n1 * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(d_kp, param, (cm ? ng : 2 * ng) * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
              (!cm || hipMemcpy(d_kp + ng, p3, ng * sizeof(double), hipMemcpyHostToDevice) == hipSuccess) &&
              hipMemset(d_rec, 0, ng * RO.size() * sizeof(double)) == hipSuccess;
    if (!ok) fail();
  }
  if (!rc) {
    switch (N) {
#define AA_BR(NN)                                                                                               \
  case NN:                                                                                                      \
    if (cm)                                                                                                     \
      hipLaunchKernelGGL(k_brdf_cx_only<NN>, dim3(ngeom), dim3(64), 0, 0, ngeom, d_tab, d_kw, d_kp, d_kp + ng,  \
                         d_g, d_g + ng, d_g + 2 * ng, d_rec, RO.size(), d_db, d_ok);                            \
    else                                                                                                        \
      hipLaunchKernelGGL(k_brdf_only<NN>, dim3(ngeom), dim3(64), 0, 0, ngeom, d_tab, d_kw, d_kp, d_g, d_g + ng, \
                         d_g + 2 * ng, d_rec, RO.size(), d_db, d_ok);                                           \
    break;
      AA_BR(1) AA_BR(2) AA_BR(3) AA_BR(4) AA_BR(5) AA_BR(6) AA_BR(7) AA_BR(8) AA_BR(9) AA_BR(10) AA_BR(11) AA_BR(12) AA_BR(13) AA_BR(14) AA_BR(15) AA_BR(16) AA_BR(17) AA_BR(18) AA_BR(19) AA_BR(20)
#undef AA_BR
    }
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) fail();
  }
  if (!rc) {
    std::vector<double> rec(ng * RO.size());
    std::vector<int> st(ng);
    if (hipMemcpy(rec.data(), d_rec, rec.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(dbounce, d_db, ng * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(st.data(), d_ok, ng * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
      fail();
    } else {
      // outputs in Fortran order: brdf_f(0:2N-1, N, N) etc. per geometry
      const size_t nm = 2 * (size_t)N;
      for (size_t g = 0; g < ng; ++g) {
        const double* r = rec.data() + g * RO.size();
        for (int M = 0; M < 2 * N; ++M) {
          for (int i = 0; i < N; ++i) {
            for (int j = 0; j < N; ++j) brdf_f[g * nm * N * N + (size_t)M + nm * (i + (size_t)N * j)] = r[RO.f(M, i, j)];
            brdf_f0[g * nm * N + (size_t)M + nm * i] = r[RO.f0(M, i)];
            user_brdf_f[g * nm * N + (size_t)M + nm * i] = r[RO.uf(M, i)];
          }
          user_brdf_f0[g * nm + M] = r[RO.uf0(M)];
        }
        status[g] = st[g];
      }
    }
  }
  return rc;
}
}  // namespace

extern "C" int aa_vbrdf_modis(int nstreams, int ngeom, const double* sza, const double* vza, const double* raa,
                              const double* kernel_wt, const double* param, double* brdf_f, double* brdf_f0,
                              double* user_brdf_f, double* user_brdf_f0, double* dbounce, int* status) {
  return vbrdf_alone(false, nstreams, ngeom, sza, vza, raa, kernel_wt, param, nullptr, brdf_f, brdf_f0, user_brdf_f,
                     user_brdf_f0, dbounce, status);
}

extern "C" int aa_vbrdf_coxmunk(int nstreams, int ngeom, const double* sza, const double* vza, const double* raa,
                                const double* u10m, const double* v10m, const double* mr, double* brdf_f,
                                double* brdf_f0, double* user_brdf_f, double* user_brdf_f0, double* dbounce,
                                int* status) {
  return vbrdf_alone(true, nstreams, ngeom, sza, vza, raa, u10m, v10m, mr, brdf_f, brdf_f0, user_brdf_f,
                     user_brdf_f0, dbounce, status);
}
