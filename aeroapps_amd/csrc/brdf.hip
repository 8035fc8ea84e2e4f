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

// Full-range table of the polarised surfaces (BPDF kernels switch the
// supplement to DO_HALF_RANGE = false, VB/vbrdf_sup_masters.f90:2176-2199):
// A = 50 azimuths, the 25 Gauss points of [0, pi] followed by their mirror
// images in [-pi, 0] with the same weights (BRDF_QUADRATURE_GAUSSIAN,
// VB/vbrdf_sup_aux.f90:356-461), cosine and sine Fourier factors, and the
// Ross-thick white-sky albedo over the full range with the factor 1/2 of
// SCALING_FOURIER_ZERO.
std::vector<double> build_brdf_table_full(int N) {
  constexpr int A = 2 * BR_NAZ;
  const BrTab T(N, A);
  std::vector<double> t((size_t)T.size(), 0.0);
  double xh[BR_NAZ], ah[BR_NAZ];
  getquad2(0.0, 1.0, BR_NAZ, xh, ah);
  double xb[A], ab[A];
  for (int k = 0; k < BR_NAZ; ++k) {
    xb[k] = xh[k]; ab[k] = ah[k];
    xb[k + BR_NAZ] = -xh[k]; ab[k + BR_NAZ] = ah[k];
  }
  for (int k = 0; k < A; ++k) {
    const double x = PIE * xb[k];
    t[T.x(k)] = x;
    t[T.cx(k)] = cos(x);
    t[T.sx(k)] = sin(x);
    t[T.a(k)] = ab[k];
  }
  for (int M = 0; M < 2 * N; ++M)
    for (int k = 0; k < A; ++k) {
      const double arg = (double)M * t[T.x(k)];
      t[T.caz(M, k)] = (M == 0) ? ab[k] : ab[k] * cos(arg);
      t[T.saz(M, k)] = (M == 0) ? 0.0 : ab[k] * sin(arg);
    }
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
      s = s + (0.5 * f) * t[T.ssw(j)];
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

// One item's polarised supplement (three Stokes components): the kernels of
// VLIDORT_LAMBERTIAN_BPDF (modis = false: Lambertian of factor `lamb`,
// BPDF-NDVI) or VLIDORT_LANDMODIS_BPDF (modis: Lambertian fiso = lamb,
// Ross-thick fvol, Li-sparse fgeo with (h/b, b/r), BPDF-NDVI) in that order,
// 50 full-range azimuths (V283/VLIDORT_SurfaceMod.F90:1822-2200).  Fourier
// components (VBRDF_FOURIER, VB/vbrdf_sup_routines.f90:2073-2491, full
// range: HELP = DELFAC / 2): cosine sums for the Mueller elements (1,1),
// (1,2), (2,1), (2,2), (3,3), sine sums with sign -HELP for (1,3), (2,3) and
// +HELP for (3,1), (3,2).  The white-sky albedo check comes after the last
// kernel (BPDF): when the total is outside (0, 1] the BPDF's Fourier terms
// are left out, its direct bounce is kept (VB/vbrdf_sup_masters.f90:2950-3090).
// Lane 0 writes the direct-bounce elements (1,1), (1,2), (1,3) to db3 and the
// check's failure to *wsa_fail.  Fourier components are skipped when `skip`
// is set or the (1,1) direct bounce is negative (the drivers do not run
// VLIDORT then).
template <int N>
__device__ void brdf_item_pol(const double* __restrict__ tab, bool modis, double lamb, const double* kw,
                              const double* kp, const double* bp, double sza, double vza, double raa,
                              double* __restrict__ rec, double* db3, int* wsa_fail, bool skip) {
#pragma clang fp contract(off)
  constexpr int A = 2 * BR_NAZ;
  constexpr BrTab T(N, A);
  constexpr BPOff PO(N);
  constexpr int NT = N * N + 2 * N;        // quadrature pairs, solar -> stream, stream -> user
  __shared__ double dang[N + 2][4];        // Li-sparse transforms of the quadrature streams, sun, view
  __shared__ double sang[BR_NSC][4];       // ... of the white-sky streams
  __shared__ double sbf[2][BR_NSC * BR_NSC];
  __shared__ double rsum[2][BR_NSC];
  const int t = threadIdx.x;
  const double fgeo = modis ? kw[1] : 0.0, fvol = modis ? kw[2] : 0.0;
  const double hb = modis ? kp[0] : 0.0, br = modis ? kp[1] : 1.0;
  const double rn = bp[0], ndvi = bp[1], cfac = bp[2];
  const double mu0 = cos(sza * DEG_TO_RAD), s0 = sqrt(1.0 - mu0 * mu0);
  const double muv = cos(vza * DEG_TO_RAD), sv = sqrt(1.0 - muv * muv);
  const double phi = raa * DEG_TO_RAD, cphi = cos(phi), sphi = sin(phi);
  // ---- exact direct bounce, kernels accumulated in order
  double kb[9];
  bpdf_ndvi(rn, ndvi, cfac, true, mu0, s0, muv, sv, cphi, sphi, kb);
  double db[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) db[q] = (q == 0) ? lamb * 1.0 : 0.0;
  if (modis) {
    db[0] = db[0] + fvol * ross_thick(mu0, s0, muv, sv, cphi);
    db[0] = db[0] + fgeo * li_sparse(hb, li_angle(br, mu0, s0), li_angle(br, muv, sv), cphi, sphi);
  }
#pragma unroll
  for (int q = 0; q < 9; ++q) db[q] = db[q] + 1.0 * kb[q];
  if (t == 0) {
#pragma unroll
    for (int q = 0; q < 9; ++q) rec[PO.db(q)] = db[q];
    db3[0] = db[0]; db3[1] = db[1]; db3[2] = db[2];
  }
  if (skip || db[0] < 0.0) return;   // uniform over the wavefront

  // ---- white-sky albedos of Li-sparse and BPDF over the full range
  if (modis && t < BR_NSC) {
    const LiAng a = li_angle(br, tab[T.sq(t)], tab[T.ssin(t)]);
    sang[t][0] = a.t; sang[t][1] = a.t2; sang[t][2] = a.x; sang[t][3] = a.sx;
  }
  if (modis && t < N + 2) {
    const double x = (t < N) ? tab[T.qs(t)] : (t == N ? mu0 : muv);
    const double sx = (t < N) ? tab[T.qsin(t)] : (t == N ? s0 : sv);
    const LiAng a = li_angle(br, x, sx);
    dang[t][0] = a.t; dang[t][1] = a.t2; dang[t][2] = a.x; dang[t][3] = a.sx;
  }
  __syncthreads();
  auto sa = [&](int i) { LiAng a; a.t = sang[i][0]; a.t2 = sang[i][1]; a.x = sang[i][2]; a.sx = sang[i][3]; return a; };
  auto da = [&](int i) { LiAng a; a.t = dang[i][0]; a.t2 = dang[i][1]; a.x = dang[i][2]; a.sx = dang[i][3]; return a; };
  for (int p = t; p < BR_NSC * BR_NSC; p += 64) {
    const int i = p / BR_NSC, j = p - i * BR_NSC;     // incident j, reflected i
    double fl = 0.0, fb = 0.0;
#pragma unroll 1
    for (int k = 0; k < A; ++k) {
      double k4[9];
      bpdf_ndvi(rn, ndvi, cfac, false, tab[T.sq(j)], tab[T.ssin(j)], tab[T.sq(i)], tab[T.ssin(i)], tab[T.cx(k)],
                tab[T.sx(k)], k4);
      fb = fb + k4[0] * tab[T.a(k)];
      if (modis) fl = fl + li_sparse(hb, sa(j), sa(i), tab[T.cx(k)], tab[T.sx(k)]) * tab[T.a(k)];
    }
    sbf[0][p] = 0.5 * fl;
    sbf[1][p] = 0.5 * fb;
  }
  __syncthreads();
  if (t < 2 * BR_NSC) {
    const int w = t / BR_NSC, i = t - w * BR_NSC;
    double sum = 0.0;
    for (int j = 0; j < BR_NSC; ++j) sum = sum + sbf[w][i * BR_NSC + j] * tab[T.ssw(j)];
    rsum[w][i] = sum;
  }
  __syncthreads();
  double hl = 0.0, hbp = 0.0;
  for (int i = 0; i < BR_NSC; ++i) { hl = hl + rsum[0][i] * tab[T.ssw(i)]; hbp = hbp + rsum[1][i] * tab[T.ssw(i)]; }
  const double wsa_li = hl * 4.0 / PIE, wsa_b = hbp * 4.0 / PIE;
  double total = 0.0 + lamb * 1.0;
  if (modis) {
    total = total + fvol * tab[T.wsa_ross()];
    total = total + fgeo * wsa_li;
  }
  total = total + 1.0 * wsa_b;
  const bool with_b = !(total <= 0.0 || total > 1.0);
  if (t == 0) *wsa_fail = with_b ? 0 : 1;

  // ---- Fourier components
  constexpr bool SINE[9] = {false, false, true, false, false, true, true, true, false};
  for (int p = t; p < NT; p += 64) {
    double xi, sxi, xj, sxj;
    int kind, a0;
    LiAng in, rf;
    if (p < N * N) {
      const int i = p / N, j = p - i * N;
      xi = tab[T.qs(j)]; sxi = tab[T.qsin(j)]; xj = tab[T.qs(i)]; sxj = tab[T.qsin(i)];
      kind = 0; a0 = p;
      if (modis) { in = da(j); rf = da(i); }
    } else if (p < N * N + N) {
      const int i = p - N * N;
      xi = mu0; sxi = s0; xj = tab[T.qs(i)]; sxj = tab[T.qsin(i)];
      kind = 1; a0 = i;
      if (modis) { in = da(N); rf = da(i); }
    } else {
      const int j = p - N * N - N;
      xi = tab[T.qs(j)]; sxi = tab[T.qsin(j)]; xj = muv; sxj = sv;
      kind = 2; a0 = j;
      if (modis) { in = da(j); rf = da(N + 1); }
    }
    double dr[2 * N], dl[2 * N], dbp[9][2 * N];
#pragma unroll
    for (int M = 0; M < 2 * N; ++M) {
      dr[M] = 0.0; dl[M] = 0.0;
#pragma unroll
      for (int q = 0; q < 9; ++q) dbp[q][M] = 0.0;
    }
#pragma unroll 1
    for (int k = 0; k < A; ++k) {
      double kv[9];
      bpdf_ndvi(rn, ndvi, cfac, true, xi, sxi, xj, sxj, tab[T.cx(k)], tab[T.sx(k)], kv);
      double kr = 0.0, kl = 0.0;
      if (modis) {
        kr = ross_thick(xi, sxi, xj, sxj, tab[T.cx(k)]);
        kl = li_sparse(hb, in, rf, tab[T.cx(k)], tab[T.sx(k)]);
      }
#pragma unroll
      for (int M = 0; M < 2 * N; ++M) {
        const double c = tab[T.caz(M, k)], sn = tab[T.saz(M, k)];
        if (modis) { dr[M] = dr[M] + kr * c; dl[M] = dl[M] + kl * c; }
#pragma unroll
        for (int q = 0; q < 9; ++q) dbp[q][M] = dbp[q][M] + kv[q] * (SINE[q] ? sn : c);
      }
    }
#pragma unroll
    for (int M = 0; M < 2 * N; ++M) {
      const double help = 0.5 * ((M == 0) ? 1.0 : 2.0);
#pragma unroll
      for (int q = 0; q < 9; ++q) {
        if (kind == 1 && (q % 3) != 0) continue;     // BRDF_F_0: first column only
        double v = (M == 0 && q == 0) ? lamb * 1.0 : 0.0;
        if (modis && q == 0) {
          v = v + fvol * (help * dr[M]);
          v = v + fgeo * (help * dl[M]);
        }
        if (with_b) {
          // (1,3), (2,3): HELPS = -HELP; (3,1), (3,2): -HELPS
          const double loc = (q == 2 || q == 5) ? (-help) * dbp[q][M]
                             : (q == 6 || q == 7) ? -((-help) * dbp[q][M]) : help * dbp[q][M];
          v = v + 1.0 * loc;
        }
        const int o = kind == 0 ? PO.f(M, a0 / N, a0 - (a0 / N) * N, q)
                    : kind == 1 ? PO.f0(M, a0, q / 3) : PO.uf(M, a0, q);
        rec[o] = v;
      }
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

// Polarised land surfaces: the supplement of every item (one wavefront each),
// the direct-bounce outputs BR, BR_Q, BR_U (GEO/VLIDORT_LAMBERT_BPDF.F90:
// 163-179, GEO/VLIDORT_BRDF_MODIS_BPDF.F90:182-197): -500 everywhere, and no
// radiance run (status 6), for a negative (1,1) direct bounce.
template <int N>
__global__ __launch_bounds__(64) void k_brdf_pol(RTWork W, int nitems) {
  const int it = blockIdx.x;
  if (it >= nitems) return;
  const int g = global_item(W, it);
  const int obs = g / W.nch, ch = g - obs * W.nch;
  const size_t oi = (size_t)ch * W.nobs + obs;
  const int st = W.status[it];
  if (st == 1) {   // missing geometry or albedo: the driver skips the element
    if (threadIdx.x == 0) { W.br_out[oi] = W.missing; W.brq_out[oi] = W.missing; W.bru_out[oi] = W.missing; }
    return;
  }
  const size_t kq = (size_t)obs * W.nch + ch;
  const bool modis = W.kernel_wt != nullptr;
  const double lamb = modis ? W.kernel_wt[kq * 3] : W.albedo[oi];
  __shared__ double dbv[3];
  __shared__ int wf;
  brdf_item_pol<N>(W.brdf_tab, modis, lamb, modis ? W.kernel_wt + kq * 3 : nullptr,
                   modis ? W.kparam + kq * 2 : nullptr, W.bpdf + kq * 3, W.sza[obs], W.vza[obs], W.raa[obs],
                   W.brdf + (size_t)it * W.brdf_stride, dbv, &wf, st != 0);
  __syncthreads();
  if (threadIdx.x == 0) {
    const bool neg = dbv[0] < 0.0;
    W.br_out[oi] = neg ? -500.0 : dbv[0];
    W.brq_out[oi] = neg ? -500.0 : dbv[1];
    W.bru_out[oi] = neg ? -500.0 : dbv[2];
    if (neg) W.status[it] = 6;
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

// Polarised supplement alone for explicit geometries (aa_vbrdf_bpdf): item g
// uses geometry g, the Lambertian factor lamb[g] (albedo or fiso), the MODIS
// kernels when kw is given, and bp[3 g ..]; the record goes to rec + g * stride.
template <int N>
__global__ __launch_bounds__(64) void k_brdf_pol_only(int n, const double* tab, const double* lamb, const double* kw,
                                                      const double* kp, const double* bp, const double* sza,
                                                      const double* vza, const double* raa, double* rec, int stride,
                                                      int* ok) {
  const int g = blockIdx.x;
  if (g >= n) return;
  __shared__ double dbv[3];
  __shared__ int wf;
  brdf_item_pol<N>(tab, kw != nullptr, lamb[g], kw ? kw + (size_t)g * 3 : nullptr, kp ? kp + (size_t)g * 2 : nullptr,
                   bp + (size_t)g * 3, sza[g], vza[g], raa[g], rec + (size_t)g * stride, dbv, &wf, false);
  __syncthreads();
  if (threadIdx.x == 0) ok[g] = (dbv[0] < 0.0) ? -1 : (wf ? 4 : 0);
}

}  // namespace

// The full-range table of the polarised surfaces, per N.
const std::vector<double>& brdf_table_full_cached(int N) {
  static std::mutex mu;
  static std::map<int, std::vector<double>> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(N);
  if (it == cache.end()) it = cache.emplace(N, build_brdf_table_full(N)).first;
  return it->second;
}

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
  if (W.bpdf) {
    static_assert(MAX_STREAMS_VECTOR == 6, "one polarised supplement per vector stream count");
    switch (W.N) {
      case 1: hipLaunchKernelGGL(k_brdf_pol<1>, dim3(nitems), dim3(64), 0, st, W, nitems); break;
      case 2: hipLaunchKernelGGL(k_brdf_pol<2>, dim3(nitems), dim3(64), 0, st, W, nitems); break;
      case 3: hipLaunchKernelGGL(k_brdf_pol<3>, dim3(nitems), dim3(64), 0, st, W, nitems); break;
      case 4: hipLaunchKernelGGL(k_brdf_pol<4>, dim3(nitems), dim3(64), 0, st, W, nitems); break;
      case 5: hipLaunchKernelGGL(k_brdf_pol<5>, dim3(nitems), dim3(64), 0, st, W, nitems); break;
      case 6: hipLaunchKernelGGL(k_brdf_pol<6>, dim3(nitems), dim3(64), 0, st, W, nitems); break;
      default: return -4;
    }
    AA_CHECK(hipGetLastError());
    return 0;
  }
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

// The polarised supplement alone: records computed on the device, returned
// in the reference's array shapes with Mueller index q = 4 o1 + o2 (the
// NSTOKES = 3 entries filled, the rest zero).
extern "C" int aa_vbrdf_bpdf(int nstreams, int ngeom, const double* sza, const double* vza, const double* raa,
                             const double* lamb, const double* kernel_wt, const double* rtls_param,
                             const double* bpdf_param, double* brdf_f, double* brdf_f0, double* user_brdf_f,
                             double* dbounce, int* status) {
  using namespace aa;
  if (nstreams < 1 || nstreams > MAX_STREAMS_VECTOR || ngeom < 0) return AA_EARG;
  if (!sza || !vza || !raa || !bpdf_param || (!lamb && !kernel_wt) || (kernel_wt && !rtls_param) || !brdf_f ||
      !brdf_f0 || !user_brdf_f || !dbounce || !status)
    return AA_EARG;
  if (ngeom == 0) return 0;
  const int N = nstreams;
  const BPOff PO(N);
  const std::vector<double>& tab = brdf_table_full_cached(N);
  const size_t ng = (size_t)ngeom;
  std::vector<double> lam(ng);
  for (size_t g = 0; g < ng; ++g) lam[g] = kernel_wt ? kernel_wt[3 * g] : lamb[g];
  std::lock_guard<std::mutex> hl(host_call_mutex());
  const size_t sz[8] = {tab.size() * sizeof(double), 3 * ng * sizeof(double), ng * sizeof(double),
                        3 * ng * sizeof(double), 2 * ng * sizeof(double), 3 * ng * sizeof(double),
                        ng * PO.size() * sizeof(double), ng * sizeof(int)};
  size_t off[8], tot = 0;
  for (int i = 0; i < 8; ++i) { off[i] = tot; tot += (sz[i] + 255) & ~(size_t)255; }
  char* base = nullptr;
  if (host_call_block(tot, &base)) return AA_EHIP;
  double *d_tab = (double*)(base + off[0]), *d_g = (double*)(base + off[1]), *d_l = (double*)(base + off[2]);
  double *d_kw = (double*)(base + off[3]), *d_kp = (double*)(base + off[4]), *d_bp = (double*)(base + off[5]);
  double* d_rec = (double*)(base + off[6]);
  int* d_ok = (int*)(base + off[7]);
  const bool ok = hipMemcpy(d_tab, tab.data(), sz[0], hipMemcpyHostToDevice) == hipSuccess &&
                  hipMemcpy(d_g, sza, ng * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
                  hipMemcpy(d_g + ng, vza, ng * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
                  hipMemcpy(d_g + 2 * ng, raa, ng * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
                  hipMemcpy(d_l, lam.data(), sz[2], hipMemcpyHostToDevice) == hipSuccess &&
                  (!kernel_wt || hipMemcpy(d_kw, kernel_wt, sz[3], hipMemcpyHostToDevice) == hipSuccess) &&
                  (!kernel_wt || hipMemcpy(d_kp, rtls_param, sz[4], hipMemcpyHostToDevice) == hipSuccess) &&
                  hipMemcpy(d_bp, bpdf_param, sz[5], hipMemcpyHostToDevice) == hipSuccess &&
                  hipMemset(d_rec, 0, sz[6]) == hipSuccess;
  if (!ok) return AA_EHIP;
  const double* kw = kernel_wt ? d_kw : nullptr;
  const double* kp = kernel_wt ? d_kp : nullptr;
  switch (N) {
#define AA_BP(NN)                                                                                               \
  case NN:                                                                                                      \
    hipLaunchKernelGGL(k_brdf_pol_only<NN>, dim3(ngeom), dim3(64), 0, 0, ngeom, d_tab, d_l, kw, kp, d_bp, d_g,  \
                       d_g + ng, d_g + 2 * ng, d_rec, PO.size(), d_ok);                                         \
    break;
    AA_BP(1) AA_BP(2) AA_BP(3) AA_BP(4) AA_BP(5) AA_BP(6)
#undef AA_BP
  }
  if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) return AA_EHIP;
  std::vector<double> rec(ng * PO.size());
  std::vector<int> st(ng);
  if (hipMemcpy(rec.data(), d_rec, rec.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(st.data(), d_ok, ng * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
    return AA_EHIP;
  const size_t nm = 2 * (size_t)N;
  for (size_t g = 0; g < ng; ++g) {
    const double* r = rec.data() + g * PO.size();
    double* F = brdf_f + g * nm * 16 * N * N;
    double* F0 = brdf_f0 + g * nm * 16 * N;
    double* UF = user_brdf_f + g * nm * 16 * N;
    std::fill(F, F + nm * 16 * N * N, 0.0);
    std::fill(F0, F0 + nm * 16 * N, 0.0);
    std::fill(UF, UF + nm * 16 * N, 0.0);
    std::fill(dbounce + g * 16, dbounce + g * 16 + 16, 0.0);
    for (int o1 = 0; o1 < 3; ++o1)
      for (int o2 = 0; o2 < 3; ++o2) {
        const int q = 4 * o1 + o2, q9 = 3 * o1 + o2;
        for (int M = 0; M < 2 * N; ++M)
          for (int i = 0; i < N; ++i) {
            for (int j = 0; j < N; ++j) F[M + nm * (q + 16 * (i + (size_t)N * j))] = r[PO.f(M, i, j, q9)];
            if (o2 == 0) F0[M + nm * (q + 16 * (size_t)i)] = r[PO.f0(M, i, o1)];
            UF[M + nm * (q + 16 * (size_t)i)] = r[PO.uf(M, i, q9)];
          }
        dbounce[g * 16 + q] = r[PO.db(q9)];
      }
    status[g] = st[g];
  }
  return 0;
}
