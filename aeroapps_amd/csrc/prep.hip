// Input preparation on the device: the steps either side of the radiance path
// in a geo_vlidort run (GEO/geo_vlidort_AI.F90:399-430).
//
//   k_rot     Rayleigh optical thickness per layer and the depolarisation
//             ratio: VLIDORT_ROT_CALC (GEO/VLIDORT_ROT.F90:19-85) ->
//             VLIDORT_Rayleigh, coef_depol, F_air
//             (RTSI/VLIDORT2p8p3/VLIDORT_ScatMod.F90:143-281); Bodhaine et
//             al. (1999) cross section, Tomasi et al. (2005) King factor.
//   k_edge    edge pressure, height above the surface and temperature from
//             layer air density and pressure thickness: getEdgeVars
//             (CB/MieObs_py.F90:27-83), float32.
//   k_qm      tracer column mass per layer, q * delp / grav: calc_qm
//             (GEO/geo_vlidort.F90:2124-2144, grav GEO/geo_vlidort_pars.F90:47),
//             float32.
//
// The reference writes its constants as default-real (float32) literals in
// real*8 expressions; F32() rounds a literal to float32 and widens it, as the
// reference's arithmetic sees it.  Compiled without floating-point
// contraction, like the reference.
#include "aa_common.h"
#include "../../include/aeroapps_amd.h"

#include <hip/hip_runtime.h>
#include <cmath>
#include <mutex>
#include <vector>

#define F32(x) ((double)(x##f))

namespace aa {
namespace {

// King factor of air (F_air, VLIDORT_ScatMod.F90:262-281); xx2 = 1/lambda^2
// (lambda in micron), c = CO2 parts per volume
__host__ __device__ inline double f_air(double xx2, double c) {
  const double depol1 = F32(1.034) + F32(3.17E-4) * xx2;                               // N2
  const double depol2 = F32(1.096) + F32(1.385E-3) * xx2 + F32(1.448E-4) * xx2 * xx2;  // O2
  const double depol3 = 1.0;                                                           // Ar
  const double depol4 = F32(1.15);                                                     // CO2
  // the first three terms of the denominator are default-real constants
  const double den0 = (double)((78.084f + 20.946f) + 0.934f);
  return (F32(78.084) * depol1 + F32(20.946) * depol2 + F32(0.934) * depol3 + c * F32(100.0) * depol4) /
         (den0 + c * F32(100.0));
}

// CO2 = 360 ppm, a default-real product
__host__ __device__ inline double co2_ppv() { return (double)(360.0f * 1e-6f); }

// depolarisation ratio (coef_depol, VLIDORT_ScatMod.F90:245-260); x in micron
__host__ __device__ inline double coef_depol(double x) {
  const double xx = 1.0 / x;
  const double xx2 = xx * xx;
  const double c = co2_ppv();
  return 6.0 * (f_air(xx2, c) - 1.0) / (7.0 * f_air(xx2, c) + 3.0);
}

// molecular density at standard pressure and temperature (molecules/cm^3)
__host__ __device__ inline double ns_std() {
  const double av = F32(6.0221367e23), ts = F32(288.15);
  return (av / F32(22.4141)) * (F32(273.15) / ts) * (double)(1.0f / 1000);
}

// Rayleigh scattering cross section (cm^2/molecule), x in micron (A, :195-240)
__device__ inline double ray_xsec(double x) {
  const double xx = 1.0 / x;
  const double xx2 = xx * xx;
  const double cm = x / F32(10000.0);
  const double pi = (double)(4.0f * atanf(1.0f));
  const double pi3 = pi * pi * pi;
  double ri = F32(8060.51) + F32(2480990.0) / (F32(132.274) - xx2) + F32(17455.7) / (F32(39.32957) - xx2);
  ri = ri * F32(1e-8) + 1.0;
  const double c = co2_ppv();
  ri = (ri - 1.0) * (1.0 + F32(0.54) * (c - F32(0.0003))) + 1.0;
  const double ri2 = ri * ri;
  const double ns = ns_std();
  return 24.0 * pi3 * f_air(xx2, c) * ((ri2 - 1.0) * (ri2 - 1.0)) /
         (ns * ns * cm * cm * cm * cm * (ri2 + 2.0) * (ri2 + 2.0));
}

// volume scattering coefficient (km^-1) at one edge
__device__ inline double vol_coef(double a, double pe_pa, double te_k) {
  const double ps = F32(1013.25), ts = F32(288.15);
  const double p = pe_pa * F32(1.E-2);   // hPa
  double v = a * ns_std() * (p / ps) * (ts / te_k);
  return v * F32(1.E5);
}

// one thread per (layer, column) of one channel: ROT(:,:,channel)
__global__ void __launch_bounds__(256) k_rot(int km, int nobs, double lam_um, const double* __restrict__ pe,
                                             const double* __restrict__ he, const double* __restrict__ te,
                                             double* __restrict__ rot) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)km * nobs) return;
  const int obs = (int)(e / km), k = (int)(e - (long)obs * km);
  const double a = ray_xsec(lam_um);
  const size_t c0 = (size_t)obs * (km + 1);
  const double v0 = vol_coef(a, pe[c0 + k], te[c0 + k]);
  const double v1 = vol_coef(a, pe[c0 + k + 1], te[c0 + k + 1]);
  const double difz = he[c0 + k] * F32(1.E-3) - he[c0 + k + 1] * F32(1.E-3);   // km
  rot[e] = difz * (v0 + v1) / 2.0;
}

// one thread per column: getEdgeVars, float32 in the reference's order
__global__ void __launch_bounds__(256) k_edge(int km, int nobs, const float* __restrict__ airdens,
                                              const float* __restrict__ delp, float ptop, float* __restrict__ pe,
                                              float* __restrict__ ze, float* __restrict__ te) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nobs) return;
  const float grav = 9.80616f, rgas = 287.f, kappa = 2.0f / 7.0f;
  const float* ad = airdens + (size_t)n * km;
  const float* dp = delp + (size_t)n * km;
  float* pen = pe + (size_t)n * (km + 1);
  float* zen = ze + (size_t)n * (km + 1);
  float* ten = te + (size_t)n * (km + 1);
  pen[0] = ptop;
  for (int k = 0; k < km; ++k) pen[k + 1] = __fadd_rn(pen[k], dp[k]);
  zen[km] = 0.0f;
  for (int k = km - 1; k >= 0; --k) zen[k] = __fadd_rn(zen[k + 1], __fdiv_rn(dp[k], __fmul_rn(ad[k], grav)));
  // layer-mean pressure and temperature, edge temperature by log-pressure
  // interpolation, isothermal top, dry adiabatic bottom
  float pm_prev = 0.0f, tm_prev = 0.0f;
  for (int k = 0; k < km; ++k) {
    const float pm = __fdiv_rn(__fadd_rn(pen[k], pen[k + 1]), 2.0f);
    const float tm = __fdiv_rn(pm, __fmul_rn(ad[k], rgas));
    if (k == 0) {
      ten[0] = tm;
    } else {
      const float alpha = __fdiv_rn(logf(__fdiv_rn(pen[k], pm_prev)), logf(__fdiv_rn(pm, pm_prev)));
      ten[k] = __fadd_rn(tm_prev, __fmul_rn(alpha, __fsub_rn(tm, tm_prev)));
    }
    if (k == km - 1) ten[km] = __fmul_rn(tm, powf(__fdiv_rn(pen[km], pm), kappa));
    pm_prev = pm;
    tm_prev = tm;
  }
}

// one thread per (layer, tracer, column): qm(km,nq,nobs) = q * delp / grav
__global__ void __launch_bounds__(256) k_qm(int km, int nq, int nobs, const float* __restrict__ q,
                                            const float* __restrict__ delp, float* __restrict__ qm) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)km * nq * nobs) return;
  const int k = (int)(e % km);
  const long n = e / ((long)km * nq);
  qm[e] = __fdiv_rn(__fmul_rn(q[e], delp[(size_t)n * km + k]), 9.81f);
}

int rot_launch(int km, int nch, int nobs, const double* channels_nm, const double* d_pe, const double* d_he,
               const double* d_te, double* d_rot, double* depol, hipStream_t st) {
  const long ne = (long)km * nobs;
  for (int i = 0; i < nch; ++i) {
    const double lam = channels_nm[i] * F32(1.E-3);   // micron
    depol[i] = coef_depol(lam);
    if (ne == 0) continue;
    hipLaunchKernelGGL(k_rot, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, km, nobs, lam, d_pe, d_he, d_te,
                       d_rot + (size_t)i * ne);
  }
  AA_CHECK(hipGetLastError());
  return 0;
}

}  // namespace

// Device memory of the synchronous host-pointer entry points (vlidort.hip):
// one grow-only block per device, used under host_call_mutex.
std::mutex& host_call_mutex();
int host_call_block(size_t bytes, char** out);

// The host-call block split into 256-byte-aligned float arrays of the given
// lengths; false when the block cannot be had.  Caller holds host_call_mutex.
template <int K>
bool host_floats(const size_t (&len)[K], float* (&out)[K]) {
  size_t off[K], tot = 0;
  for (int i = 0; i < K; ++i) { off[i] = tot; tot += (len[i] * sizeof(float) + 255) & ~(size_t)255; }
  char* base = nullptr;
  if (host_call_block(tot, &base)) return false;
  for (int i = 0; i < K; ++i) out[i] = (float*)(base + off[i]);
  return true;
}
}  // namespace aa

extern "C" {

int aa_rot_calc(int km, int nch, int nobs, const double* channels, const double* pe, const double* he,
                const double* te, double missing, int verbose, double* rot, double* depol, int* rc) {
  (void)missing; (void)verbose;
  if (!rc) return AA_EARG;
  *rc = 0;
  if (km <= 0 || nch <= 0 || nobs < 0 || !channels || !pe || !he || !te || !rot || !depol) return AA_EARG;
  if (km > aa::MAX_LAYERS) { *rc = 3; return AA_EDIMS; }   // VLIDORT_Init layer-dimension check, as the drivers
  const size_t ne = (size_t)(km + 1) * nobs, nr = (size_t)km * nobs * nch;
  std::lock_guard<std::mutex> hl(aa::host_call_mutex());
  const size_t se = (ne * sizeof(double) + 255) & ~(size_t)255;
  char* base = nullptr;
  if (aa::host_call_block(3 * se + nr * sizeof(double), &base)) return AA_EHIP;
  double *dpe = (double*)base, *dhe = (double*)(base + se), *dte = (double*)(base + 2 * se);
  double* drot = (double*)(base + 3 * se);
  AA_CHECK(hipMemcpy(dpe, pe, sizeof(double) * ne, hipMemcpyHostToDevice));
  AA_CHECK(hipMemcpy(dhe, he, sizeof(double) * ne, hipMemcpyHostToDevice));
  AA_CHECK(hipMemcpy(dte, te, sizeof(double) * ne, hipMemcpyHostToDevice));
  int e = aa::rot_launch(km, nch, nobs, channels, dpe, dhe, dte, drot, depol, 0);
  if (e) return e;
  AA_CHECK(hipMemcpy(rot, drot, sizeof(double) * nr, hipMemcpyDeviceToHost));
  return 0;
}

int aa_rot_calc_dev(int km, int nch, int nobs, const double* channels, const double* d_pe, const double* d_he,
                    const double* d_te, double* d_rot, double* depol, void* stream) {
  if (km <= 0 || km > aa::MAX_LAYERS || nch <= 0 || nobs < 0 || !channels || !d_pe || !d_he || !d_te || !d_rot ||
      !depol)
    return AA_EARG;
  return aa::rot_launch(km, nch, nobs, channels, d_pe, d_he, d_te, d_rot, depol, (hipStream_t)stream);
}

int aa_edge_vars_dev(int km, int nobs, const float* d_airdens, const float* d_delp, float ptop, float* d_pe,
                     float* d_ze, float* d_te, void* stream) {
  if (km <= 0 || nobs < 0 || !d_airdens || !d_delp || !d_pe || !d_ze || !d_te) return AA_EARG;
  if (nobs == 0) return 0;
  hipLaunchKernelGGL(aa::k_edge, dim3((nobs + 255) / 256), dim3(256), 0, (hipStream_t)stream, km, nobs, d_airdens,
                     d_delp, ptop, d_pe, d_ze, d_te);
  AA_CHECK(hipGetLastError());
  return 0;
}

int aa_edge_vars(int km, int nobs, const float* airdens, const float* delp, float ptop, float* pe, float* ze,
                 float* te) {
  if (km <= 0 || nobs < 0 || !airdens || !delp || !pe || !ze || !te) return AA_EARG;
  const size_t nl = (size_t)km * nobs, ne = (size_t)(km + 1) * nobs;
  std::lock_guard<std::mutex> hl(aa::host_call_mutex());
  float* d[5];
  if (!aa::host_floats({nl, nl, ne, ne, ne}, d)) return AA_EHIP;
  AA_CHECK(hipMemcpy(d[0], airdens, sizeof(float) * nl, hipMemcpyHostToDevice));
  AA_CHECK(hipMemcpy(d[1], delp, sizeof(float) * nl, hipMemcpyHostToDevice));
  int e = aa_edge_vars_dev(km, nobs, d[0], d[1], ptop, d[2], d[3], d[4], nullptr);
  if (e) return e;
  AA_CHECK(hipMemcpy(pe, d[2], sizeof(float) * ne, hipMemcpyDeviceToHost));
  AA_CHECK(hipMemcpy(ze, d[3], sizeof(float) * ne, hipMemcpyDeviceToHost));
  AA_CHECK(hipMemcpy(te, d[4], sizeof(float) * ne, hipMemcpyDeviceToHost));
  return 0;
}

int aa_calc_qm_dev(int km, int nq, int nobs, const float* d_q, const float* d_delp, float* d_qm, void* stream) {
  if (km <= 0 || nq <= 0 || nobs < 0 || !d_q || !d_delp || !d_qm) return AA_EARG;
  const long ne = (long)km * nq * nobs;
  if (ne == 0) return 0;
  hipLaunchKernelGGL(aa::k_qm, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, (hipStream_t)stream, km, nq, nobs,
                     d_q, d_delp, d_qm);
  AA_CHECK(hipGetLastError());
  return 0;
}

int aa_calc_qm(int km, int nq, int nobs, const float* q, const float* delp, float* qm) {
  if (km <= 0 || nq <= 0 || nobs < 0 || !q || !delp || !qm) return AA_EARG;
  const size_t nqe = (size_t)km * nq * nobs, nl = (size_t)km * nobs;
  std::lock_guard<std::mutex> hl(aa::host_call_mutex());
  float* d[3];
  if (!aa::host_floats({nqe, nl, nqe}, d)) return AA_EHIP;
  AA_CHECK(hipMemcpy(d[0], q, sizeof(float) * nqe, hipMemcpyHostToDevice));
  AA_CHECK(hipMemcpy(d[1], delp, sizeof(float) * nl, hipMemcpyHostToDevice));
  int e = aa_calc_qm_dev(km, nq, nobs, d[0], d[1], d[2], nullptr);
  if (e) return e;
  AA_CHECK(hipMemcpy(qm, d[2], sizeof(float) * nqe, hipMemcpyDeviceToHost));
  return 0;
}

}  // extern "C"
