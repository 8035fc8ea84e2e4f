// Column setup of the discrete-ordinate radiance path: layer optics and Greek
// assembly, delta-M scaling, pseudo-spherical beam geometry (Chapman factors),
// transmittances, beam multipliers and the first-order (single scatter +
// direct bounce) TOA field with outgoing sphericity.
//
// Reference (V283 = src/Components/rtms/RTSI/VLIDORT2p8p3, REG =
// V283/vlidort_main/regular, FOC = V283/vlidort_focode):
//   Greek assembly     V283/VLIDORT_ScatMod.F90:440-551 (vector), 1140-1202 (scalar)
//   input checks       REG/vlidort_inputs.f90:2439-3825
//   Chapman factors    REG/vlidort_geometry.f90:119-318
//   DELTAMSCALE        REG/vlidort_miscsetups.f90:370-758
//   QSPREP / PREPTRANS REG/vlidort_miscsetups.f90:762-1292
//   EMULT_MASTER       REG/vlidort_miscsetups.f90:3187-3421
//   first order        REG/vlidort_vfo_interface.f90:128-728, FOC/FO_WPgeometry_Routines.f90,
//                      FOC/FO_geometry_Generic.f90:100-330, FOC/FO_VectorSS_Spherfuncs.f90:93-300,
//                      FOC/FO_VectorSS_RTCalcs_I.f90:116-785
//
// Thread mapping: one thread per (item, layer) for the layer-local work, one
// thread per item for the vertical recursions.  All FP64.
#include "rt_work.h"
#include "rt_device.h"
#include "brdf_dev.h"
#include "prof.h"

namespace aa {

namespace {

constexpr double FO_CUTOFF = 88.0;   // FOC/FO_VectorSS_RTCalcs_I.f90 exponential cutoff
constexpr double SMALL_SPH = 1.0e-12;

__device__ __forceinline__ bool is_missing(double x, double missing) { return fabs(x / missing - 1.0) < 0.001; }

// Heights (km) of the column edges: the driver's ze*1.E-3 with a default-real
// literal (V283/VLIDORT_ScatMod.F90:437-438, 1128-1129).
struct ColumnGeo {
  const double* he;   // column's edges (metres)
  int L;
  __device__ double h(int k) const { return he[k] * (double)1.e-3f; }
  __device__ double radius(int k) const { return k <= L ? EARTH_RADIUS + h(k) : 0.0; }
};

// Scattering-angle cosine of the observation geometry
// (FOC/FO_WPgeometry_Routines.f90 ObsGeom_*; identical for the three FO modes
// when sza < 90).
__device__ inline double cos_scat(double sza, double vza, double raa) {
  const double dtr = DEG_TO_RAD;
  double calpha, salpha;
  if (vza == 90.0) { calpha = 0.0; salpha = 1.0; }
  else { calpha = cos(vza * dtr); salpha = sin(vza * dtr); }
  double stheta = sin(sza * dtr), ctheta = cos(sza * dtr);
  double cphi = cos(raa * dtr);
  if (sza == 0.0) {
    double term2 = calpha;
    double cs = -term2;
    if (term2 == 0.0) cs = term2;
    return cs;
  }
  double term1 = salpha * stheta * cphi, term2 = calpha * ctheta;
  return -term2 + term1;
}

// FOC/FO_geometry_Generic.f90:105-160
__device__ inline void find_sun(bool donadir, bool overhead, double radius, const double* sdir, double cumangle,
                                double theta_boa_r, double& theta, double& stheta, double& ctheta, bool& dirsun) {
  if (donadir) { dirsun = true; theta = theta_boa_r; ctheta = cos(theta_boa_r); stheta = sin(theta_boa_r); return; }
  if (overhead) { dirsun = true; ctheta = cos(cumangle); stheta = sin(cumangle); theta = cumangle; return; }
  double px0 = -radius * sin(cumangle), px2 = radius * cos(cumangle);
  double b = px0 * sdir[0] + 0.0 * sdir[1] + px2 * sdir[2];
  ctheta = -b / radius;
  dirsun = (ctheta >= 0.0);
  stheta = sqrt(1.0 - ctheta * ctheta);
  theta = acos(ctheta);
}

// For each channel c in the mask cm: s[c] = sum_k ext(k, c) * sunpath(k) for a
// direct sun (FOC/FO_geometry_Generic.f90:164-213).  The path lengths depend on
// the column only and are computed once for all channels; each channel's sum
// keeps the reference's term order.
#define AA_FOR_CH(c) _Pragma("unroll") for (int c = 0; c < CH; ++c) if ((cm >> c) & 1u)
template <int CH, class Ext>
__device__ inline void sunpaths_d_sum(bool zerosunboa, double radstart, const ColumnGeo& G, double thstart,
                                      double sthstart, int n, const Ext& ext, unsigned cm, double (&s)[CH]) {
  const int n1 = n - 1;
  if (zerosunboa) {
    const double p = G.radius(n1) - radstart;
    AA_FOR_CH(c) s[c] = ext(n, c) * p;
    for (int k = n1; k >= 1; --k) {
      const double q = G.radius(k - 1) - G.radius(k);
      AA_FOR_CH(c) s[c] += ext(k, c) * q;
    }
    return;
  }
  // First segment (radstart up to the top of layer n) exactly as the
  // reference: sin(th0 - th1) r0 / sin(th1) with th1 = asin(sin(th0) r0/r1).
  // th0 and sin(th0) come from acos(c) and sqrt(1 - c*c) of one cosine c, and
  // near an overhead sun their rounding differs by ~1e-8 relative, which
  // th0 - th1 ~ th0 (r1 - r0)/r1 magnifies to ~1e-3 of this segment; the
  // radiance follows the reference only if this segment keeps its arithmetic.
  // Above it the ray is parametrised consistently and the reference's layer
  // path equals c1 - c0 with c = sqrt(r^2 - b^2), b = r sin(th) the ray's
  // impact parameter (law of sines in the triangle of the Earth's centre and
  // the ray's crossings of radii r0 < r1); written (r1 - r0)(r1 + r0)/(c1 + c0)
  // it needs one sqrt and one division per layer and has no cancellation.
  const double b = sthstart * radstart;
  double r0, c0;
  {
    const double r1 = G.radius(n1);
    const double sth1 = sthstart * radstart / r1;
    const double th1 = asin(sth1);
    const double p = sin(thstart - th1) * radstart / sth1;
    AA_FOR_CH(c) s[c] = ext(n, c) * p;
    r0 = r1; c0 = sqrt((r1 - b) * (r1 + b));
  }
  for (int k = n1; k >= 1; --k) {
    const double r1 = G.radius(k - 1), c1 = sqrt((r1 - b) * (r1 + b));
    const double p = (r1 - r0) * (r1 + r0) / (c1 + c0);
    AA_FOR_CH(c) s[c] += ext(k, c) * p;
    r0 = r1; c0 = c1;
  }
}

// Same for a sun behind the tangent point (FOC/FO_geometry_Generic.f90:217-290);
// the path sum covers levels 1..nt.
template <int CH, class Ext>
__device__ inline void sunpaths_t_sum(double radstart, const ColumnGeo& G, double thstart, double sthstart, int n,
                                      int L, const Ext& ext, unsigned cm, double (&s)[CH]) {
  const int n1 = n - 1;
  const double tanr = sthstart * radstart;
  int k = n1; bool trawl = true;
  while (k >= n1 && trawl && k <= L + 1) { trawl = (G.radius(k) > tanr); k = k + 1; }
  const int nt = k - 1;
  AA_FOR_CH(c) s[c] = 0.0;
  double spn = 0.0;
  double sth0, th0, sth1, th1, ks1;
  if (nt > n) {
    th0 = PIE - thstart; sth0 = sthstart;
    sth1 = sth0 * radstart / G.radius(n);
    th1 = asin(sth1); ks1 = th0 - th1;
    spn = 2.0 * sin(ks1) * radstart / sth1;
    sth0 = sth1; th0 = th1;
    for (int kk = n + 1; kk <= nt - 1; ++kk) {
      sth1 = sth0 * G.radius(kk - 1) / G.radius(kk);
      th1 = asin(sth1); ks1 = th0 - th1;
      const double p = 2.0 * sin(ks1) * G.radius(kk) / sth0;
      AA_FOR_CH(c) s[c] += ext(kk, c) * p;
      sth0 = sth1; th0 = th1;
    }
    ks1 = 0.5 * PIE - th0;
    const double p = 2.0 * sin(ks1) * G.radius(nt - 1);
    AA_FOR_CH(c) s[c] += ext(nt, c) * p;
  } else if (nt == n) {
    spn = -2.0 * radstart * cos(thstart);
  }
  th0 = PIE - thstart; sth0 = sthstart;
  sth1 = sth0 * radstart / G.radius(n1);
  th1 = asin(sth1); ks1 = th0 - th1;
  spn = spn + sin(ks1) * radstart / sth1;
  sth0 = sth1; th0 = th1;
  if (n <= nt) { AA_FOR_CH(c) s[c] += ext(n, c) * spn; }
  for (int kk = n1; kk >= 1; --kk) {
    sth1 = sth0 * G.radius(kk) / G.radius(kk - 1);
    th1 = asin(sth1); ks1 = th0 - th1;
    if (kk <= nt) {
      const double p = sin(ks1) * G.radius(kk) / sth1;
      AA_FOR_CH(c) s[c] += ext(kk, c) * p;
    }
    sth0 = sth1; th0 = th1;
  }
}

struct ItemIdx {
  int g, obs, ch;
};
__device__ __forceinline__ ItemIdx item_of(const RTWork& W, int it) {
  ItemIdx r; r.g = global_item(W, it); r.obs = r.g / W.nch; r.ch = r.g - r.obs * W.nch; return r;
}

// ---------------------------------------------------------------- kernels

// First item that switches the FO correction to nadir mode: the reference
// driver latches DO_FOCORR_NADIR once a column with SZA == 0 is processed
// (V283/VLIDORT_ScatMod.F90:432-435, 1123-1126); items are processed in
// (obs, channel) order.
// The MODIS driver (src/Components/rtms/geo_vlidort/VLIDORT_BRDF_MODIS.F90:
// 160-170) runs VLIDORT only where the direct-bounce BRDF is not negative, so
// only such items latch.
__global__ void k_latch(int nitems, int nch, int nobs, const double* sza, const double* vza, const double* raa,
                        const double* albedo, const double* kernel_wt, const double* kparam, const double* u10m,
                        const double* v10m, const double* mr, const double* bpdf, double missing, int* latch) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nitems) return;
  int obs = g / nch, ch = g - obs * nch;
  if (sza[obs] != 0.0) return;
  if (is_missing(sza[obs], missing) || is_missing(vza[obs], missing) || is_missing(raa[obs], missing)) return;
  if (albedo && is_missing(albedo[(size_t)ch * nobs + obs], missing)) return;
  // the Cox-Munk drivers skip elements with missing wind or refractive index
  // (src/Components/rtms/geo_vlidort/VLIDORT_BRDF_CX.F90:506-512)
  if (u10m && (is_missing(u10m[obs], missing) || is_missing(v10m[obs], missing) || is_missing(mr[ch], missing)))
    return;
  if (bpdf) {
    // polarised land surfaces: the (1,1) direct bounce of the Lambertian (or
    // MODIS) kernels plus BPDF-NDVI
    const size_t kq = (size_t)obs * nch + ch;
    const double* bp = bpdf + kq * 3;
    const double mu0 = cos(sza[obs] * DEG_TO_RAD), s0 = sqrt(1.0 - mu0 * mu0);
    const double muv = cos(vza[obs] * DEG_TO_RAD), sv = sqrt(1.0 - muv * muv);
    const double phi = raa[obs] * DEG_TO_RAD;
    double kb[9];
    bpdf_ndvi(bp[0], bp[1], bp[2], false, mu0, s0, muv, sv, cos(phi), sin(phi), kb);
    double db;
    if (kernel_wt) {
      const double* kw = kernel_wt + kq * 3;
      const double* kp = kparam + kq * 2;
      db = modis_db(kw[0], kw[1], kw[2], kp[0], kp[1], sza[obs], vza[obs], raa[obs]);
    } else {
      db = albedo[(size_t)ch * nobs + obs] * 1.0;
    }
    if (db + 1.0 * kb[0] < 0.0) return;
  } else if (kernel_wt) {
    const double* kw = kernel_wt + ((size_t)obs * nch + ch) * 3;
    const double* kp = kparam + ((size_t)obs * nch + ch) * 2;
    if (modis_db(kw[0], kw[1], kw[2], kp[0], kp[1], sza[obs], vza[obs], raa[obs]) < 0.0) return;
  }
  atomicMin(latch, g);
}

// Input checks (REG/vlidort_inputs.f90 VLIDORT_CHECK_INPUT, the ones the driver
// inputs can trip) and MISSING propagation of the driver
// (src/Components/rtms/geo_vlidort/VLIDORT_LAMBERT.F90:95-110).
__global__ void k_check(RTWork W, int nitems) {
  int it = blockIdx.x * blockDim.x + threadIdx.x;
  if (it >= nitems) return;
  ItemIdx I = item_of(W, it);
  const double sza = W.sza[I.obs], vza = W.vza[I.obs], raa = W.raa[I.obs];
  int st = 0;
  if (is_missing(sza, W.missing) || is_missing(vza, W.missing) || is_missing(raa, W.missing)) st = 1;
  else if (W.albedo && is_missing(W.albedo[(size_t)I.ch * W.nobs + I.obs], W.missing)) st = 1;
  else if (W.u10m && (is_missing(W.u10m[I.obs], W.missing) || is_missing(W.v10m[I.obs], W.missing) ||
                      is_missing(W.mr[I.ch], W.missing)))
    st = 1;
  int L = W.nlayers ? W.nlayers[I.obs] : W.km;
  if (L < 1 || L > W.km) { if (!st) st = 4; L = 1; }
  if (!st) {
    if (sza < 0.0 || sza >= 90.0) st = 3;
    if (raa > 360.0 || raa < 0.0) st = 3;
    if (vza > 90.0 || vza < 0.0) st = 3;
    ColumnGeo G{W.he + (size_t)I.obs * (W.km + 1), L};
    for (int i = 1; i <= L; ++i) if (G.h(i - 1) <= G.h(i)) st = 3;
  }
  W.status[it] = st;
  W.L_it[it] = L;
  W.csza[it] = cos(sza * DEG_TO_RAD);
  W.usec[it] = 1.0 / cos(DEG_TO_RAD * vza);
}

// Greek matrix of one layer from the driver optics (V283/VLIDORT_ScatMod.F90:
// 1185-1198 scalar, 456-530 vector).  Returns G(l, k) for Greek index k.
template <int S, typename TO>
struct LayerGreek {
  const TO* pm;          // pmom at (layer, ch, obs): moment l, pol p at pm[(p*nmom + l) * pstride]
  size_t pstride;
  int nmom;
  double aerswt, rayswt, raysmom2, gammamom2, alphamom2, deltamom1;
  __device__ double P(int l, int p) const { return (l <= nmom - 1) ? (double)pm[((size_t)p * nmom + l) * pstride] : 0.0; }
  __device__ double operator()(int l, int k) const {
    if (S == 1) {
      if (k != 0) return 0.0;
      if (l == 0) return 1.0;
      if (l == 1) return P(1, 0) * aerswt;
      if (l == 2) return raysmom2 * rayswt + P(2, 0) * aerswt;
      return P(l, 0) * aerswt;
    } else {
      if (l == 0 && k == 0) return 1.0;
      double aer;
      switch (k) {
        case 0: aer = P(l, 0); break;
        case 1: case 4: aer = P(l, 1); break;
        case 10: aer = P(l, 2); break;
        case 11: aer = P(l, 3); break;
        case 14: aer = -P(l, 3); break;
        case 5: aer = P(l, 4); break;
        case 15: aer = P(l, 5); break;
        default: aer = 0.0;
      }
      if (l > 2) return aer * aerswt;
      double ray = 0.0;
      if (l == 0 && k == 0) ray = 1.0;
      if (l == 2 && k == 0) ray = raysmom2;
      if (l == 2 && (k == 1 || k == 4)) ray = gammamom2;
      if (l == 2 && k == 5) ray = alphamom2;
      if (l == 1 && k == 15) ray = deltamom1;
      return ray * rayswt + aer * aerswt;
    }
  }
};

// One thread per (item, layer): optics, delta-M, omega*Greek, scattering flags
// and the first-order phase function x TMS factor.
template <int S, typename TO>
__global__ void k_optics(RTWork W, int nitems) {
  constexpr int LMX = 2 * MAX_STREAMS;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int it = t / W.km, n = t - it * W.km;
  if (it >= nitems) return;
  if (W.status[it] != 0 || n >= W.L_it[it]) return;
  ItemIdx I = item_of(W, it);
  const int km = W.km, nch = W.nch, nobs = W.nobs, nmom = W.nmom;
  const TO* tau = (const TO*)W.tau;
  const TO* ssa = (const TO*)W.ssa;
  const TO* pmom = (const TO*)W.pmom;
  const size_t ti = ((size_t)(I.obs - W.opt_obs0) * nch + I.ch) * km + n;
  const size_t ri = ((size_t)I.ch * nobs + I.obs) * km + n;
  const double ray_l = W.rot[ri];
  const double alpha_l = W.alpha ? W.alpha[ri] : 0.0;
  const double tau_l = (double)tau[ti], ssa_l = (double)ssa[ti];
  const double tau_ext = alpha_l + ray_l + tau_l;
  const double tau_scat = ray_l + ssa_l * tau_l;
  double omega = tau_scat / tau_ext;
  if (omega > (double)0.99999f) omega = (double)0.99999f;
  if (omega > 1.0 - OMEGA_SMALLNUM || omega < OMEGA_SMALLNUM) W.status[it] = 3;
  const double d = W.depol[I.ch];
  LayerGreek<S, TO> G;
  G.pm = pmom + ti;
  G.pstride = (size_t)km * nch * W.opt_nobs;
  G.nmom = nmom;
  G.aerswt = ssa_l * tau_l / tau_scat;
  G.rayswt = ray_l / tau_scat;
  G.raysmom2 = (1.0 - d) / (2.0 + d);
  G.gammamom2 = -((double)sqrtf(6.0f) * (1 - d) / (2 + d));
  G.alphamom2 = 6 * (1 - d) / (2 + d);
  G.deltamom1 = 3 * (1 - 2 * d) / (2 + d);

  // DELTAMSCALE (REG/vlidort_miscsetups.f90:471-500)
  const int NMOM = W.nmoments;
  const int nm1 = NMOM + 1;
  const double dnm1 = (double)(2 * nm1 + 1);
  const double fdel = G(nm1, 0) / dnm1;
  const double fac2 = 1.0 - fdel;
  const double fac1 = 1.0 - fdel * omega;
  const double dtau = tau_ext * fac1;
  double omega_s = omega * fac2 / fac1;
  if (omega_s > 0.999999999) omega_s = 0.999999999;
  else if (omega_s < 1.0e-9) omega_s = 1.0e-9;
  const int LM = 2 * W.N;
  double* og = W.og + ((size_t)it * km + n) * (size_t)(LM * S * S);
  for (int l = 0; l < LM; ++l) {
    double gm[S][S];
    const double fdnl1 = fdel * (double)(2 * l + 1);
#pragma unroll
    for (int o1 = 0; o1 < S; ++o1)
#pragma unroll
      for (int o2 = 0; o2 < S; ++o2) {
        const int k = 4 * o1 + o2;
        double v = 0.0;
        if (l <= NMOM) {
          if (S == 1) v = (G(l, 0) - fdnl1) / fac2;
          else if (k == 0 || k == 5 || k == 10 || k == 15) v = (G(l, k) - fdnl1) / fac2;
          else if (k == 1 || k == 4 || k == 11 || k == 14) v = G(l, k) / fac2;
        }
        gm[o1][o2] = v;
      }
    if (l == 0) gm[0][0] = 1.0;
    if (l == 1) {
      double g1 = gm[0][0] / 3.0;
      if (g1 > 0.999999999) g1 = 0.999999999;
      else if (g1 < -0.999999999) g1 = -0.999999999;
      else if (g1 >= 0.0 && g1 < 1.0e-9) g1 = 1.0e-9;
      else if (g1 < 0.0 && g1 > -1.0e-9) g1 = -1.0e-9;
      gm[0][0] = 3.0 * g1;
    }
#pragma unroll
    for (int o1 = 0; o1 < S; ++o1)
#pragma unroll
      for (int o2 = 0; o2 < S; ++o2) og[(l * S + o1) * S + o2] = (l <= NMOM) ? omega_s * gm[o1][o2] : 0.0;
  }
  (void)LMX;
  // layer scattering flags (REG/vlidort_inputs.f90:4136-4160): bit m set when
  // some unscaled moment l in [m, NMOM] is nonzero; always set for m < 3
  unsigned mask = 7u;
  {
    int lastnz = -1;
    for (int l = 3; l <= NMOM; ++l) if (G(l, 0) != 0.0) lastnz = l;
    for (int m = 3; m <= NMOM; ++m) if (m <= lastnz) mask |= 1u << m;
  }
  W.lsc[(size_t)it * km + n] = mask;
  W.deltau[(size_t)it * km + n] = tau_ext;
  W.fac1[(size_t)it * km + n] = fac1;
  W.tr[((size_t)it * km + n) * TR_NFIELDS + TR_DTAU] = dtau;
  {
    // layer extinction per km for the first-order sun paths (n < L here)
    const ColumnGeo Gc{W.he + (size_t)I.obs * (km + 1), W.L_it[it]};
    W.ext[(size_t)it * km + n] = dtau / (Gc.h(n) - Gc.h(n + 1));
  }

  // first-order phase function at the scattering angle, all nmom moments
  // (FOC/FO_VectorSS_Spherfuncs.f90:93-300, FOC/FO_VectorSS_RTCalcs_I.f90)
  const double sza = W.sza[I.obs], vza = W.vza[I.obs];
  double raa = W.raa[I.obs];
  if (raa < 0.0) raa = -raa;
  if (raa > 360.0) raa = 360.0 - raa - 360.0;
  const double mus = cos_scat(sza, vza, raa);
  const double tms = omega / (1.0 - fdel * omega);
  double* ex = W.exact + ((size_t)it * km + n) * 3;
  if (S == 1) {
    double gm2 = 1.0, gm1 = mus, f1 = gm2 * G(0, 0);
    if (nmom >= 1) f1 = f1 + gm1 * G(1, 0);
    for (int l = 2; l <= nmom; ++l) {
      double gh1 = (double)(2 * l - 1) / (double)l, gh2 = (double)(l - 1) / (double)l;
      double gl = gh1 * gm1 * mus - gh2 * gm2;
      f1 = f1 + gl * G(l, 0);
      gm2 = gm1; gm1 = gl;
    }
    ex[0] = f1 * tms;
  } else {
    const double dtr = DEG_TO_RAD;
    const double ctheta = cos(dtr * sza), stheta = sin(dtr * sza);
    const double calpha = cos(dtr * vza), salpha = sin(dtr * vza);
    const double cphi = cos(dtr * raa);
    const double help = 1.0 - mus * mus;
    const double sinscat = (help <= 0.0) ? SMALL_SPH : sqrt(help);
    double csig1, csig2;
    if (fabs(sinscat) <= SMALL_SPH) { csig1 = 0.0; csig2 = 0.0; }
    else {
      csig1 = (stheta == 0.0) ? -cphi : (1.0 * calpha + ctheta * mus) / sinscat / stheta;
      csig2 = (salpha == 0.0) ? -cphi : (-ctheta - 1.0 * calpha * mus) / sinscat / salpha;
    }
    csig1 = fmin(1.0, fmax(-1.0, csig1));
    csig2 = fmin(1.0, fmax(-1.0, csig2));
    (void)csig1;
    const double csig2_2 = 2.0 * csig2;
    const double ssig2 = (fabs(csig2 - 1.0) < SMALL_SPH) ? 0.0 : sqrt(1.0 - csig2 * csig2);
    const double rot3 = csig2_2 * csig2 - 1.0;
    const double rot4 = (raa <= 180.0) ? csig2_2 * ssig2 : -csig2_2 * ssig2;
    const double qroot6 = -0.25 * sqrt(6.0);
    double p0m2 = 0.0, p0m1 = 0.0, f1 = 0.0, f2 = 0.0;   // P00 recurrence
    double p2m2 = 0.0, p2m1 = 0.0, sql41 = 0.0;          // P02 recurrence
    for (int l = 0; l <= nmom; ++l) {
      double p00;
      if (l == 0) p00 = 1.0;
      else if (l == 1) p00 = mus;
      else {
        double gh1 = (double)(2 * l - 1) / (double)l, gh2 = (double)(l - 1) / (double)l;
        p00 = gh1 * mus * p0m1 - gh2 * p0m2;
      }
      p0m2 = p0m1; p0m1 = p00;
      f1 = f1 + p00 * G(l, 0);
      double p02 = 0.0;
      if (l == 2) {
        double mum1 = 1.0 - mus, mup1 = 1.0 + mus;
        p02 = qroot6 * mum1 * mup1;
        p2m2 = 0.0; p2m1 = p02;
        sql41 = 0.0;
      } else if (l > 2) {
        double dl = l, sql4 = sql41;
        sql41 = sqrt(dl * dl - 4.0);
        double gh3 = (double)(2 * l - 1) / sql41, gh4 = sql4 / sql41;
        p02 = gh3 * mus * p2m1 - gh4 * p2m2;
        p2m2 = p2m1; p2m1 = p02;
      }
      f2 = f2 + p02 * G(l, 1);
    }
    ex[0] = tms * (+f1);
    ex[1] = tms * (-f2 * rot3);
    ex[2] = tms * (+f2 * rot4);
  }
}

// Chapman slant optical depth to level n and the first-order sun paths
// (outgoing-sphericity quadrature) of the items ita + c, c in the mask cm: the
// channels of one column (same geometry and L), so the slant-path lengths are
// computed once and summed with each channel's extinction.
template <int CH>
__device__ void geom_level(const RTWork& W, int ita, unsigned cm, int n) {
  const int itf = ita + (__ffs(cm) - 1);
  const ItemIdx I = item_of(W, itf);
  const int km = W.km;
  const int L = W.L_it[itf];
  ColumnGeo G{W.he + (size_t)I.obs * (km + 1), L};
  const double *deltau[CH], *fac1[CH], *extc[CH];
  double* lev[CH];
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const size_t it = (size_t)(ita + c);
    deltau[c] = W.deltau + it * km;
    fac1[c] = W.fac1 + it * km;
    extc[c] = W.ext + it * km;
    lev[c] = W.lev + (it * (km + 1) + n) * 4;
  }
  const double sza = W.sza[I.obs], vza = W.vza[I.obs];
  double raa = W.raa[I.obs];
  if (raa < 0.0) raa = -raa;
  if (raa > 360.0) raa = 360.0 - raa - 360.0;
  const double dtr = DEG_TO_RAD;

  // ---- Chapman factors (REG/vlidort_geometry.f90:119-318, non-refractive)
  {
    const double th_toa = sza * DEG_TO_RAD, mu_toa = cos(th_toa), sn_toa = sqrt(1.0 - mu_toa * mu_toa);
    double tau[CH];
    AA_FOR_CH(c) tau[c] = 0.0;
    if (W.plane_parallel) {
      for (int k = 1; k <= n; ++k) { AA_FOR_CH(c) tau[c] += deltau[c][k - 1] * (1.0 / mu_toa) * fac1[c][k - 1]; }
    } else if (sn_toa == 0.0) {
      for (int k = 1; k <= n; ++k) { AA_FOR_CH(c) tau[c] += deltau[c][k - 1] * 1.0 * fac1[c][k - 1]; }
    } else {
      const double r0 = G.h(0) + EARTH_RADIUS;
      double sinth1 = sn_toa * (G.h(n) + EARTH_RADIUS) / r0;
      double sth1 = asin(sinth1);
      double re_upper = r0;
      for (int k = 1; k <= n; ++k) {
        double delz = G.h(k - 1) - G.h(k);
        double re_lower = re_upper - delz;
        double sinth2 = re_upper * sinth1 / re_lower;
        double sth2 = asin(sinth2);
        double phi = sth2 - sth1;
        double sinphi = sin(phi);
        double dist = re_upper * sinphi / sinth2;
        double chap = dist / delz;
        AA_FOR_CH(c) tau[c] += deltau[c][k - 1] * chap * fac1[c][k - 1];
        re_upper = re_lower; sinth1 = sinth2; sth1 = sth2;
      }
    }
    AA_FOR_CH(c) lev[c][0] = tau[c];
  }

  // ---- first-order geometry and sun-path optical depths
  // layer extinction per km (k_optics: scaled optical thickness over the
  // layer depth); levels below the surface (a ray tangent under the ground)
  // contribute nothing
  auto ext = [&](int k, int c) { return k <= L ? extc[c][k - 1] : 0.0; };
  const bool planpar = W.plane_parallel != 0;
  // enhanced (outgoing-sphericity) mode unless plane-parallel or past the nadir latch
  unsigned em = 0u;
  if (!planpar) {
    const int latch = *W.latch;
    AA_FOR_CH(c) if (global_item(W, ita + c) < latch) em |= 1u << c;
  }
  double suntau[CH], multenh[CH], lostrans[CH];
#pragma unroll
  for (int c = 0; c < CH; ++c) { suntau[c] = 0.0; multenh[c] = 0.0; lostrans[c] = 0.0; }
  if (planpar) {
    // ObsGeom_PlanPar_WP (FOC/FO_WPgeometry_Routines.f90:126-311)
    const double utheta = 1.0 / cos(sza * dtr);
    for (int k = 1; k <= n; ++k) {
      const double q = (G.h(k - 1) - G.h(k)) * utheta;
      AA_FOR_CH(c) suntau[c] += ext(k, c) * q;
    }
  }
  if (!planpar && (cm & ~em)) {
    // Obsgeom_RegularPS_WP (FOC/FO_WPgeometry_Routines.f90:776-991)
    const double theta_r = sza * dtr;
    double stheta = sin(theta_r);
    const bool overhead = (sza == 0.0);
    sunpaths_d_sum<CH>(overhead, G.radius(n), G, theta_r, stheta, n, ext, cm & ~em, suntau);
  }
  if (em) {
    // LosOut_EnhancedPS_Initial/Quadrature + SolarIn_EnhancedPS_ObsGeom_SunPaths
    // (FOC/FO_WPgeometry_Routines.f90:1511-1971, 2409-2859)
    const unsigned cm = em;
    const bool donadir = (vza == 0.0);
    double salpha_boa, calpha1;
    if (vza == 90.0) { salpha_boa = 1.0; calpha1 = 0.0; }
    else { salpha_boa = sin(vza * dtr); calpha1 = cos(vza * dtr); }
    const double raycon = salpha_boa * G.radius(L);
    auto alpha_at = [&](int k) -> double {   // alpha(k), 0 in nadir mode
      if (donadir) return 0.0;
      if (k == L) return vza * dtr;
      return asin(raycon / G.radius(k));
    };
    auto cosa_at = [&](int k) -> double {
      if (donadir) return 0.0;
      if (k == L) return calpha1;
      return cos(asin(raycon / G.radius(k)));
    };
    double losw;
    if (donadir) losw = G.radius(n - 1) - G.radius(n);
    else losw = G.radius(n - 1) * cosa_at(n - 1) - G.radius(n) * cosa_at(n);
    double tfine[NFINE], afine[NFINE], radiif[NFINE], alphaf[NFINE];
    if (donadir) {
      getquad2(0.0, G.radius(n - 1) - G.radius(n), NFINE, tfine, afine);
      for (int j = 0; j < NFINE; ++j) { radiif[j] = G.radius(n) + tfine[j]; alphaf[j] = 0.0; }
    } else {
      const double rcn = G.radius(n) * cosa_at(n);
      getquad2(0.0, losw, NFINE, tfine, afine);
      for (int j = 0; j < NFINE; ++j) {
        double tanfine = raycon / (rcn + tfine[j]);
        double af = atan(tanfine);
        double sf = sin(af);
        alphaf[j] = af;
        radiif[j] = raycon / sf;
      }
    }
    const bool overhead = (sza == 0.0);
    const double alphaL = alpha_at(L);
    const double theta_boa_r = sza * dtr;
    double ctheta_boa, stheta_boa;
    if (sza == 90.0) { ctheta_boa = 0.0; stheta_boa = 1.0; }
    else { stheta_boa = sin(theta_boa_r); ctheta_boa = cos(theta_boa_r); }
    const double phi_boa_r = raa * dtr, cphi_boa = cos(phi_boa_r), sphi_boa = sin(phi_boa_r);
    double sdir[3];
    if (overhead) { sdir[0] = sdir[1] = sdir[2] = 0.0; }
    else { sdir[0] = -stheta_boa * cphi_boa * 1.0; sdir[1] = -stheta_boa * sphi_boa; sdir[2] = -ctheta_boa; }
    const bool zerosunboa = overhead && (n == L || donadir);
    const double radstart = G.radius(n);
    double theta_n, stheta, ctheta; bool directsun;
    if (n == L) { theta_n = theta_boa_r; stheta = stheta_boa; ctheta = ctheta_boa; directsun = true; }
    else find_sun(donadir, overhead, G.radius(n), sdir, alphaL - alpha_at(n), theta_boa_r, theta_n, stheta, ctheta, directsun);
    if (directsun) sunpaths_d_sum<CH>(zerosunboa, radstart, G, theta_n, stheta, n, ext, cm, suntau);
    else sunpaths_t_sum<CH>(radstart, G, theta_n, stheta, n, L, ext, cm, suntau);
    double kn[CH], sum[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) { kn[c] = 0.0; sum[c] = 0.0; }
    AA_FOR_CH(c) kn[c] = ext(n, c);
    for (int j = 0; j < NFINE; ++j) {
      double thf, sthf, cthf; bool dsf;
      find_sun(donadir, overhead, radiif[j], sdir, alphaL - alphaf[j], theta_boa_r, thf, sthf, cthf, dsf);
      double sf[CH];
      if (dsf) sunpaths_d_sum<CH>(zerosunboa, radiif[j], G, thf, sthf, n, ext, cm, sf);
      else sunpaths_t_sum<CH>(radiif[j], G, thf, sthf, n, L, ext, cm, sf);
      const double dj = losw - tfine[j];
      AA_FOR_CH(c) {
        double attenf = (sf[c] < FO_CUTOFF) ? exp(-sf[c]) : 0.0;
        double tran = exp(-kn[c] * dj);
        sum[c] = sum[c] + attenf * tran * afine[j];
      }
    }
    AA_FOR_CH(c) {
      multenh[c] = sum[c] * kn[c];
      const double lostau = kn[c] * losw;
      if (lostau < FO_CUTOFF) lostrans[c] = exp(-lostau);
    }
  }
  AA_FOR_CH(c) {
    lev[c][1] = suntau[c];
    lev[c][2] = (suntau[c] < FO_CUTOFF) ? exp(-suntau[c]) : 0.0;
    lev[c][3] = multenh[c];
    W.lostrans[(size_t)(ita + c) * (km + 1) + n] = lostrans[c];
  }
}

// One thread per (column, level n in 1..L) over the column's items in the
// chunk, GEO_CH channels at a time; with an item list (stream-count buckets)
// one thread per (item, level).  The path loops of level n run over the n
// layers above it, so a wavefront holds 8 columns x 8 adjacent levels: its
// lanes' trip counts differ by at most 7.
constexpr int GEO_LV = 8, GEO_CH = 4;
__global__ __launch_bounds__(64, 2) void k_geom(RTWork W, int nitems, int obs_first, int nruns) {
  const int nlg = (W.km + GEO_LV - 1) / GEO_LV;
  const int grp = blockIdx.x / nlg, lg = blockIdx.x - grp * nlg;
  const int r = grp * (64 / GEO_LV) + threadIdx.x / GEO_LV, n = lg * GEO_LV + threadIdx.x % GEO_LV + 1;
  if (r >= nruns) return;
  int ita, itb;
  if (W.item_list) {
    ita = r; itb = r + 1;
  } else {
    const long g0 = (long)(obs_first + r) * W.nch - W.item0;
    ita = (int)(g0 > 0 ? g0 : 0L);
    itb = (int)(g0 + W.nch < nitems ? g0 + W.nch : (long)nitems);
  }
  for (int a = ita; a < itb; a += GEO_CH) {
    unsigned cm = 0u;
#pragma unroll
    for (int c = 0; c < GEO_CH; ++c) {
      const int it = a + c;
      if (it < itb && W.status[it] == 0 && n <= W.L_it[it]) cm |= 1u << c;
    }
    if (cm) geom_level<GEO_CH>(W, a, cm, n);
  }
}

// One thread per item: beam cutoff and average secants (QSPREP), layer
// transmittances (PREPTRANS), beam multipliers (EMULT), the first-order TOA
// recursion, and the Fourier state.  Failed items get MISSING outputs.
template <int S>
__global__ void k_item(RTWork W, int nitems, int* active, int* nactive) {
  const int it = blockIdx.x * blockDim.x + threadIdx.x;
  if (it >= nitems) return;
  ItemIdx I = item_of(W, it);
  const size_t oi = (size_t)I.ch * W.nobs + I.obs;
  const int km = W.km;
  if (W.status[it] != 0) {
    // status 6: negative direct-bounce BRDF, reported as -500 by the MODIS driver
    const double fill = (W.status[it] == 6) ? -500.0 : W.missing;
    W.radiance[oi] = fill; W.reflectance[oi] = fill;
    // (the polarised drivers write -500 into Q and U as well)
    const double qfill = (W.status[it] == 6 && W.bpdf) ? -500.0 : W.missing;
    if (S > 1) { W.qout[oi] = qfill; W.uout[oi] = qfill; }
    W.nfourier[oi] = 0;
    return;
  }
  const int L = W.L_it[it];
  const double usec = W.usec[it], csza = W.csza[it];
  double* trr = W.tr + (size_t)it * km * TR_NFIELDS;
  auto TR = [&](int k, int f) -> double& { return trr[(size_t)k * TR_NFIELDS + f]; };
  const double* lev = W.lev + (size_t)it * (km + 1) * 4;
  auto tauslant = [&](int nl) { return nl == 0 ? 0.0 : lev[(size_t)nl * 4 + 0]; };
  auto suntau = [&](int nl) { return nl == 0 ? 0.0 : lev[(size_t)nl * 4 + 1]; };
  auto atten = [&](int nl) { return nl == 0 ? 1.0 : lev[(size_t)nl * 4 + 2]; };
  // ---- QSPREP (REG/vlidort_miscsetups.f90:813-1010)
  int beam_cutoff = L;
  double tau_solar;
  for (int nl = 0; nl < L; ++nl) { TR(nl, TR_AVSEC) = 0.0; TR(nl, TR_ITRANS) = 0.0; }
  if (W.plane_parallel) {
    const double sec0 = 1.0 / csza;
    double taugrid = 0.0;
    for (int nl = 1; nl <= L; ++nl) {
      double tprev = taugrid;
      taugrid = taugrid + TR(nl - 1, TR_DTAU);
      double tsl = taugrid * sec0;
      if (nl <= beam_cutoff) {
        if (tsl > MAX_TAU_PATH) beam_cutoff = nl;
        TR(nl - 1, TR_AVSEC) = sec0; TR(nl - 1, TR_ITRANS) = exp(-tprev * sec0);
      }
    }
    tau_solar = taugrid * sec0;
  } else {
    tau_solar = tauslant(L);
    double s_t_1 = 0.0, s_t_0 = 1.0;
    for (int nl = 1; nl <= L; ++nl) {
      if (nl <= beam_cutoff) {
        if (tauslant(nl) > MAX_TAU_PATH) beam_cutoff = nl;
        else s_t_1 = exp(-tauslant(nl));
        TR(nl - 1, TR_AVSEC) = (tauslant(nl) - tauslant(nl - 1)) / TR(nl - 1, TR_DTAU);
        TR(nl - 1, TR_ITRANS) = s_t_0;
        s_t_0 = s_t_1;
      }
    }
  }
  W.tsb[it] = (tau_solar < MAX_TAU_PATH) ? exp(-tau_solar) : 0.0;
  W.beam_cutoff[it] = beam_cutoff;
  // ---- PREPTRANS + EMULT (REG/vlidort_miscsetups.f90:1014-1292, 3187-3421)
  for (int nl = 1; nl <= L; ++nl) {
    const double dt = TR(nl - 1, TR_DTAU);
    double tdm;
    if (nl > beam_cutoff) tdm = 0.0;
    else { double sp = dt * TR(nl - 1, TR_AVSEC); tdm = (sp > MAX_TAU_PATH) ? 0.0 : exp(-sp); }
    TR(nl - 1, TR_TDMUBAR) = tdm;
    const double itu = TR(nl - 1, TR_ITRANS) * usec;
    TR(nl - 1, TR_ITRANSU) = itu;
    double sp = dt * usec;
    const double tdu = (sp > MAX_TAU_PATH) ? 0.0 : exp(-sp);
    TR(nl - 1, TR_TDUSER) = tdu;
    const double sigp = TR(nl - 1, TR_AVSEC) + usec;
    TR(nl - 1, TR_SIGP) = sigp;
    double em;
    if (nl > beam_cutoff) em = 0.0;
    else { double wudel = tdm * tdu; double su = (1.0 - wudel) / sigp; em = itu * su; }
    TR(nl - 1, TR_EMULTUP) = em;
  }
  // ---- first-order TOA recursion (FOC/FO_VectorSS_RTCalcs_I.f90 SSV_Integral_I_UP)
  const double sza = W.sza[I.obs], vza = W.vza[I.obs];
  const double dtr = DEG_TO_RAD;
  const bool latched = global_item(W, it) >= *W.latch;
  const bool planpar = W.plane_parallel != 0;
  const bool enhanced = !planpar && !latched;
  double mu0, mu1;
  if (planpar) { mu0 = cos(sza * dtr); mu1 = (vza == 90.0) ? 0.0 : cos(vza * dtr); }
  else {
    mu0 = (sza == 90.0) ? 0.0 : cos(sza * dtr);
    mu1 = (vza == 90.0) ? 0.0 : cos(vza * dtr);
  }
  const double fvec1 = 1.0;
  double cums[3] = {0, 0, 0};
  // surface term of the direct bounce: M4*albedo*atten*fvec (Lambertian) or
  // M4*(BRDF . fvec)*atten with the exact direct-bounce BRDF
  // (FOC/FO_VectorSS_RTCalcs_I.f90:698-706)
  double cdbv[3] = {0.0, 0.0, 0.0};
  if (S == 3 && W.bpdf) {
    // polarised surface: the first column of the reflection matrix times the
    // unpolarised flux vector, for each Stokes component
    const double* rec = W.brdf + (size_t)it * W.brdf_stride;
    for (int o = 0; o < S; ++o) cdbv[o] = 4.0 * mu0 * (rec[BPOff(W.N).db(o * 3)] * fvec1) * atten(L);
  } else if (W.brdf) {
    cdbv[0] = 4.0 * mu0 * (W.brdf[(size_t)it * W.brdf_stride + BROff(W.N).db()] * fvec1) * atten(L);
  } else {
    cdbv[0] = 4.0 * mu0 * W.albedo[oi] * atten(L) * fvec1;
  }
  const double* lt = W.lostrans + (size_t)it * (km + 1);
  const double* ex = W.exact + (size_t)it * km * 3;
  for (int nl = L; nl >= 1; --nl) {
    double ltr = 0.0, mult = 0.0;
    if (!enhanced) {
      if (mu1 > 0.0) {
        double lostau = TR(nl - 1, TR_DTAU) / mu1;
        if (lostau < FO_CUTOFF) ltr = exp(-lostau);
        double factor1 = atten(nl - 1) - atten(nl) * ltr;
        double factor2 = 1.0 + (suntau(nl) - suntau(nl - 1)) / lostau;
        mult = factor1 / factor2;
      }
    } else {
      ltr = lt[nl];
      mult = lev[(size_t)nl * 4 + 3];
    }
#pragma unroll
    for (int o = 0; o < S; ++o) {
      double src = ex[(size_t)(nl - 1) * 3 + o] * fvec1 * mult;
      cdbv[o] = ltr * cdbv[o];
      cums[o] = ltr * cums[o] + src;
    }
  }
  const double ff = W.flux ? W.flux[(size_t)I.obs * W.nch + I.ch] : 1.0;
  const double flux = ff / PI4;
#pragma unroll
  for (int o = 0; o < 3; ++o) {
    W.ss[(size_t)it * 3 + o] = (o < S) ? flux * cums[o] : 0.0;
    W.db[(size_t)it * 3 + o] = (o < S) ? flux * cdbv[o] : 0.0;
    W.stokes[(size_t)it * 3 + o] = 0.0;
  }
  W.testconv[it] = 0;
  W.fsaved[it] = 0;
  W.nfour[it] = 0;
}

// The first Fourier term's item lists and (item, layer) slots: one thread per
// item, wavefront-aggregated appends.
__global__ __launch_bounds__(64) void k_enlist(RTWork W, int nitems, int* active, int* nactive) {
  const int it = blockIdx.x * 64 + threadIdx.x;
  const bool take = it < nitems && W.status[it] == 0;
  append_items(W, it, take ? W.L_it[it] : 0, take, nactive, active, W.lcnt);
}

// Owners of the first Fourier term's (item, layer) slots: one thread per
// (item, layer), so a wavefront writes consecutive slots.
__global__ void k_owner(RTWork W, int nitems) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int it = t / W.km, l = t - it * W.km;
  if (it >= nitems || W.status[it] != 0 || l >= W.L_it[it]) return;
  W.lown[W.loff[it] + l] = it;
}

}  // namespace

// Runs the setup stage for items [W.item0, W.item0 + nitems) and fills the
// first active list.
int rt_brdf(const RTWork& W, int nitems, hipStream_t st);

int rt_setup(const RTWork& W, int nitems, int* d_active, int* d_nactive, hipStream_t st) {
  const int B = 256;
  const int gi = (nitems + B - 1) / B;
  const int gl = (int)(((size_t)nitems * W.km + B - 1) / B);
  prof_begin(PROF_CHECK, st);
  hipLaunchKernelGGL(k_check, dim3(gi), dim3(B), 0, st, W, nitems);
  prof_end(PROF_CHECK, st);
  if (W.kernel_wt || W.u10m || W.bpdf) {
    prof_begin(PROF_BRDF, st);
    if (rt_brdf(W, nitems, st)) return -2;
    prof_end(PROF_BRDF, st);
  }
  prof_begin(PROF_OPTICS, st);
  if (W.S == 1) {
    if (W.optics_f32) hipLaunchKernelGGL((k_optics<1, float>), dim3(gl), dim3(B), 0, st, W, nitems);
    else hipLaunchKernelGGL((k_optics<1, double>), dim3(gl), dim3(B), 0, st, W, nitems);
  } else {
    if (W.optics_f32) hipLaunchKernelGGL((k_optics<3, float>), dim3(gl), dim3(B), 0, st, W, nitems);
    else hipLaunchKernelGGL((k_optics<3, double>), dim3(gl), dim3(B), 0, st, W, nitems);
  }
  prof_end(PROF_OPTICS, st);
  prof_begin(PROF_GEOM, st);
  {
    // runs: the chunk's columns (items are (obs, channel) in order), or single items
    int obs_first = 0, nruns = nitems;
    if (!W.item_list && nitems > 0) {
      obs_first = W.item0 / W.nch;
      nruns = (W.item0 + nitems - 1) / W.nch - obs_first + 1;
    }
    const size_t gg = (size_t)((nruns + 64 / GEO_LV - 1) / (64 / GEO_LV)) * ((W.km + GEO_LV - 1) / GEO_LV);
    if (gg > 0) hipLaunchKernelGGL(k_geom, dim3((unsigned)gg), dim3(64), 0, st, W, nitems, obs_first, nruns);
  }
  prof_end(PROF_GEOM, st);
  prof_begin(PROF_ITEM, st);
  if (W.S == 1) hipLaunchKernelGGL((k_item<1>), dim3(gi), dim3(B), 0, st, W, nitems, d_active, d_nactive);
  else hipLaunchKernelGGL((k_item<3>), dim3(gi), dim3(B), 0, st, W, nitems, d_active, d_nactive);
  hipLaunchKernelGGL(k_enlist, dim3((nitems + 63) / 64), dim3(64), 0, st, W, nitems, d_active, d_nactive);
  {
    const size_t go = ((size_t)nitems * W.km + 255) / 256;
    hipLaunchKernelGGL(k_owner, dim3((unsigned)go), dim3(256), 0, st, W, nitems);
  }
  prof_end(PROF_ITEM, st);
  AA_CHECK(hipGetLastError());
  return 0;
}

int rt_latch(int nitems, int nch, int nobs, const double* sza, const double* vza, const double* raa,
             const double* albedo, const double* kernel_wt, const double* kparam, const double* u10m,
             const double* v10m, const double* mr, const double* bpdf, double missing, int* d_latch,
             hipStream_t st) {
  const int B = 256;
  hipLaunchKernelGGL(k_latch, dim3((nitems + B - 1) / B), dim3(B), 0, st, nitems, nch, nobs, sza, vza, raa, albedo,
                     kernel_wt, kparam, u10m, v10m, mr, bpdf, missing, d_latch);
  AA_CHECK(hipGetLastError());
  return 0;
}

}  // namespace aa
