// Fourier-term kernels of the discrete-ordinate radiance path.
//
//   k_layer: one thread per (item, layer).  Homogeneous solutions
//            (VLIDORT_QHOM_SOLUTION + ASYMTX, REG/vlidort_solutions.f90:150-1047,
//            REG/vlidort_aux.f90:386-959), their normalisation
//            (VLIDORT_QHOM_NORMS :1051-1163), user-angle solutions and
//            multipliers (VLIDORT_UHOM_SOLUTION :1167-1505, HMULT :1651-1790),
//            Green's-function beam solution (VLIDORT_GBEAM_SOLUTION :2649-2932)
//            and the layer's upwelling source coefficients
//            (REG/vlidort_PostProcessing.f90:139-451).
//   k_bvp:   one wavefront per item.  The banded boundary-value problem
//            (REG/vlidort_bvproblem.f90:147-2311): matrix rows are generated on
//            the fly from the layer solutions, factorised by partial-pivoting
//            band elimination with the same pivot choices as DGBTF2
//            (REG/lapack_tools.f90:156-357) while the right-hand side is carried
//            along (DGBTRS forward part, :7354); back-substitution then yields
//            the integration constants column by column and the TOA source
//            recursion (REG/vlidort_intensity.f90:145-643) and the Fourier
//            convergence test (REG/vlidort_converge.f90:825-1231) are folded in.
//
// Band LU layout: lanes hold the columns of the active window (width
// 2*KL+1, circular over Q = 64*CPL-1 virtual lanes, the last virtual lane
// holds the right-hand side); every lane keeps the KL+1 candidate rows of its
// column in registers.  The pivot search runs in the lane owning the pivot
// column and the multipliers are broadcast with readlane, so a step needs no
// LDS and no barrier.  Eliminated rows (U and the transformed right-hand
// side) go to a per-item buffer that the back-substitution reads in reverse.
#pragma once
#include <type_traits>
#include "rt_work.h"
#include "rt_device.h"
#include "prof.h"

namespace aa {

namespace {

constexpr double DMS[4] = {1.0, 1.0, -1.0, -1.0};   // DMAT diagonal
constexpr double VLIDORT_ACCURACY_D = (double)0.0001f;   // V283/VLIDORT_Mod.F90:232

// Row slots of one band column live in up to five vectors of 8 doubles
// (P##0..P##4).  With a wave-uniform slot index these macros compile to
// indexed register moves; the vectors must stay plain local variables (no
// references or lambdas) for that to happen.
typedef double d8v __attribute__((ext_vector_type(8)));
#define AA_SLOT_GET(P, r, s)                                 \
  {                                                          \
    const int k_ = (s) >> 3, e_ = (s) & 7;                   \
    r = P##0[e_];                                            \
    if (NV > 1 && k_ == 1) r = P##1[e_];                     \
    if (NV > 2 && k_ == 2) r = P##2[e_];                     \
    if (NV > 3 && k_ == 3) r = P##3[e_];                     \
    if (NV > 4 && k_ == 4) r = P##4[e_];                     \
  }
#define AA_SLOT_SET(P, s, x)                                 \
  {                                                          \
    const int k_ = (s) >> 3, e_ = (s) & 7;                   \
    if (k_ == 0) P##0[e_] = (x);                             \
    if (NV > 1 && k_ == 1) P##1[e_] = (x);                   \
    if (NV > 2 && k_ == 2) P##2[e_] = (x);                   \
    if (NV > 3 && k_ == 3) P##3[e_] = (x);                   \
    if (NV > 4 && k_ == 4) P##4[e_] = (x);                   \
  }
#define AA_XV(K) d8v{xc[K * 8 + 0], xc[K * 8 + 1], xc[K * 8 + 2], xc[K * 8 + 3], \
                     xc[K * 8 + 4], xc[K * 8 + 5], xc[K * 8 + 6], xc[K * 8 + 7]}
#define AA_SLOT_AXPY(P, w)                                   \
  {                                                          \
    P##0 = P##0 - AA_XV(0) * (w);                            \
    if (NV > 1) P##1 = P##1 - AA_XV(1) * (w);                \
    if (NV > 2) P##2 = P##2 - AA_XV(2) * (w);                \
    if (NV > 3) P##3 = P##3 - AA_XV(3) * (w);                \
    if (NV > 4) P##4 = P##4 - AA_XV(4) * (w);                \
  }

__device__ __forceinline__ double readlane_d(double v, int lane) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_readlane(lo, lane);
  hi = __builtin_amdgcn_readlane(hi, lane);
  return __hiloint2double(hi, lo);
}

// Lane exchange of a double within 16-lane rows (DPP) and across the two rows
// of a 32-lane half (swizzle); used by the max-reductions of the band LU.
template <int CTRL>
__device__ __forceinline__ double dpp_d(double x) {
  int lo = __double2loint(x), hi = __double2hiint(x);
  lo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xF, 0xF, false);
  hi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double swz16_d(double x) {
  int lo = __double2loint(x), hi = __double2hiint(x);
  lo = __builtin_amdgcn_ds_swizzle(lo, 0x401F);
  hi = __builtin_amdgcn_ds_swizzle(hi, 0x401F);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }

// ------------------------------------------------------------------ k_layer
template <int N, int S, bool MIX>
__global__ __launch_bounds__(64) void k_layer(RTWork W, int M, const int* __restrict__ active,
                                              const int* __restrict__ nactive) {
  constexpr int n = N * S, LM = 2 * N;
  // MIX: SAB/DAB come from the FP32 MFMA contraction, the source-assembly
  // contractions accumulate in FP32; eigensystem, exponentials and
  // normalisations stay FP64.
  using TR = typename std::conditional<MIX, float, double>::type;
  constexpr LOff O(N, S);
  constexpr QOff QO(N, S);
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int km = W.km;
  const int a = t / km, layer = t - a * km;
  if (a >= *nactive) return;
  const int it = active[a];
  const int L = W.L_it[it];
  if (layer >= L) return;
  const int NMOM = W.nmoments;
  const double* __restrict__ Qt = W.quad;
  auto XQ = [&](int l, int i, int s1, int s2) -> double { return Qt[QO.xq(M, l, i, s1, s2)]; };
  auto pisign = [&](int l) -> double { return (((l - M) & 1) == 0) ? 1.0 : -1.0; };
  auto inl = [&](int l) -> bool { return l >= M && l <= NMOM; };
  double qs[N], qw[N], qhw[N], qsw[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    qs[i] = Qt[QO.qs(i)]; qw[i] = Qt[QO.qw(i)];
    qhw[i] = 0.5 * qw[i]; qsw[i] = qs[i] * qw[i];
  }
  const size_t il = (size_t)it * km + layer;
  const double* trr = W.tr + il * TR_NFIELDS;
  const double dtau = trr[TR_DTAU], tduser = trr[TR_TDUSER], avsec = trr[TR_AVSEC], itrans = trr[TR_ITRANS];
  const double tdmubar = trr[TR_TDMUBAR], itransu = trr[TR_ITRANSU], emultup = trr[TR_EMULTUP], sigp = trr[TR_SIGP];
  const bool lsc = (M < 3) ? true : (((W.lsc[il] >> M) & 1u) != 0u);
  const bool beam = (layer + 1 <= W.beam_cutoff[it]) && lsc;
  const double usec = W.usec[it];
  double og[LM][S][S];
  {
    const double* ogp = W.og + il * (size_t)(LM * S * S);
#pragma unroll
    for (int l = 0; l < LM; ++l)
#pragma unroll
      for (int o1 = 0; o1 < S; ++o1)
#pragma unroll
        for (int o2 = 0; o2 < S; ++o2) og[l][o1][o2] = ogp[(l * S + o1) * S + o2];
  }

  // ---- SAB / DAB (REG/vlidort_solutions.f90:439-475)
  // (dp, dm) of stream pair (i,o1), (j,o2)
  auto dpdm = [&](int i, int o1, int j, int o2, double& dp, double& dm) {
    dp = 0.0; dm = 0.0;
#pragma unroll
    for (int l = 0; l < LM; ++l) {
      if (!inl(l)) continue;
      double sp = 0.0, sm = 0.0;
#pragma unroll
      for (int o = 0; o < S; ++o) {
        double h2p = 0.0, h2m = 0.0;
#pragma unroll
        for (int o3 = 0; o3 < S; ++o3) {
          const double xqv = XQ(l, j, o3, o2);
          h2p += og[l][o][o3] * xqv;
          h2m += og[l][o][o3] * (DMS[o3] * xqv * pisign(l));
        }
        const double xi = XQ(l, i, o1, o);
        sp += xi * h2p; sm += xi * h2m;
      }
      dp += sp; dm += sm;
    }
  };
  const float* s32 = MIX ? W.sab32 + il * (size_t)(2 * n * n) : nullptr;
  double sab[n][n];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const double xinv = -1.0 / qs[i];
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const double fac = xinv * qhw[j];
#pragma unroll
      for (int o1 = 0; o1 < S; ++o1)
#pragma unroll
        for (int o2 = 0; o2 < S; ++o2) {
          if (MIX) {
            sab[i * S + o1][j * S + o2] = (double)s32[(i * S + o1) * n + j * S + o2];
          } else {
            double dp, dm; dpdm(i, o1, j, o2, dp, dm);
            sab[i * S + o1][j * S + o2] = fac * (dp + dm);
          }
        }
    }
#pragma unroll
    for (int o1 = 0; o1 < S; ++o1) sab[i * S + o1][i * S + o1] -= xinv;
  }

  // ---- eigensystem of DAB.SAB with the reference's tolerance retries (:545-575)
  double evec[n][n], eval[n];
  int fail = 0, ier = 0;
  for (int attempt = 0; attempt < 3; ++attempt) {
    const double tol = attempt == 0 ? ASYMTX_TOL : (attempt == 1 ? 1.0e-10 : 1.0e-8);
    double am[n][n];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const double xinv = -1.0 / qs[i];
#pragma unroll
      for (int o1 = 0; o1 < S; ++o1) {
        const int r = i * S + o1;
        double drow[n];
#pragma unroll
        for (int j = 0; j < N; ++j) {
          const double fac = xinv * qhw[j];
#pragma unroll
          for (int o2 = 0; o2 < S; ++o2) {
            if (MIX) {
              drow[j * S + o2] = (double)s32[n * n + r * n + j * S + o2];
            } else {
              double dp, dm; dpdm(i, o1, j, o2, dp, dm);
              drow[j * S + o2] = fac * (dp - dm);
            }
          }
        }
        drow[r] -= xinv;
#pragma unroll
        for (int c = 0; c < n; ++c) {
          double s = 0.0;
#pragma unroll
          for (int o = 0; o < S; ++o)
#pragma unroll
            for (int j = 0; j < N; ++j) s += drow[j * S + o] * sab[j * S + o][c];
          am[r][c] = s;
        }
      }
    }
#pragma unroll
    for (int r = 0; r < n; ++r)
#pragma unroll
      for (int c = 0; c < n; ++c) evec[r][c] = 0.0;
    ier = asymtx_dev<n>(tol, am, evec, eval, &fail);
    if (fail || ier == 0) break;
  }
  if (fail || ier) { W.status[it] = 4; return; }

  // ---- normalised eigenvectors, X+ / X- (:580-700)
#pragma unroll
  for (int c = 0; c < n; ++c) {
    double s = 0.0;
#pragma unroll
    for (int r = 0; r < n; ++r) s += evec[r][c] * evec[r][c];
    const double nr = sqrt(s);
#pragma unroll
    for (int r = 0; r < n; ++r) evec[r][c] = evec[r][c] / nr;
  }
  double keig[n], tdel[n];
  double xp[n][n], xn[n][n];   // X(i,o,k) for i < N at row i*S+o, column k
#pragma unroll
  for (int k = 0; k < n; ++k) {
    const double kv = sqrt(eval[k]);
    keig[k] = kv;
    const double sc_r = 1.0 / kv;
    double s = 0.0;
#pragma unroll
    for (int r = 0; r < n; ++r) s += evec[r][k] * evec[r][k];
    const double norm_r = sqrt(s);
    double sum[n];
#pragma unroll
    for (int r = 0; r < n; ++r) sum[r] = evec[r][k] / norm_r;
#pragma unroll
    for (int r = 0; r < n; ++r) {
      double h = 0.0;
#pragma unroll
      for (int o = 0; o < S; ++o)
#pragma unroll
        for (int j = 0; j < N; ++j) h += sab[r][j * S + o] * sum[j * S + o];
      const double dif = h * sc_r;
      xp[r][k] = 0.5 * (sum[r] + dif);
      xn[r][k] = 0.5 * (sum[r] - dif);
    }
    const double help = kv * dtau;
    tdel[k] = (help > MAX_TAU_PATH) ? 0.0 : exp(-help);
  }
  // X for the lower half-space streams (i + N)
  auto XPF = [&](int i, int o, int k) -> double {
    if (i < N) return xp[i * S + o][k];
    const int ii = i - N;
    return (o < 2) ? xn[ii * S + o][k] : -xn[ii * S + o][k];
  };
  auto XNF = [&](int i, int o, int k) -> double {
    if (i < N) return xn[i * S + o][k];
    const int ii = i - N;
    return (o < 2) ? xp[ii * S + o][k] : -xp[ii * S + o][k];
  };

  // ---- PI matrices at the user stream and the solar cosine
  const int obs = global_item(W, it) / W.nch;
  const double ustream = cos(DEG_TO_RAD * W.vza[obs]);
  double xup[LM][S][S], x0p[LM][S][S];
  pimatrix_x<S, LM>(M, NMOM, ustream, xup);
  pimatrix_x<S, LM>(M, NMOM, W.csza[it], x0p);

  // ---- UHOM + HMULT (:1167-1505, :1651-1790)
  double updn[n][S], upup[n][S], hm1[n], hm2[n];
#pragma unroll
  for (int k = 0; k < n; ++k) {
    const double rho_p = usec + keig[k], rho_m = usec - keig[k];
    const double zeta_p = 1.0 / rho_p;
    const bool hsingo = fabs(rho_m) < TAYLOR_SMALL;
    const double zeta_m = hsingo ? rho_m : 1.0 / rho_m;
    if (!lsc) {
#pragma unroll
      for (int o = 0; o < S; ++o) { updn[k][o] = 0.0; upup[k][o] = 0.0; }
    } else {
      TR gaux[LM][S];
#pragma unroll
      for (int l = 0; l < LM; ++l) {
        TR hs[S];
#pragma unroll
        for (int o1 = 0; o1 < S; ++o1) {
          TR sps = 0.0;
          if (inl(l)) {
#pragma unroll
            for (int j = 0; j < N; ++j) {
              const TR a5 = (TR)qhw[j];
              TR sp = 0.0, sm = 0.0;
#pragma unroll
              for (int o = 0; o < S; ++o) {
                const TR xqv = (TR)XQ(l, j, o1, o);
                sp += xqv * a5 * (TR)xp[j * S + o][k];
                sm += ((TR)DMS[o1] * xqv * (TR)pisign(l)) * a5 * (TR)xn[j * S + o][k];
              }
              sps = sps + sp + sm;
            }
          }
          hs[o1] = sps;
        }
#pragma unroll
        for (int o1 = 0; o1 < S; ++o1) {
          TR s = 0.0;
#pragma unroll
          for (int o = 0; o < S; ++o) s += (TR)og[l][o1][o] * hs[o];
          gaux[l][o1] = s;
        }
      }
#pragma unroll
      for (int o1 = 0; o1 < S; ++o1) {
        TR sdn = 0.0, sup = 0.0;
#pragma unroll
        for (int o = 0; o < S; ++o)
#pragma unroll
          for (int l = 0; l < LM; ++l) {
            if (!inl(l)) continue;
            const TR xum = (TR)(DMS[o1] * xup[l][o1][o] * DMS[o] * pisign(l));
            const TR xupre = (TR)(DMS[o1] * xup[l][o1][o]);
            sdn += xum * gaux[l][o];
            sup += xupre * gaux[l][o];
          }
        updn[k][o1] = sdn; upup[k][o1] = sup;
      }
    }
    const double udel = tduser;
    const double zdel = tdel[k], zudel = zdel * udel;
    const double th2 = 1.0 - zudel, th1 = zdel - udel;
    hm1[k] = hsingo ? taylor1(zeta_m, dtau, udel, usec) : usec * th1 * zeta_m;
    hm2[k] = usec * th2 * zeta_p;
  }

  // ---- norms (:1051-1163)
  double normv[n];
#pragma unroll
  for (int k = 0; k < n; ++k) {
    double nv = 0.0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
      double t1 = 0.0, t2 = 0.0;
#pragma unroll
      for (int o = 0; o < S; ++o) { t1 += xp[j * S + o][k] * xp[j * S + o][k]; t2 += xn[j * S + o][k] * xn[j * S + o][k]; }
      nv = nv + qsw[j] * (t1 - t2);
    }
    normv[k] = nv;
  }

  // ---- Green's-function beam solution (:2649-2932)
  double wu[2 * N][S], wl[2 * N][S];
  double aterm[n], bterm[n], gamp[n], gamm[n];
#pragma unroll
  for (int k = 0; k < n; ++k) { aterm[k] = 0.0; bterm[k] = 0.0; gamp[k] = 0.0; gamm[k] = 0.0; }
#pragma unroll
  for (int i = 0; i < 2 * N; ++i)
#pragma unroll
    for (int o = 0; o < S; ++o) { wu[i][o] = 0.0; wl[i][o] = 0.0; }
  if (beam) {
    const int g = global_item(W, it);
    const int ch = g - obs * W.nch;
    const double F1 = (W.flux ? W.flux[(size_t)obs * W.nch + ch] : 1.0) / PI4;
    const double secbar = avsec, cst = itrans, wdel = tdmubar;
    double cfunc[n], dfunc[n];
#pragma unroll
    for (int k = 0; k < n; ++k) {
      const double zdel = tdel[k], zwdel = zdel * wdel;
      gamp[k] = secbar + keig[k];
      gamm[k] = secbar - keig[k];
      const double eps = gamm[k];
      if (fabs(eps) < TAYLOR_SMALL) cfunc[k] = taylor1(eps, dtau, wdel, cst);
      else cfunc[k] = cst * (zdel - wdel) / gamm[k];
      dfunc[k] = cst * (1.0 - zwdel) / gamp[k];
    }
    TR hq[LM][S];
#pragma unroll
    for (int o1 = 0; o1 < S; ++o1)
#pragma unroll
      for (int l = 0; l < LM; ++l) {
        TR gsum = 0.0;
#pragma unroll
        for (int o2 = 0; o2 < S; ++o2) gsum = gsum + (TR)og[l][o1][o2] * (TR)x0p[l][o2][0];
        hq[l][o1] = gsum;
      }
    TR dpi[N][S], dmi[N][S];
#pragma unroll
    for (int o1 = 0; o1 < S; ++o1)
#pragma unroll
      for (int i = 0; i < N; ++i) {
        TR sp = 0.0, sm = 0.0;
#pragma unroll
        for (int o = 0; o < S; ++o)
#pragma unroll
          for (int l = 0; l < LM; ++l) {
            if (!inl(l)) continue;
            const TR xqv = (TR)XQ(l, i, o1, o);
            sp += xqv * hq[l][o];
            sm += (xqv * (TR)DMS[o] * (TR)pisign(l)) * hq[l][o];
          }
        dpi[i][o1] = (TR)F1 * sp; dmi[i][o1] = (TR)F1 * sm;
      }
    double gup[n], gdn[n];
#pragma unroll
    for (int k = 0; k < n; ++k) {
      TR av = 0.0, bv = 0.0;
#pragma unroll
      for (int i = 0; i < N; ++i) {
        TR h1 = 0.0, h2 = 0.0, h1b = 0.0, h2b = 0.0;
#pragma unroll
        for (int o = 0; o < S; ++o) h1 += dpi[i][o] * (TR)xp[i * S + o][k];
#pragma unroll
        for (int o = 0; o < S; ++o) h1b += dmi[i][o] * (TR)xn[i * S + o][k];
        h1 = h1 + h1b;
#pragma unroll
        for (int o = 0; o < S; ++o) h2 += dmi[i][o] * (TR)xp[i * S + o][k];
#pragma unroll
        for (int o = 0; o < S; ++o) h2b += dpi[i][o] * (TR)xn[i * S + o][k];
        h2 = h2 + h2b;
        av += (TR)qw[i] * h1; bv += (TR)qw[i] * h2;
      }
      aterm[k] = (double)av / normv[k];
      bterm[k] = (double)bv / normv[k];
      gdn[k] = cfunc[k] * aterm[k];
      gup[k] = dfunc[k] * bterm[k];
    }
#pragma unroll
    for (int o1 = 0; o1 < S; ++o1)
#pragma unroll
      for (int i = 0; i < N; ++i) {
        double su = 0.0, su1 = 0.0, sl1 = 0.0, sl = 0.0;
#pragma unroll
        for (int k = 0; k < n; ++k) {
          su += gup[k] * xn[i * S + o1][k]; su1 += gup[k] * xp[i * S + o1][k];
          sl1 += gdn[k] * xn[i * S + o1][k]; sl += gdn[k] * xp[i * S + o1][k];
        }
        wu[i][o1] = su; wu[i + N][o1] = (o1 > 1) ? -su1 : su1;
        wl[i + N][o1] = (o1 > 1) ? -sl1 : sl1; wl[i][o1] = sl;
      }
  }

  // ---- write the layer record
  double* rec = W.lbuf + il * (size_t)W.lstride;
#pragma unroll
  for (int i = 0; i < 2 * N; ++i)
#pragma unroll
    for (int o = 0; o < S; ++o)
#pragma unroll
      for (int k = 0; k < n; ++k) rec[O.xp(i, o, k)] = XPF(i, o, k);
#pragma unroll
  for (int i = 0; i < 2 * N; ++i)
#pragma unroll
    for (int o = 0; o < S; ++o)
#pragma unroll
      for (int k = 0; k < n; ++k) rec[O.xn(i, o, k)] = XNF(i, o, k);
#pragma unroll
  for (int k = 0; k < n; ++k) rec[O.tdel(k)] = tdel[k];
#pragma unroll
  for (int i = 0; i < 2 * N; ++i)
#pragma unroll
    for (int o = 0; o < S; ++o) { rec[O.wu(i, o)] = wu[i][o]; rec[O.wl(i, o)] = wl[i][o]; }
  // upwelling source: src(o) = sum_k lcon_k A(o,k) + mcon_k B(o,k) + C(o)
#pragma unroll
  for (int o = 0; o < S; ++o) {
    double spar = 0.0;
#pragma unroll
    for (int k = 0; k < n; ++k) {
      rec[O.coef(o, k)] = lsc ? updn[k][o] * hm2[k] : 0.0;
      rec[O.coef(o, n + k)] = lsc ? upup[k][o] * hm1[k] : 0.0;
      if (beam) {
        const double wdel = tdmubar, itw = -itrans * wdel;
        double sd;
        if (fabs(gamm[k]) < TAYLOR_SMALL) sd = itransu * taylor2(gamm[k], sigp, dtau, 1.0, wdel * tduser, 1.0);
        else sd = (itrans * hm2[k] - emultup) / gamm[k];
        const double su = (itw * hm1[k] + emultup) / gamp[k];
        spar = spar + updn[k][o] * sd * aterm[k] + upup[k][o] * su * bterm[k];
      }
    }
    rec[O.csrc(o)] = beam ? spar : 0.0;
  }
  // ---- Lambertian surface terms of the bottom layer at Fourier 0
  //      (BVP_SURFACE_SETUP_HOM/BEAM, REG/vlidort_bvproblem.f90; GET_BOA_SOURCE,
  //      REG/vlidort_intensity.f90:1096-1496)
  if (layer == L - 1 && M == 0) {
    const int g = global_item(W, it);
    const int ch = g - obs * W.nch;
    const double albedo = W.albedo[(size_t)ch * W.nobs + obs];
    if (albedo != 0.0) {
      constexpr BOff BO(n);
      double* bb = W.bbuf + (size_t)it * W.bstride;
      const double kmult = 2.0 * albedo;
#pragma unroll
      for (int k = 0; k < n; ++k) {
        double sp = 0.0, sm = 0.0, ba = 0.0, bbv = 0.0;
#pragma unroll
        for (int j = 0; j < N; ++j) {
          sp += qsw[j] * xp[j * S][k]; sm += qsw[j] * xn[j * S][k];
          ba += qsw[j] * xp[j * S][k] * tdel[k]; bbv += qsw[j] * xn[j * S][k];
        }
        bb[BO.r2p(k)] = sp * kmult; bb[BO.r2m(k)] = sm * kmult;
        bb[BO.bcoef(k)] = ba; bb[BO.bcoef(n + k)] = bbv;
      }
      double rb = 0.0;
#pragma unroll
      for (int j = 0; j < N; ++j) rb += qsw[j] * wl[j][0];
      bb[BO.r2beam()] = rb * kmult;
      bb[BO.bc()] = rb;
    }
  }
}

// ------------------------------------------------------------------ k_bvp
template <int N, int S>
__global__ __launch_bounds__(64) void k_bvp(RTWork W, int M, const int* __restrict__ active,
                                            const int* __restrict__ nactive, int* __restrict__ next,
                                            int* __restrict__ nnext) {
  constexpr int n = N * S, n2 = 2 * n, KL = 3 * n - 1, R = KL + 1, WD = 2 * KL + 1;
  constexpr int CPL = (WD + 1 > 64) ? 2 : 1;
  constexpr int P = 64 * CPL, Q = P - 1;
  constexpr int PF = 4;    // back-substitution: right-hand sides loaded PF steps ahead
  static_assert(WD + PF + 1 <= P, "window does not fit the wavefront");
  constexpr LOff O(N, S);
  constexpr BOff BO(n);
  constexpr int BIG = 0x3fffffff;
  const int ai = blockIdx.x;
  if (ai >= *nactive) return;
  const int it = active[ai];
  const int lane = threadIdx.x;
  const int km = W.km;
  const int g = global_item(W, it), obs = g / W.nch, ch = g - obs * W.nch;
  const size_t oi = (size_t)ch * W.nobs + obs;
  if (W.status[it] != 0) {
    if (lane == 0) {
      W.radiance[oi] = W.missing; W.reflectance[oi] = W.missing;
      if (S > 1) { W.qout[oi] = W.missing; W.uout[oi] = W.missing; }
      W.nfourier[oi] = 0;
    }
    return;
  }
  const int L = W.L_it[it];
  const int NT = n2 * L;
  const int LS = W.lstride, UW = W.ustride;
  const double* __restrict__ lb = W.lbuf + (size_t)it * km * LS;
  const double* __restrict__ bb = W.bbuf + (size_t)it * W.bstride;
  double* __restrict__ ub = W.ubuf + (size_t)it * (size_t)n2 * km * UW;
  const double albedo = W.albedo[oi];
  const bool surf = (M == 0) && (albedo != 0.0);
  const double tsb = W.tsb[it];
  const double ff = W.flux ? W.flux[(size_t)obs * W.nch + ch] : 1.0;
  double rf_direct = 0.0;
  if (surf && tsb != 0.0) {   // DIRECTRADIANCE (REG/vlidort_miscsetups.f90:1296-1614), Lambertian
    double x0 = 4.0 * W.csza[it] / 1.0;
    x0 = ff * x0 / PI4;
    rf_direct = x0 * tsb * albedo;
  }
  const double r2beam = surf ? bb[BO.r2beam()] : 0.0;

  // BVP matrix element (row r, column c), 0-based (BVP_MATRIX_SETUP,
  // REG/vlidort_bvproblem.f90:627-1176)
  auto elem = [&](int r, int c) -> double {
    if (r < n) {
      const int i = r / S, o1 = r - (r / S) * S;
      if (c < n) return lb[O.xp(i, o1, c)];
      if (c < n2) return lb[O.xn(i, o1, c - n)] * lb[O.tdel(c - n)];
      return 0.0;
    }
    if (r >= NT - n) {
      const int rr = r - (NT - n);
      const int i = rr / S, o1 = rr - (rr / S) * S;
      const int q = c - (NT - n2);
      if (q < 0 || q >= n2) return 0.0;
      const double* lr = lb + (size_t)(L - 1) * LS;
      if (q < n) {
        double xpn = lr[O.xp(i + N, o1, q)];
        if (surf && o1 == 0) xpn = xpn - bb[BO.r2p(q)];
        return lr[O.tdel(q)] * xpn;
      }
      double xmn = lr[O.xn(i + N, o1, q - n)];
      if (surf && o1 == 0) xmn = xmn - bb[BO.r2m(q - n)];
      return xmn;
    }
    const int tt = r - n;
    const int nl0 = tt / n2, irow = tt - nl0 * n2;
    const int i = irow / S, o1 = irow - (irow / S) * S;
    const int q = c - nl0 * n2;
    if (q < 0 || q >= 2 * n2) return 0.0;
    const double* lu = lb + (size_t)nl0 * LS;
    const double* ll = lu + LS;
    if (q < n) return lu[O.tdel(q)] * lu[O.xp(i, o1, q)];
    if (q < n2) return lu[O.xn(i, o1, q - n)];
    if (q < 3 * n) return -ll[O.xp(i, o1, q - n2)];
    return -ll[O.tdel(q - 3 * n)] * ll[O.xn(i, o1, q - 3 * n)];
  };
  // right-hand side (BVP_COLUMN_SETUP + BVP_SURFACE_SETUP_BEAM)
  auto rhs = [&](int r) -> double {
    if (r < n) { const int i = r / S, o1 = r - (r / S) * S; return -lb[O.wu(i, o1)]; }
    if (r >= NT - n) {
      const int rr = r - (NT - n);
      const int i = rr / S, o1 = rr - (rr / S) * S;
      double v = -lb[(size_t)(L - 1) * LS + O.wl(i + N, o1)];
      if (surf && o1 == 0) v = v + r2beam;
      if (surf && o1 == 0) v = v + rf_direct;
      return v;
    }
    const int tt = r - n;
    const int nl0 = tt / n2, irow = tt - nl0 * n2;
    const int i = irow / S, o1 = irow - (irow / S) * S;
    return lb[(size_t)(nl0 + 1) * LS + O.wu(i, o1)] - lb[(size_t)nl0 * LS + O.wl(i, o1)];
  };
  // value of row r at the column held by virtual lane v when the window starts
  // at column j0 (jm0 = j0 mod Q)
  auto gen = [&](int v, int r, int j0, int jm0) -> double {
    if (r >= NT) return 0.0;
    if (v == Q) return rhs(r);
    int off = v - jm0;
    if (off < 0) off += Q;
    const int c = j0 + off;
    if (off >= WD || c >= NT) return 0.0;
    return elem(r, c);
  };

  // ---------------- band LU with partial pivoting + forward substitution
  // Step j: the lane owning column j copies it to LDS; lanes 0..R-1 read it
  // back transposed (lane = row slot) and a butterfly finds the pivot: largest
  // |a(r,j)|, ties to the smallest row position (IDAMAX order of DGBTF2).  Every
  // lane then eliminates its column with the broadcast column:
  //   a(s,c) -= a(s,j) * (a(p,c) / a(p,j)).
  // The column leaving the window is zeroed exactly (its own scale is 1), so a
  // lane re-entering as a new column starts from zeros.  The row entering the
  // window after step j (position j+R) is independent of the pivoting and is
  // loaded PD steps ahead.
  constexpr int PD = 4;
  constexpr int NV = (R + 7) / 8;
  static_assert(NV <= 5 && R <= 64, "candidate rows exceed the register layout");
  __shared__ double colbuf[NV * 8];
  d8v a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, a4 = 0.0;   // column slot 0
  d8v b0 = 0.0, b1 = 0.0, b2 = 0.0, b3 = 0.0, b4 = 0.0;   // column slot 1 (two columns per lane)
#pragma unroll
  for (int s = 0; s < R; ++s) {
    AA_SLOT_SET(a, s, gen(lane, s, 0, 0));
    if (CPL > 1) AA_SLOT_SET(b, s, gen(lane + 64, s, 0, 0));
  }
  int posl = (lane < R && lane < NT) ? lane : BIG;   // row position held by slot `lane`
  if (lane < NV * 8) colbuf[lane] = 0.0;
  auto fetch_row = [&](int j, double (&dst)[CPL]) {
    const int j1 = j + 1;
    const int jm1x = j1 % Q;
#pragma unroll
    for (int cs = 0; cs < CPL; ++cs) dst[cs] = gen(lane + 64 * cs, j + R, j1, jm1x);
  };
  double nvq[PD][CPL];
#pragma unroll
  for (int d = 0; d < PD; ++d) fetch_row(d, nvq[d]);
  bool singular = false;
  int jm = 0;   // j mod Q
  __syncthreads();
  for (int jb = 0; jb < NT; jb += PD) {
    double nvn[PD][CPL];
#pragma unroll
    for (int d = 0; d < PD; ++d) fetch_row(jb + PD + d, nvn[d]);
#pragma unroll
    for (int d = 0; d < PD; ++d) {
      const int j = jb + d;
      if (j >= NT) break;
      const int jm1 = (jm + 1 == Q) ? 0 : jm + 1;
      const int rnew = j + R;
      const int lj = jm & 63, csj = jm >> 6;
      // column j to LDS
      if (lane == lj) {
#pragma unroll
        for (int s = 0; s < R; ++s) {
          double x;
          if (CPL == 1 || csj == 0) AA_SLOT_GET(a, x, s) else AA_SLOT_GET(b, x, s)
          colbuf[s] = x;
        }
      }
      __syncthreads();
      // pivot search, lane = row slot: max |a(s,j)| over the candidate rows,
      // ties resolved to the smallest row position
      const double xt = (lane < R) ? colbuf[lane] : 0.0;
      double am = (lane < R && posl < BIG) ? fabs(xt) : -1.0;
      am = dmax(am, dpp_d<0xB1>(am));
      am = dmax(am, dpp_d<0x4E>(am));
      am = dmax(am, dpp_d<0x141>(am));
      am = dmax(am, dpp_d<0x140>(am));
      if (R > 16) am = dmax(am, swz16_d(am));
      double mx = readlane_d(am, 0);
      if (R > 32) mx = dmax(mx, readlane_d(am, 32));
      unsigned long long cand = __ballot(lane < R && posl < BIG && fabs(xt) == mx);
      int bsu = (int)__builtin_ctzll(cand);
      if (__popcll(cand) > 1) {   // equal magnitudes: smallest position wins
        int best = BIG;
        for (unsigned long long m = cand; m; m &= m - 1) {
          const int l = (int)__builtin_ctzll(m);
          const int pl = __builtin_amdgcn_readlane(posl, l);
          if (pl < best) { best = pl; bsu = l; }
        }
      }
      const int jp = __builtin_amdgcn_readlane(posl, bsu);
      const int sj = (int)__builtin_ctzll(__ballot(lane < R && posl == j));
      const double pv = readlane_d(xt, bsu);
      if (pv == 0.0) singular = true;
      const double rp = 1.0 / pv;
      double xc[NV * 8];
#pragma unroll
      for (int s = 0; s < NV * 8; ++s) xc[s] = colbuf[s];
      // eliminate, store the pivot row (U row j and transformed RHS), and
      // put the next row into the pivot slot
      {
        double u;
        AA_SLOT_GET(a, u, bsu);
        const int v = lane;
        const double w = (v == jm) ? 1.0 : u * rp;
        AA_SLOT_AXPY(a, w);
        if (v == Q) ub[(size_t)j * UW + WD] = u;
        else {
          int off = v - jm;
          if (off < 0) off += Q;
          if (off < WD && j + off < NT) ub[(size_t)j * UW + off] = u;
        }
        AA_SLOT_SET(a, bsu, nvq[d][0]);
      }
      if (CPL > 1) {
        double u;
        AA_SLOT_GET(b, u, bsu);
        const int v = lane + 64;
        const double w = (v == jm) ? 1.0 : u * rp;
        AA_SLOT_AXPY(b, w);
        if (v == Q) ub[(size_t)j * UW + WD] = u;
        else {
          int off = v - jm;
          if (off < 0) off += Q;
          if (off < WD && j + off < NT) ub[(size_t)j * UW + off] = u;
        }
        AA_SLOT_SET(b, bsu, nvq[d][CPL - 1]);
      }
      // row interchange bookkeeping: the row at position j moves to position jp
      if (lane == sj && sj != bsu) posl = jp;
      if (lane == bsu) posl = (rnew < NT) ? rnew : BIG;
      __syncthreads();
      jm = jm1;
    }
#pragma unroll
    for (int d = 0; d < PD; ++d)
#pragma unroll
      for (int cs = 0; cs < CPL; ++cs) nvq[d][cs] = nvn[d][cs];
  }

  // ---------------- back-substitution fused with the TOA source recursion
  // Row-oriented: lanes hold the solution values of the columns in the window
  // (j, j+WD-1] (same circular lane map as the elimination) and load row j of
  // U coalesced; x_j = (y_j - sum_c U(j,c) x_c) / U(j,j) with a wave
  // reduction.  Rows are loaded PB steps ahead.
  constexpr int PB = 8;
  double xv[CPL];
#pragma unroll
  for (int cs = 0; cs < CPL; ++cs) xv[cs] = 0.0;
  const double kmult = 2.0 * albedo;
  double acc[S], cum[S];
#pragma unroll
  for (int o = 0; o < S; ++o) { acc[o] = 0.0; cum[o] = 0.0; }
  double bacc = 0.0;
  const double* trr = W.tr + (size_t)it * km * TR_NFIELDS;
  // U(j, c) for this lane's column(s) when the solve is at row j
  auto fetch_u = [&](int j, double (&dst)[CPL]) {
    const int jmq = (j >= 0) ? j % Q : 0;
#pragma unroll
    for (int cs = 0; cs < CPL; ++cs) {
      const int v = lane + 64 * cs;
      int off = v - jmq;
      if (off < 0) off += Q;
      dst[cs] = (j >= 0 && v != Q && off >= 1 && off < WD && j + off < NT) ? ub[(size_t)j * UW + off] : 0.0;
    }
  };
  double uq[PB][CPL], dq[PB], yq[PB];
#pragma unroll
  for (int d = 0; d < PB; ++d) {
    const int j = NT - 1 - d;
    fetch_u(j, uq[d]);
    dq[d] = (j >= 0) ? ub[(size_t)j * UW] : 1.0;
    yq[d] = (j >= 0) ? ub[(size_t)j * UW + WD] : 0.0;
  }
  for (int jb = NT - 1; jb >= 0; jb -= PB) {
    double un[PB][CPL], dn[PB], yn[PB];
#pragma unroll
    for (int d = 0; d < PB; ++d) {
      const int j = jb - PB - d;
      fetch_u(j, un[d]);
      dn[d] = (j >= 0) ? ub[(size_t)j * UW] : 1.0;
      yn[d] = (j >= 0) ? ub[(size_t)j * UW + WD] : 0.0;
    }
#pragma unroll
    for (int d = 0; d < PB; ++d) {
      const int j = jb - d;
      if (j < 0) break;
      double pr = 0.0;
#pragma unroll
      for (int cs = 0; cs < CPL; ++cs) pr += uq[d][cs] * xv[cs];
      pr += dpp_d<0xB1>(pr);
      pr += dpp_d<0x4E>(pr);
      pr += dpp_d<0x141>(pr);
      pr += dpp_d<0x140>(pr);
      pr += swz16_d(pr);
      const double sum = readlane_d(pr, 0) + readlane_d(pr, 32);
      const double xj = (yq[d] - sum) / dq[d];
      const int jv = j % Q;
#pragma unroll
      for (int cs = 0; cs < CPL; ++cs) if (lane + 64 * cs == jv) xv[cs] = xj;
      // TOA source: layer nl, constant index q (lcon for q < n, mcon after)
      const int nl = j / n2, q = j - nl * n2;
      const double* lr = lb + (size_t)nl * LS;
#pragma unroll
      for (int o = 0; o < S; ++o) acc[o] += xj * lr[O.coef(o, q)];
      if (surf && nl == L - 1) bacc += xj * bb[BO.bcoef(q)];
      if (q == 0) {
        if (nl == L - 1) {
#pragma unroll
          for (int o = 0; o < S; ++o) cum[o] = 0.0;
          if (surf) cum[0] = kmult * (bacc + bb[BO.bc()]);
        }
        const double tdu = trr[(size_t)nl * TR_NFIELDS + TR_TDUSER];
#pragma unroll
        for (int o = 0; o < S; ++o) {
          const double src = acc[o] + lr[O.csrc(o)];
          cum[o] = src + tdu * cum[o];
          acc[o] = 0.0;
        }
      }
    }
#pragma unroll
    for (int d = 0; d < PB; ++d) {
      dq[d] = dn[d]; yq[d] = yn[d];
#pragma unroll
      for (int cs = 0; cs < CPL; ++cs) uq[d][cs] = un[d][cs];
    }
  }

  // ---------------- Fourier convergence (VLIDORT_CONVERGE_OBSGEO)
  if (lane != 0) return;
  const double flux_multiplier = (M == 0) ? 1.0 : 2.0;
  double* stk = W.stokes + (size_t)it * 3;
  bool done = false;
  if (singular) {
    W.status[it] = 4;
    W.radiance[oi] = W.missing; W.reflectance[oi] = W.missing;
    if (S > 1) { W.qout[oi] = W.missing; W.uout[oi] = W.missing; }
    W.nfourier[oi] = 0;
    return;
  }
  double sf[S];
#pragma unroll
  for (int o = 0; o < S; ++o) sf[o] = flux_multiplier * cum[o];
  int testconv = W.testconv[it], fsaved = W.fsaved[it];
  if (M == 0) {
#pragma unroll
    for (int o = 0; o < S; ++o) stk[o] = sf[o] + W.ss[(size_t)it * 3 + o] + W.db[(size_t)it * 3 + o];
  } else {
    const double azm = W.raa[obs] * (double)M * DEG_TO_RAD;
    const double azmfac[3] = {cos(azm), cos(azm), sin(azm)};
    double tazm[S], tnew[S];
#pragma unroll
    for (int o = 0; o < S; ++o) { tazm[o] = azmfac[o] * sf[o]; tnew[o] = stk[o] + tazm[o]; }
    int count = 0;
    if (tazm[0] != 0.0) { const double accur = fabs(tazm[0] / tnew[0]); if (accur < VLIDORT_ACCURACY_D) count++; }
    else count++;
#pragma unroll
    for (int o = 0; o < S; ++o) stk[o] = tnew[o];
    if (count == 1) {
      testconv++;
      if (testconv == 2) { done = true; fsaved = M; }
    } else { testconv = 0; fsaved = 2 * N - 1; }
  }
  W.testconv[it] = testconv;
  W.fsaved[it] = fsaved;
  if (M >= W.nmoments) done = true;
  if (done) {
    const float pif = 4.0f * atanf(1.0f);
    const double pi = (double)pif;
    const double I = stk[0];
    W.radiance[oi] = I;
    W.reflectance[oi] = (pi * I) / (cos(W.sza[obs] * pi / (double)180.0f) * ff);
    if (S > 1) { W.qout[oi] = stk[1]; W.uout[oi] = stk[2]; }
    W.nfourier[oi] = fsaved + 1;
  } else {
    const int slot = atomicAdd(nnext, 1);
    next[slot] = it;
  }
}

// ------------------------------------------------------------------ k_sabdab
// Mixed precision: SAB and DAB of every active (item, layer) at Fourier term M
// as one FP32 GEMM on the matrix cores,
//   C[row][pair] = sum_k Y_M[row][k] * og[pair][k],
// rows = 2n^2 (SAB then DAB entries), k = (l, o, o3) over the omega-Greek
// coefficients (REG/vlidort_solutions.f90:439-475).  One wavefront computes
// 16 pairs x all rows with v_mfma_f32_16x16x4f32 tiles.
typedef float f4v __attribute__((ext_vector_type(4)));
template <int N, int S>
__global__ __launch_bounds__(64) void k_sabdab(RTWork W, int M, const int* __restrict__ active,
                                               const int* __restrict__ nactive) {
  constexpr int n = N * S, LM = 2 * N;
  constexpr YOff YO(N, S);
  constexpr int ROWS = YO.rows(), KK = YO.K;
  constexpr int RT = (ROWS + 15) / 16, KT = (KK + 3) / 4;
  const int lane = threadIdx.x;
  const int km = W.km;
  const long p = (long)blockIdx.x * 16 + (lane & 15);   // pair of this lane's B column / C column
  const int a = (int)(p / km), layer = (int)(p - (long)a * km);
  const int na = *nactive;
  bool valid = a < na;
  int it = 0;
  if (valid) { it = active[a]; valid = layer < W.L_it[it]; }
  if (__ballot(valid) == 0ull) return;
  const size_t il = (size_t)it * km + layer;
  const double* ogp = W.og + il * (size_t)(LM * S * S);
  const float* Y = W.ytab;
  const int kq = lane >> 4;
  float bk[KT];
#pragma unroll
  for (int ks = 0; ks < KT; ++ks) {
    const int k = 4 * ks + kq;
    bk[ks] = (valid && k < KK) ? (float)ogp[k] : 0.0f;
  }
  float* out = W.sab32 + il * (size_t)ROWS;
#pragma unroll
  for (int rt = 0; rt < RT; ++rt) {
    f4v c = {0.0f, 0.0f, 0.0f, 0.0f};
    const int row = rt * 16 + (lane & 15);
#pragma unroll
    for (int ks = 0; ks < KT; ++ks) {
      const int k = 4 * ks + kq;
      const float av = (row < ROWS && k < KK) ? Y[YO.at(M, row, k)] : 0.0f;
      c = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bk[ks], c, 0, 0, 0);
    }
    if (valid) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = rt * 16 + 4 * kq + v;
        if (r < ROWS) out[r] = c[v];
      }
    }
  }
}

template <int N, int S>
int fourier_loop(const RTWork& W, int nitems, int* const lists[2], int* counts, hipStream_t st) {
  const int nl = (int)(((size_t)nitems * W.km + 63) / 64);
  const int ns = (int)(((size_t)nitems * W.km + 15) / 16);
  for (int M = 0; M <= W.nmoments; ++M) {
    const int* in = lists[M & 1];
    int* out = lists[(M + 1) & 1];
    prof_begin(PROF_LAYER, st);
    if (W.precision == 1) {
      hipLaunchKernelGGL((k_sabdab<N, S>), dim3(ns), dim3(64), 0, st, W, M, in, counts + M);
      hipLaunchKernelGGL((k_layer<N, S, true>), dim3(nl), dim3(64), 0, st, W, M, in, counts + M);
    } else {
      hipLaunchKernelGGL((k_layer<N, S, false>), dim3(nl), dim3(64), 0, st, W, M, in, counts + M);
    }
    prof_end(PROF_LAYER, st);
    prof_begin(PROF_BVP, st);
    hipLaunchKernelGGL((k_bvp<N, S>), dim3(nitems), dim3(64), 0, st, W, M, in, counts + M, out, counts + M + 1);
    prof_end(PROF_BVP, st);
  }
  AA_CHECK(hipGetLastError());
  return 0;
}

}  // namespace
}  // namespace aa
