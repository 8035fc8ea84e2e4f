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
//   k_bvp:   the banded boundary-value problem (REG/vlidort_bvproblem.f90:
//            147-2311) for 64/G items per wavefront (G lanes per item): matrix
//            rows are built from the layer solutions, factorised by partial
//            pivoting with the pivot choices of DGBTF2
//            (REG/lapack_tools.f90:156-357) while the right-hand side is carried
//            along (DGBTRS forward part, :7354); back-substitution in DTBSV's
//            column order then yields the integration constants layer by layer,
//            with the TOA source recursion (REG/vlidort_intensity.f90:145-643)
//            and the Fourier convergence test
//            (REG/vlidort_converge.f90:825-1231) folded in.
//
// Panel LU: with the unknowns ordered layer by layer (2n per layer) the n TOA
// rows touch layer 0 only, the 2n rows of boundary k touch layers k and k+1,
// and the n BOA rows touch the bottom layer.  DGBTF2's band window at a column
// of layer k therefore holds, besides structural zeros, exactly the n rows left
// over from layer k-1 and the 2n rows of boundary k: one panel of 3n rows over
// the 4n columns of layers k, k+1 and the right-hand side.  Each lane of an
// item's group holds up to three of these rows in registers (row slots); the
// pivot search is a group reduction, the pivot row is broadcast through LDS,
// and the rows of the next boundary move into the slots whose rows became
// pivots.  Row positions are tracked as DGBTF2 numbers them (ties go to the
// smallest position), so pivots, interchanges and updated values are those of
// the band factorisation without its arithmetic on structural zeros.  U rows
// (with 1/pivot) go, in compact form, to a per-item buffer that the
// back-substitution reads layer by layer.
#pragma once
#include <algorithm>
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

// Lane exchange of a double within 16-lane rows (DPP); used by the lane
// reductions of the boundary-value solve.
template <int CTRL>
__device__ __forceinline__ double dpp_d(double x) {
  int lo = __double2loint(x), hi = __double2hiint(x);
  lo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xF, 0xF, false);
  hi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// ------------------------------------------------------------------ k_layer
template <int N, int S, bool MIX>
__global__ __launch_bounds__(64) void k_layer(RTWork W, int M) {
  constexpr int n = N * S, LM = 2 * N;
  // MIX (the mixed path): SAB/DAB come from the FP32 MFMA contraction; from
  // N = 8 streams on, the source assembly -- the user-angle homogeneous
  // solutions (Legendre sums of X+- with the quadrature and user PI matrices)
  // and the phase-matrix sums of the beam source -- also runs in FP32 (type
  // TR), where its smaller registers pay.  Below N = 8 it stays FP64: without
  // packed FP32 the FP32 VALU rate is the FP64 one, so the conversions only
  // cost there.  The eigensystem, the normalisation and the particular
  // solution's coefficients are FP64 throughout; without MIX everything is.
  using TR = typename std::conditional<MIX && (N >= 8), float, double>::type;
  constexpr LOff O(N, S);
  constexpr QOff QO(N, S);
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int km = W.km;
  if (t >= W.lcnt[M]) return;   // dense (item, layer) slots of this term
  const int it = W.lown[t];
  const int layer = t - W.loff[it];
  const int L = W.L_it[it];
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

  // ---- SAB / DAB (REG/vlidort_solutions.f90:439-475): the omega-Greek
  // contraction of every (item, layer) comes from the matrix-core kernel
  // k_sabdab (FP32 in the mixed path, FP64 otherwise); the diagonal term is
  // added here
  using TSD = typename std::conditional<MIX, float, double>::type;
  const TSD* srec;
  if constexpr (MIX) srec = W.sab32 + il * (size_t)(2 * n * n);
  else srec = W.sab64 + il * (size_t)(2 * n * n);
  double sab[n][n];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const double xinv = -1.0 / qs[i];
#pragma unroll
    for (int j = 0; j < N; ++j)
#pragma unroll
      for (int o1 = 0; o1 < S; ++o1)
#pragma unroll
        for (int o2 = 0; o2 < S; ++o2) sab[i * S + o1][j * S + o2] = (double)srec[(i * S + o1) * n + j * S + o2];
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
        for (int c = 0; c < n; ++c) drow[c] = (double)srec[n * n + r * n + c];
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
    const double rn = 1.0 / sqrt(s);
#pragma unroll
    for (int r = 0; r < n; ++r) evec[r][c] = evec[r][c] * rn;
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
    const double rnorm_r = 1.0 / sqrt(s);
    double sum[n];
#pragma unroll
    for (int r = 0; r < n; ++r) sum[r] = evec[r][k] * rnorm_r;
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

  // ---- PI matrices at the user stream and the solar cosine
  const int obs = global_item(W, it) / W.nch;
  const double ustream = cos(DEG_TO_RAD * W.vza[obs]);
  double xup[LM][S][S], x0p[LM][S][S];
  pimatrix_x<S, LM>(M, NMOM, ustream, xup);
  pimatrix_x<S, LM>(M, NMOM, W.csza[it], x0p);

  // ---- UHOM + HMULT (:1167-1505, :1651-1790)
  // The Legendre sums of the user-angle solutions, with the quadrature
  // weights folded into the PI products and the sums over X+ and X- taken
  // apart:  hs(l, o1, k) = P(l, o1, k) + DMS(o1) pisign(l) Q(l, o1, k),
  //   P = sum_{j,o} (XQ(l, j, o1, o) a5_j) X+(j, o, k),  Q likewise with X-
  // (the reference forms each XQ X+- product inside one sum; the terms are
  // the same, the summation order differs at rounding level).
  double updn[n][S], upup[n][S], hm1[n], hm2[n];
  {
    TR ud[n][S], uu[n][S];
#pragma unroll
    for (int k = 0; k < n; ++k)
#pragma unroll
      for (int o = 0; o < S; ++o) { ud[k][o] = 0.0; uu[k][o] = 0.0; }
    if (lsc) {
#pragma unroll
      for (int l = 0; l < LM; ++l) {
        if (!inl(l)) continue;
        TR pa[n][S], qa[n][S];
#pragma unroll
        for (int k = 0; k < n; ++k)
#pragma unroll
          for (int o1 = 0; o1 < S; ++o1) { pa[k][o1] = 0.0; qa[k][o1] = 0.0; }
#pragma unroll
        for (int j = 0; j < N; ++j)
#pragma unroll
          for (int o1 = 0; o1 < S; ++o1)
#pragma unroll
            for (int o = 0; o < S; ++o) {
              const TR aq = (TR)XQ(l, j, o1, o) * (TR)qhw[j];
#pragma unroll
              for (int k = 0; k < n; ++k) {
                pa[k][o1] += aq * (TR)xp[j * S + o][k];
                qa[k][o1] += aq * (TR)xn[j * S + o][k];
              }
            }
#pragma unroll
        for (int k = 0; k < n; ++k) {
          TR hs[S], ga[S];
#pragma unroll
          for (int o1 = 0; o1 < S; ++o1) hs[o1] = pa[k][o1] + (TR)(DMS[o1] * pisign(l)) * qa[k][o1];
#pragma unroll
          for (int o1 = 0; o1 < S; ++o1) {
            TR g = 0.0;
#pragma unroll
            for (int o = 0; o < S; ++o) g += (TR)og[l][o1][o] * hs[o];
            ga[o1] = g;
          }
#pragma unroll
          for (int o1 = 0; o1 < S; ++o1) {
            TR sdn = 0.0, sup = 0.0;
#pragma unroll
            for (int o = 0; o < S; ++o) {
              const TR xum = (TR)(DMS[o1] * xup[l][o1][o] * DMS[o] * pisign(l));
              const TR xupre = (TR)(DMS[o1] * xup[l][o1][o]);
              sdn += xum * ga[o];
              sup += xupre * ga[o];
            }
            ud[k][o1] += sdn;
            uu[k][o1] += sup;
          }
        }
      }
    }
#pragma unroll
    for (int k = 0; k < n; ++k)
#pragma unroll
      for (int o = 0; o < S; ++o) { updn[k][o] = (double)ud[k][o]; upup[k][o] = (double)uu[k][o]; }
  }
#pragma unroll
  for (int k = 0; k < n; ++k) {
    const double rho_p = usec + keig[k], rho_m = usec - keig[k];
    const double zeta_p = 1.0 / rho_p;
    const bool hsingo = fabs(rho_m) < TAYLOR_SMALL;
    const double zeta_m = hsingo ? rho_m : 1.0 / rho_m;
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

  // ---- layer record: X+-, transmittances and the homogeneous source
  // coefficients now; the beam solution's terms are written as they are
  // formed below (so that they hold no registers past their use)
  double* lrec = W.lbuf + il * (size_t)W.lstride;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int o = 0; o < S; ++o)
#pragma unroll
      for (int k = 0; k < n; ++k) lrec[O.xp(i, o, k)] = xp[i * S + o][k];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int o = 0; o < S; ++o)
#pragma unroll
      for (int k = 0; k < n; ++k) lrec[O.xn(i, o, k)] = xn[i * S + o][k];
#pragma unroll
  for (int k = 0; k < n; ++k) lrec[O.tdel(k)] = tdel[k];
  // upwelling source: src(o) = sum_k lcon_k A(o,k) + mcon_k B(o,k) + C(o)
#pragma unroll
  for (int o = 0; o < S; ++o)
#pragma unroll
    for (int k = 0; k < n; ++k) {
      lrec[O.coef(o, k)] = lsc ? updn[k][o] * hm2[k] : 0.0;
      lrec[O.coef(o, n + k)] = lsc ? upup[k][o] * hm1[k] : 0.0;
    }

  // ---- Green's-function beam solution (:2649-2932); its particular source
  // C(o) (spar) accumulates per eigenvector as the A, B terms are formed
  double spar[S];
#pragma unroll
  for (int o = 0; o < S; ++o) spar[o] = 0.0;
  if (!beam) {
#pragma unroll
    for (int i = 0; i < 2 * N; ++i)
#pragma unroll
      for (int o = 0; o < S; ++o) { lrec[O.wu(i, o)] = 0.0; lrec[O.wl(i, o)] = 0.0; }
  } else {
    const int g = global_item(W, it);
    const int ch = g - obs * W.nch;
    const double F1 = (W.flux ? W.flux[(size_t)obs * W.nch + ch] : 1.0) / PI4;
    const double secbar = avsec, cst = itrans, wdel = tdmubar;
    double cfunc[n], dfunc[n];
#pragma unroll
    for (int k = 0; k < n; ++k) {
      const double zdel = tdel[k], zwdel = zdel * wdel;
      const double gamp = secbar + keig[k];
      const double gamm = secbar - keig[k];
      const double eps = gamm;
      if (fabs(eps) < TAYLOR_SMALL) cfunc[k] = taylor1(eps, dtau, wdel, cst);
      else cfunc[k] = cst * (zdel - wdel) / gamm;
      dfunc[k] = cst * (1.0 - zwdel) / gamp;
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
    double dpi[N][S], dmi[N][S];
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
        dpi[i][o1] = F1 * (double)sp; dmi[i][o1] = F1 * (double)sm;
      }
    double gup[n], gdn[n];
    const double itw = -itrans * wdel;
#pragma unroll
    for (int k = 0; k < n; ++k) {
      const double gamp = secbar + keig[k], gamm = secbar - keig[k];
      double av = 0.0, bv = 0.0;
#pragma unroll
      for (int i = 0; i < N; ++i) {
        double h1 = 0.0, h2 = 0.0, h1b = 0.0, h2b = 0.0;
#pragma unroll
        for (int o = 0; o < S; ++o) h1 += dpi[i][o] * xp[i * S + o][k];
#pragma unroll
        for (int o = 0; o < S; ++o) h1b += dmi[i][o] * xn[i * S + o][k];
        h1 = h1 + h1b;
#pragma unroll
        for (int o = 0; o < S; ++o) h2 += dmi[i][o] * xp[i * S + o][k];
#pragma unroll
        for (int o = 0; o < S; ++o) h2b += dpi[i][o] * xn[i * S + o][k];
        h2 = h2 + h2b;
        av += qw[i] * h1; bv += qw[i] * h2;
      }
      const double aterm = (double)av / normv[k];
      const double bterm = (double)bv / normv[k];
      gdn[k] = cfunc[k] * aterm;
      gup[k] = dfunc[k] * bterm;
      double sd;
      if (fabs(gamm) < TAYLOR_SMALL) sd = itransu * taylor2(gamm, sigp, dtau, 1.0, wdel * tduser, 1.0);
      else sd = (itrans * hm2[k] - emultup) / gamm;
      const double su = (itw * hm1[k] + emultup) / gamp;
#pragma unroll
      for (int o = 0; o < S; ++o) spar[o] = spar[o] + updn[k][o] * sd * aterm + upup[k][o] * su * bterm;
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
        lrec[O.wu(i, o1)] = su; lrec[O.wu(i + N, o1)] = (o1 > 1) ? -su1 : su1;
        lrec[O.wl(i + N, o1)] = (o1 > 1) ? -sl1 : sl1; lrec[O.wl(i, o1)] = sl;
      }
  }
#pragma unroll
  for (int o = 0; o < S; ++o) lrec[O.csrc(o)] = beam ? spar[o] : 0.0;

  // ---- surface terms of the bottom layer (BVP_SURFACE_SETUP_HOM/BEAM,
  //      REG/vlidort_bvproblem.f90:440-626, 1284+; GET_BOA_SOURCE,
  //      REG/vlidort_intensity.f90:1096-1496).  Lambertian: Fourier 0 only,
  //      the same reflected term in every quadrature row.  BRDF: every Fourier
  //      term, AXBID(i,j) = w_j mu_j BRDF_F(M,i,j) per row, and the BOA source
  //      weights w_j mu_j USER_BRDF_F(M,j).  With three Stokes components the
  //      Lambertian and Cox-Munk surfaces reflect I into I only (their
  //      reflection matrices, vsup/vbrdf/vbrdf_sup_kernels.f90:201-233,
  //      1018-1141, have only the (1,1) element), so their terms act on Stokes
  //      component 0; the polarised BPDF surfaces fill every Stokes row.
  const bool brdf = W.brdf != nullptr;
  if constexpr (S == 3) {
    // polarised BRDF (BPDF land surfaces): every Stokes row of the BOA
    // reflects every Stokes component of the downwelling field through the
    // 3 x 3 Mueller blocks of BRDF_F / USER_BRDF_F (the R2 and AXBID sums of
    // REG/vlidort_bvproblem.f90:575-619, 1700-1735 and the BOA source of
    // REG/vlidort_intensity.f90:1375-1395 with MUELLER_INDEX(o1, o2))
    if (W.bpdf && layer == L - 1) {
      constexpr BOff BO(N, n, S);
      constexpr BPOff PO(N);
      double* bb = W.bbuf + (size_t)it * W.bstride;
      const double* rec = W.brdf + (size_t)it * W.brdf_stride;
      const double sfac = (M == 0) ? 2.0 : 1.0;
#pragma unroll
      for (int o = 0; o < S; ++o) {
#pragma unroll
        for (int k = 0; k < n; ++k) {
          double ba = 0.0, bbv = 0.0;
#pragma unroll
          for (int j = 0; j < N; ++j)
#pragma unroll
            for (int o2 = 0; o2 < S; ++o2) {
              const double cw = qsw[j] * rec[PO.uf(M, j, o * 3 + o2)];
              ba += cw * xp[j * S + o2][k] * tdel[k];
              bbv += cw * xn[j * S + o2][k];
            }
          bb[BO.bcoef(o, k)] = ba; bb[BO.bcoef(o, n + k)] = bbv;
        }
        double rb = 0.0;
#pragma unroll
        for (int j = 0; j < N; ++j)
#pragma unroll
          for (int o2 = 0; o2 < S; ++o2) rb += qsw[j] * rec[PO.uf(M, j, o * 3 + o2)] * lrec[O.wl(j, o2)];
        bb[BO.bc(o)] = rb;
      }
#pragma unroll 1
      for (int i = 0; i < N; ++i) {
#pragma unroll
        for (int o1 = 0; o1 < S; ++o1) {
#pragma unroll
          for (int k = 0; k < n; ++k) {
            double sp = 0.0, sm = 0.0;
#pragma unroll
            for (int o2 = 0; o2 < S; ++o2)
#pragma unroll
              for (int j = 0; j < N; ++j) {
                const double ax = qsw[j] * rec[PO.f(M, i, j, o1 * 3 + o2)];
                sp += ax * xp[j * S + o2][k];
                sm += ax * xn[j * S + o2][k];
              }
            bb[BO.r2p(i, o1, k)] = sfac * sp; bb[BO.r2m(i, o1, k)] = sfac * sm;
          }
          double rbi = 0.0;
#pragma unroll
          for (int j = 0; j < N; ++j)
#pragma unroll
            for (int o2 = 0; o2 < S; ++o2) rbi += qsw[j] * rec[PO.f(M, i, j, o1 * 3 + o2)] * lrec[O.wl(j, o2)];
          bb[BO.r2beam(i, o1)] = sfac * rbi;
        }
      }
      return;
    }
  }
  if (layer == L - 1 && (brdf || M == 0)) {
    const int g = global_item(W, it);
    const int ch = g - obs * W.nch;
    const double albedo = brdf ? 0.0 : W.albedo[(size_t)ch * W.nobs + obs];
    if (brdf || albedo != 0.0) {
      constexpr BOff BO(N, n, S);
      constexpr BROff RO(N);
      double* bb = W.bbuf + (size_t)it * W.bstride;
      const double* rec = brdf ? W.brdf + (size_t)it * W.brdf_stride : nullptr;
      const double sfac = (M == 0) ? 2.0 : 1.0;
      const double kmult = 2.0 * albedo;
      double cw[N];
#pragma unroll
      for (int j = 0; j < N; ++j) cw[j] = brdf ? qsw[j] * rec[RO.uf(M, j)] : qsw[j];
#pragma unroll
      for (int k = 0; k < n; ++k) {
        double ba = 0.0, bbv = 0.0;
#pragma unroll
        for (int j = 0; j < N; ++j) { ba += cw[j] * xp[j * S][k] * tdel[k]; bbv += cw[j] * xn[j * S][k]; }
        bb[BO.bcoef(0, k)] = ba; bb[BO.bcoef(0, n + k)] = bbv;
      }
      double rb = 0.0;
#pragma unroll
      for (int j = 0; j < N; ++j) rb += cw[j] * lrec[O.wl(j, 0)];
      bb[BO.bc(0)] = rb;
      if (!brdf) {
        double rq = 0.0;
#pragma unroll
        for (int j = 0; j < N; ++j) rq += qsw[j] * lrec[O.wl(j, 0)];
#pragma unroll
        for (int k = 0; k < n; ++k) {
          double sp = 0.0, sm = 0.0;
#pragma unroll
          for (int j = 0; j < N; ++j) { sp += qsw[j] * xp[j * S][k]; sm += qsw[j] * xn[j * S][k]; }
#pragma unroll
          for (int i = 0; i < N; ++i) { bb[BO.r2p(i, 0, k)] = sp * kmult; bb[BO.r2m(i, 0, k)] = sm * kmult; }
        }
#pragma unroll
        for (int i = 0; i < N; ++i) bb[BO.r2beam(i, 0)] = rq * kmult;
      } else {
#pragma unroll
        for (int i = 0; i < N; ++i) {
          double ax[N];
#pragma unroll
          for (int j = 0; j < N; ++j) ax[j] = qsw[j] * rec[RO.f(M, i, j)];
#pragma unroll
          for (int k = 0; k < n; ++k) {
            double sp = 0.0, sm = 0.0;
#pragma unroll
            for (int j = 0; j < N; ++j) { sp += ax[j] * xp[j * S][k]; sm += ax[j] * xn[j * S][k]; }
            bb[BO.r2p(i, 0, k)] = sfac * sp; bb[BO.r2m(i, 0, k)] = sfac * sm;
          }
          double rbi = 0.0;
#pragma unroll
          for (int j = 0; j < N; ++j) rbi += ax[j] * lrec[O.wl(j, 0)];
          bb[BO.r2beam(i, 0)] = sfac * rbi;
        }
      }
    }
  }
}

// ------------------------------------------------------------------ k_bvp
// Group reductions over the G lanes that hold one item's rows (G = 4, 8, 16,
// 32 or 64; every lane of the group ends with the same value).
// Values of the two 16-lane rows of each 32-lane half, in both rows
// (v_permlane16_swap on x against itself): for every lane, {r.x, r.y} holds
// the value of its own row and of the partner row at the same position.
__device__ __forceinline__ void xrow16_d(double x, double& p, double& q) {
  const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)__double2loint(x), (unsigned)__double2loint(x), false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)__double2hiint(x), (unsigned)__double2hiint(x), false, false);
  p = __hiloint2double((int)hi[0], (int)lo[0]);
  q = __hiloint2double((int)hi[1], (int)lo[1]);
}
template <int G>
__device__ __forceinline__ double grp_sum(double x) {
  x += dpp_d<0xB1>(x);
  x += dpp_d<0x4E>(x);
  if (G >= 8) x += dpp_d<0x141>(x);
  if (G >= 16) x += dpp_d<0x140>(x);
  if (G >= 32) { double p, q; xrow16_d(x, p, q); x = p + q; }
  if (G == 64) x = readlane_d(x, 0) + readlane_d(x, 32);
  return x;
}
// (the DPP moves take INT_MAX, the identity, as the old value so they fold
// into v_min_i32's DPP form)
template <int G>
__device__ __forceinline__ int grp_min_i(int x) {
  x = min(x, __builtin_amdgcn_update_dpp(0x7fffffff, x, 0xB1, 0xF, 0xF, false));
  x = min(x, __builtin_amdgcn_update_dpp(0x7fffffff, x, 0x4E, 0xF, 0xF, false));
  if (G >= 8) x = min(x, __builtin_amdgcn_update_dpp(0x7fffffff, x, 0x141, 0xF, 0xF, false));
  if (G >= 16) x = min(x, __builtin_amdgcn_update_dpp(0x7fffffff, x, 0x140, 0xF, 0xF, false));
  if (G >= 32) {
    const auto r = __builtin_amdgcn_permlane16_swap((unsigned)x, (unsigned)x, false, false);
    x = min((int)r[0], (int)r[1]);
  }
  if (G == 64) x = min(__builtin_amdgcn_readlane(x, 0), __builtin_amdgcn_readlane(x, 32));
  return x;
}
// unsigned max; the DPP moves use 0 (the identity) as the old value so they
// fold into v_max_u32's DPP form
template <int G>
__device__ __forceinline__ unsigned grp_max_u(unsigned x) {
  x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0xB1, 0xF, 0xF, false));
  x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x4E, 0xF, 0xF, false));
  if (G >= 8) x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x141, 0xF, 0xF, false));
  if (G >= 16) x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x140, 0xF, 0xF, false));
  if (G >= 32) {
    const auto r = __builtin_amdgcn_permlane16_swap(x, x, false, false);
    x = max((unsigned)r[0], (unsigned)r[1]);
  }
  if (G == 64) x = max((unsigned)__builtin_amdgcn_readlane((int)x, 0), (unsigned)__builtin_amdgcn_readlane((int)x, 32));
  return x;
}

// Lanes per item in k_bvp (G, bvp_group / bvp_group_alt) and panel rows per
// lane (SL): the 3n rows of a panel fill SL row slots per lane (8 lanes and
// up to 3 slots for n <= 8, so a wavefront carries 8 items); the adjoint
// accumulation of the TOA radiance updates columns t + G b on lane t.

// A pivot row goes through LDS from column (i + 1) & ~1 (the columns right of
// its pivot, from an even column so the pairs stay 16-byte aligned).
__host__ __device__ constexpr int urow_start(int i) { return (i + 1) & ~1; }

// Pivot of each group from per-slot candidate flags: the candidate of the
// group is the one in the lowest slot, and in it the lowest lane (a group
// reduction over the code slot * 64 + lane); -1 if the group has none.
// `multi` (wave-uniform) reports a group with more than one candidate.
// Returns the lane of the lane's own group and sets `slot`.
template <int G, int SL>
__device__ __forceinline__ int grp_pivot(const bool (&eq)[SL], int g, int t, bool& multi, int& slot) {
  constexpr int NONE = 0x7fff;
  int code = NONE;
#pragma unroll
  for (int s = SL - 1; s >= 0; --s) code = eq[s] ? s * 64 + t : code;
  const int cmin = grp_min_i<G>(code);
  bool other = false;
#pragma unroll
  for (int s = 0; s < SL; ++s) other = other || (eq[s] && s * 64 + t != cmin);
  multi = __ballot(other) != 0ull;
  slot = (cmin == NONE) ? 0 : (cmin >> 6);
  return (cmin == NONE) ? -1 : g * G + (cmin & 63);
}

template <int N, int S, int G>
__global__ __launch_bounds__(64, (N * S <= 6) ? 2 : 1) void k_bvp(RTWork W, int M, const int* __restrict__ active,
                                            const int* __restrict__ nactive, int* __restrict__ next,
                                            int* __restrict__ nnext) {
  constexpr int n = N * S, n2 = 2 * n, NC = 4 * n, RH = NC, UW = NC + 2;
  constexpr int IPW = 64 / G, SL = bvp_slots(n, G);
  constexpr int NAC = (NC + G - 1) / G;   // accumulator columns per lane in the adjoint of the TOA recursion
  static_assert(2 * n <= 64 && SL <= 5 && G >= 4, "panel rows exceed the wavefront");
  constexpr LOff O(N, S);
  constexpr BOff BO(N, n, S);
  constexpr int BIG = 0x3fffffff;
  // The current pivot row of each group (columns right of the pivot, RHS,
  // 1/pivot), two rows in turn; column c of row i is at us[g][(i & 1) * UW + c].
  // The groups' blocks are padded by two doubles so that the groups'
  // broadcast reads of the same column fall in different LDS banks.
  constexpr int SR = 2 * UW + 2;
  __shared__ __align__(16) double us[IPW][SR];
  // adjoint state of each group: acc[o][h][c] (h = layer parity, c < 2n) and
  // the current layer's c_o; padded like us
  constexpr int AS = 3 * S * n2 + 2;
  __shared__ double adj[IPW][AS];
  const int lane = threadIdx.x, g = lane / G, t = lane - g * G;
  const unsigned long long gm = (G == 64) ? ~0ull : (((1ull << (G & 63)) - 1ull) << (g * G));
  const unsigned long long below = (1ull << lane) - 1ull;
  // this wavefront's IPW items: the term's items are in LBINS lists by layer
  // count, each list's wavefronts in turn
  int first = -1, cntb = 0;
  const int* lst = active;
  {
    int acc = 0;
#pragma unroll
    for (int b = 0; b < LBINS; ++b) {
      const int c = nactive[b], nbk = (c + IPW - 1) / IPW;
      if (first < 0 && (int)blockIdx.x < acc + nbk) {
        first = ((int)blockIdx.x - acc) * IPW;
        cntb = c;
        lst = active + (size_t)b * W.cap;
      }
      acc += nbk;
    }
  }
  if (first < 0) return;   // uniform over the wavefront
  const int ai = first + g;
  bool valid = ai < cntb;
  const int it = valid ? lst[ai] : lst[first];
  const int km = W.km;
  const int gi = global_item(W, it), obs = gi / W.nch, ch = gi - obs * W.nch;
  const size_t oi = (size_t)ch * W.nobs + obs;
  if (valid && W.status[it] != 0) {
    if (t == 0) {
      // status 6: negative direct-bounce BRDF, which the MODIS driver reports as -500
      const double fill = (W.status[it] == 6) ? -500.0 : W.missing;
      W.radiance[oi] = fill; W.reflectance[oi] = fill;
      if (S > 1) { W.qout[oi] = W.missing; W.uout[oi] = W.missing; }
      W.nfourier[oi] = 0;
    }
    valid = false;
  }
  const int L = valid ? W.L_it[it] : 0;
  int Lmax = 0;
#pragma unroll
  for (int q = 0; q < IPW; ++q) Lmax = max(Lmax, __builtin_amdgcn_readlane(L, q * G));
  const int LS = W.lstride;
  const double* __restrict__ lb = W.lbuf + (size_t)it * km * LS;
  const double* __restrict__ bb = W.bbuf + (size_t)it * W.bstride;
  constexpr BROff RO(N);
  constexpr BPOff PO(N);
  const bool brdf = W.brdf != nullptr;
  const bool pol = S == 3 && W.bpdf != nullptr;   // polarised BRDF: every Stokes row reflects
  const double* __restrict__ brec = brdf ? W.brdf + (size_t)it * W.brdf_stride : nullptr;
  const double albedo = brdf ? 0.0 : W.albedo[oi];
  const bool surf = brdf || ((M == 0) && (albedo != 0.0));
  const double tsb = W.tsb[it];
  const double ff = W.flux ? W.flux[(size_t)obs * W.nch + ch] : 1.0;
  // DIRECTRADIANCE (REG/vlidort_miscsetups.f90:1296-1614): reflected direct
  // beam ATTN * albedo (Lambertian, Fourier 0) or ATTN * BRDF_F_0(M, i)
  double attn = 0.0;
  if (surf && tsb != 0.0) {
    double x0 = 4.0 * W.csza[it] / ((M == 0) ? 1.0 : 2.0);
    x0 = ff * x0 / PI4;
    attn = x0 * tsb;
  }
  const double rf_direct = attn * albedo;

  // ---------------- panel LU with partial pivoting + forward substitution
  // Each lane of the group holds SL rows of the panel: a[s][0..2n) the columns
  // of layer k, a[s][2n..4n) those of layer k+1, a[s][4n] the right-hand side.
  double a[SL][NC + 1];
  bool act[SL];
  int posl[SL];
#pragma unroll
  for (int sl = 0; sl < SL; ++sl) {
#pragma unroll
    for (int c = 0; c <= NC; ++c) a[sl][c] = 0.0;
    act[sl] = false;
    posl[sl] = BIG;
  }
  bool sing = false;

  // ---------------- adjoint of the TOA source recursion
  // The Fourier component of the TOA radiance at the user angle
  // (REG/vlidort_intensity.f90:145-643) is linear in the integration
  // constants x of the boundary-value problem:  I_o = c_o . x + d_o, with
  //   c_o(layer k, unknown r) = P_k coef_k(o, r)   (+ P_L kmult bcoef(o, r) on
  //                                                  the bottom layer: o = 0, or
  //                                                  every o for a polarised BRDF)
  //   d_o = sum_k P_k csrc_k(o)                     (+ P_L kmult bc(o), the same)
  // and P_k the product of the user-angle transmittances of the layers above
  // layer k (the source recursion cum = (sv + cs) + tdu cum unrolled).  With
  // the band LU of the elimination, U x = b' (b' the transformed right-hand
  // side), so c.x = w.b' where U^T w = c, and w_j = (c_j - sum_{p<j} U_pj
  // w_p) / U_jj is known as soon as row j of U is.  The rows of U are used
  // once, in the elimination, and never stored.  The sums acc_c of the
  // current and the next layer's columns and c of the current layer are kept
  // per group in LDS; lane t updates columns t + G b.  Every lane of the
  // group carries the same w_j and the same running w . b'.
  double ia[S], dsum[S];
#pragma unroll
  for (int o = 0; o < S; ++o) {
    ia[o] = 0.0;
    dsum[o] = 0.0;
  }
  for (int e = t; e < AS; e += G) adj[g][e] = 0.0;
  double ptr = 1.0;   // P_k
  const double kmult = brdf ? ((M == 0) ? 2.0 : 1.0) : 2.0 * albedo;
  const double* __restrict__ trr = W.tr + (size_t)it * km * TR_NFIELDS;
  // TOA rows (BVP_MATRIX_SETUP, upper boundary), layer 0: row r to lane
  // r % G, slot r / G
#pragma unroll
  for (int sl = 0; sl < SL; ++sl) {
    const int r = t + G * sl;
    if (valid && r < n) {
      const int i = r / S, o1 = r - (r / S) * S;
#pragma unroll
      for (int q = 0; q < n; ++q) a[sl][q] = lb[O.xp(i, o1, q)];
#pragma unroll
      for (int q = 0; q < n; ++q) a[sl][n + q] = lb[O.xn(i, o1, q)] * lb[O.tdel(q)];
      a[sl][RH] = -lb[O.wu(i, o1)];
      act[sl] = true;
      posl[sl] = r;
    }
  }

  // one panel: eliminate its 2n columns; LAST: the bottom layer (no layer k+1)
  auto lu_panel = [&](auto lastc, int k) __attribute__((always_inline)) {
    constexpr bool LAST = decltype(lastc)::value;
    constexpr int NCOL = LAST ? n2 : NC;
    const int j0 = n2 * k, cur = k & 1;
#pragma unroll
    for (int i = 0; i < n2; ++i) {
      const int j = j0 + i;
      // pivot: largest |a(r,j)| among the candidate rows, ties to the smallest
      // row position (IDAMAX order of DGBTF2).  |a| orders as its bit
      // pattern: the high words decide unless two candidates share one.
      unsigned hs[SL], key = 0u;
      bool eq[SL];
#pragma unroll
      for (int sl = 0; sl < SL; ++sl) {
        hs[sl] = act[sl] ? ((unsigned)__double2hiint(a[sl][i]) & 0x7fffffffu) : 0u;
        key = max(key, hs[sl]);
      }
      const unsigned mh = grp_max_u<G>(key);
#pragma unroll
      for (int sl = 0; sl < SL; ++sl) eq[sl] = act[sl] && hs[sl] == mh;
      bool multi;
      int psl;
      int sel = grp_pivot<G, SL>(eq, g, t, multi, psl);
      if (multi) {   // wave-uniform: some group has equal high words
        unsigned ls[SL], k2 = 0u;
#pragma unroll
        for (int sl = 0; sl < SL; ++sl) {
          ls[sl] = eq[sl] ? (unsigned)__double2loint(a[sl][i]) : 0u;
          k2 = max(k2, ls[sl]);
        }
        const unsigned ml = grp_max_u<G>(k2);
#pragma unroll
        for (int sl = 0; sl < SL; ++sl) eq[sl] = eq[sl] && ls[sl] == ml;
        sel = grp_pivot<G, SL>(eq, g, t, multi, psl);
        if (multi) {
          int k3 = BIG;
#pragma unroll
          for (int sl = 0; sl < SL; ++sl) k3 = min(k3, eq[sl] ? posl[sl] : BIG);
          const int kp = grp_min_i<G>(k3);
#pragma unroll
          for (int sl = 0; sl < SL; ++sl) eq[sl] = eq[sl] && posl[sl] == kp;
          sel = grp_pivot<G, SL>(eq, g, t, multi, psl);
        }
      }
      bool ispiv[SL];
#pragma unroll
      for (int sl = 0; sl < SL; ++sl) ispiv[sl] = lane == sel && psl == sl;
      // pivot value and position from the pivot lane (LDS crossbar), so the
      // reciprocal is ready while the pivot row makes its way through LDS
      const int src = 4 * (sel < 0 ? lane : sel);
      double myv = a[0][i];
      int myp = posl[0];
#pragma unroll
      for (int sl = 1; sl < SL; ++sl) {
        myv = (psl == sl) ? a[sl][i] : myv;
        myp = (psl == sl) ? posl[sl] : myp;
      }
      const double pv = __hiloint2double(__builtin_amdgcn_ds_bpermute(src, __double2hiint(myv)),
                                         __builtin_amdgcn_ds_bpermute(src, __double2loint(myv)));
      const int ppos = __builtin_amdgcn_ds_bpermute(src, myp);
      const double rp = 1.0 / pv;
      double* __restrict__ urow = &us[g][(i & 1) * UW];   // column-indexed
#pragma unroll
      for (int sl = 0; sl < SL; ++sl) {
        if (ispiv[sl]) {   // aligned pairs from the row's first stored column
#pragma unroll
          for (int c = 0; c < NCOL; c += 2)
            if (c >= urow_start(i)) *reinterpret_cast<double2*>(&urow[c]) = make_double2(a[sl][c], a[sl][c + 1]);
          *reinterpret_cast<double2*>(&urow[RH]) = make_double2(a[sl][RH], rp);
        }
      }
      // the workgroup is one wavefront, whose LDS operations complete in
      // order: a wavefront-scope fence (no wait on outstanding memory) keeps
      // the compiler from moving the reads above the pivot row's writes
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      double u[NC + 1];
#pragma unroll
      for (int c = 0; c < NCOL; ++c)
        if (c > i) u[c] = urow[c];
      u[RH] = urow[RH];
#pragma unroll
      for (int sl = 0; sl < SL; ++sl) {
        const double l = (act[sl] && !ispiv[sl]) ? a[sl][i] * rp : 0.0;
#pragma unroll
        for (int c = 0; c < NCOL; ++c)
          if (c > i) a[sl][c] = fma(-l, u[c], a[sl][c]);
        a[sl][RH] = fma(-l, u[RH], a[sl][RH]);
        if (ispiv[sl]) {
          act[sl] = false;
        } else if (act[sl] && posl[sl] == j) {
          posl[sl] = ppos;   // row interchange: the row at position j moves to the pivot's position
        }
      }
      sing = sing || (pv == 0.0);
      // adjoint: every lane of the group forms w_j = (c_j - acc_j) / U_jj
      // and adds w_j b'_j (b'_j = u[RH], the transformed right-hand side of
      // the pivot row); each lane adds U_jc w_j to the accumulators of its
      // columns c > j (the rest of this layer, then the next)
#pragma unroll
      for (int o = 0; o < S; ++o) {
        double* __restrict__ acc = &adj[g][o * 2 * n2];
        const double w = (adj[g][2 * S * n2 + o * n2 + i] - acc[cur * n2 + i]) * rp;
        ia[o] = fma(w, u[RH], ia[o]);
#pragma unroll
        for (int b = 0; b < NAC; ++b) {
          const int c = t + G * b;
          if (c > i && c < NCOL) {
            const int e = (c < n2) ? cur * n2 + c : (cur ^ 1) * n2 + (c - n2);
            acc[e] = fma(urow[c], w, acc[e]);
          }
        }
      }
    }
    if (!LAST) {
#pragma unroll
      for (int sl = 0; sl < SL; ++sl)
#pragma unroll
        for (int c = 0; c < n2; ++c) { a[sl][c] = a[sl][c + n2]; a[sl][c + n2] = 0.0; }
    }
  };

  // row `rank` of boundary k (between layers k and k+1) or, for the bottom
  // layer, of the BOA rows with the Lambertian reflection (BVP_SURFACE_SETUP)
  auto load_row = [&](double (&r)[NC + 1], int k, bool lastp, int rank) __attribute__((always_inline)) {
    const double* lu = lb + (size_t)k * LS;
    const int i = rank / S, o1 = rank - (rank / S) * S;
    double sp, sm;
    if (!lastp) {
      const double* ll = lu + LS;
      const double* xpr = lu + O.xrow(true, i, o1, sp);
      const double* xnr = lu + O.xrow(false, i, o1, sm);
      const double* xpl = ll + O.xrow(true, i, o1, sp);
      const double* xnl = ll + O.xrow(false, i, o1, sm);
#pragma unroll
      for (int q = 0; q < n; ++q) r[q] = lu[O.tdel(q)] * (sp * xpr[q]);
#pragma unroll
      for (int q = 0; q < n; ++q) r[n + q] = sm * xnr[q];
#pragma unroll
      for (int q = 0; q < n; ++q) r[n2 + q] = -(sp * xpl[q]);
#pragma unroll
      for (int q = 0; q < n; ++q) r[3 * n + q] = -ll[O.tdel(q)] * (sm * xnl[q]);
      r[RH] = ll[O.wu(i, o1)] - lu[O.wl(i, o1)];
    } else {
      const bool sr = surf && (o1 == 0 || pol);
      const double* xpr = lu + O.xrow(true, i + N, o1, sp);
      const double* xnr = lu + O.xrow(false, i + N, o1, sm);
#pragma unroll
      for (int q = 0; q < n; ++q) {
        double xpn = sp * xpr[q];
        if (sr) xpn = xpn - bb[BO.r2p(i, o1, q)];
        r[q] = lu[O.tdel(q)] * xpn;
      }
#pragma unroll
      for (int q = 0; q < n; ++q) {
        double xmn = sm * xnr[q];
        if (sr) xmn = xmn - bb[BO.r2m(i, o1, q)];
        r[n + q] = xmn;
      }
      double v = -lu[O.wl(i + N, o1)];
      if (sr) v = v + bb[BO.r2beam(i, o1)];
      if (sr) v = v + (pol ? attn * brec[PO.f0(M, i, o1)] : brdf ? attn * brec[RO.f0(M, i)] : rf_direct);
      r[RH] = v;
    }
  };

  for (int k = 0; k < Lmax; ++k) {
    const bool here = valid && k < L;
    const bool lastp = here && k == L - 1;
    // the rows of boundary k (or the BOA rows) go to the row slots retired so
    // far, slot 0 of every lane before slot 1
    const int nnew = lastp ? n : n2;
    int base = 0;
#pragma unroll
    for (int sl = 0; sl < SL; ++sl) {
      const unsigned long long fb = __ballot(here && !act[sl]) & gm;
      const int rank = base + __popcll(fb & below);
      if (here && !act[sl] && rank < nnew) {
        load_row(a[sl], k, lastp, rank);
        act[sl] = true;
        posl[sl] = n + n2 * k + rank;
      }
      base += __popcll(fb);
    }
    if (here) {   // c and d of layer k; the next layer's accumulators start at 0
      const double* lr = lb + (size_t)k * LS;
      const double tdu = trr[(size_t)k * TR_NFIELDS + TR_TDUSER];
      const double pl = ptr * tdu;   // P_L on the bottom layer
      const int nxt = (k & 1) ^ 1;
#pragma unroll
      for (int o = 0; o < S; ++o) {
        for (int c = t; c < n2; c += G) {
          double cv = ptr * lr[O.coef(o, c)];
          if ((o == 0 || pol) && lastp && surf) cv = fma(pl * kmult, bb[BO.bcoef(o, c)], cv);
          adj[g][2 * S * n2 + o * n2 + c] = cv;
          if (k > 0) adj[g][o * 2 * n2 + nxt * n2 + c] = 0.0;
        }
        dsum[o] = fma(ptr, lr[O.csrc(o)], dsum[o]);
      }
      if (lastp && surf) {
#pragma unroll
        for (int o = 0; o < S; ++o)
          if (o == 0 || pol) dsum[o] = fma(pl * kmult, bb[BO.bc(o)], dsum[o]);
      }
      ptr = pl;
    }
    if (__ballot(here && !lastp) != 0ull) {
      if (here && !lastp) lu_panel(std::integral_constant<bool, false>(), k);
    }
    if (__ballot(lastp) != 0ull) {
      if (lastp) lu_panel(std::integral_constant<bool, true>(), k);
    }
  }
  double cum[S];
#pragma unroll
  for (int o = 0; o < S; ++o) cum[o] = ia[o] + dsum[o];


  // ---------------- Fourier convergence (VLIDORT_CONVERGE_OBSGEO), on the
  // group's lane 0; an item that goes on gets its place in the next term's
  // list and its (item, layer) slots, which the group's lanes then fill
  bool goes_on = false;
  if (t == 0 && valid) {
    const double flux_multiplier = (M == 0) ? 1.0 : 2.0;
    double* stk = W.stokes + (size_t)it * 3;
    bool done = false;
    if (sing) {
      W.status[it] = 4;
      W.radiance[oi] = W.missing; W.reflectance[oi] = W.missing;
      if (S > 1) { W.qout[oi] = W.missing; W.uout[oi] = W.missing; }
      W.nfourier[oi] = 0;
    } else {
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
        goes_on = true;
      }
    }
  }
  const int noff = append_items(W, it, L, goes_on, nnext, next, W.lcnt + M + 1);
  const int goff = __builtin_amdgcn_ds_bpermute(4 * (lane - t), noff);
  if (goff >= 0)
    for (int l = t; l < L; l += G) W.lown[goff + l] = it;
}

// ------------------------------------------------------------------ k_sabdab
// Mixed precision: SAB and DAB of every active (item, layer) at Fourier term M
// as one FP32 GEMM on the matrix cores,
//   C[row][pair] = sum_k Y_M[row][k] * og[pair][k],
// rows = 2n^2 (SAB then DAB entries), k = (l, o, o3) over the omega-Greek
// coefficients (REG/vlidort_solutions.f90:439-475).  One wavefront computes
// sab_groups(N, S) groups of 16 pairs x all rows with v_mfma_f32_16x16x4f32
// tiles.  When the A tiles (Y_M) are few they are loaded once per wavefront,
// and the pairs' B columns of all groups are loaded up front; large n (the
// vector path) keeps one group and loads the A tiles as it goes.
typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d4v __attribute__((ext_vector_type(4)));

// T = float: the mixed path (v_mfma_f32_16x16x4f32); T = double: the FP64
// path (v_mfma_f64_16x16x4f64).  Both take A and B in the same lane layout;
// the accumulator rows of lane l are 4*(l/16) + v for f32 but l/16 + 4*v for
// f64 (v = 0..3; checked on the device with a probe of known products).
template <int N, int S, typename T, int PG>
__global__ __launch_bounds__(64) void k_sabdab(RTWork W, int M) {
  constexpr bool F64 = std::is_same<T, double>::value;
  using V4 = typename std::conditional<F64, d4v, f4v>::type;
  constexpr int n = N * S, LM = 2 * N;
  constexpr YOff YO(N, S);
  constexpr int ROWS = YO.rows(), KK = YO.K;
  constexpr int RT = (ROWS + 15) / 16, KT = (KK + 3) / 4;
  constexpr int SAB_PG = PG;
  constexpr bool PRE = sab_tiles(N, S) <= 32;   // A tiles held in registers
  const int lane = threadIdx.x;
  const int km = W.km;
  const int nl = W.lcnt[M];
  const int kq = lane >> 4;
  bool valid[SAB_PG];
  size_t il[SAB_PG];
  bool any = false;
#pragma unroll
  for (int q = 0; q < SAB_PG; ++q) {
    // pair of this lane's B column / C column in group q
    const long p = ((long)blockIdx.x * SAB_PG + q) * 16 + (lane & 15);
    const bool v = p < nl;
    const int it = v ? W.lown[p] : 0;
    const int layer = v ? (int)(p - W.loff[it]) : 0;
    valid[q] = v;
    il[q] = (size_t)it * km + layer;
    any = any || v;
  }
  if (__ballot(any) == 0ull) return;
  T bk[SAB_PG][KT];
#pragma unroll
  for (int q = 0; q < SAB_PG; ++q) {
    const double* ogp = W.og + il[q] * (size_t)(LM * S * S);
#pragma unroll
    for (int ks = 0; ks < KT; ++ks) {
      const int k = 4 * ks + kq;
      bk[q][ks] = (valid[q] && k < KK) ? (T)ogp[k] : (T)0;
    }
  }
  const T* Y;
  if constexpr (F64) Y = W.ytab64;
  else Y = W.ytab;
  auto atile = [&](int rt, int ks) -> T {
    const int row = rt * 16 + (lane & 15), k = 4 * ks + kq;
    return (row < ROWS && k < KK) ? Y[YO.at(M, row, k)] : (T)0;
  };
  T at[PRE ? RT : 1][PRE ? KT : 1];
  if constexpr (PRE) {
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int ks = 0; ks < KT; ++ks) at[rt][ks] = atile(rt, ks);
  }
#pragma unroll
  for (int q = 0; q < SAB_PG; ++q) {
  T* out;
  if constexpr (F64) out = W.sab64 + il[q] * (size_t)ROWS;
  else out = W.sab32 + il[q] * (size_t)ROWS;
#pragma unroll
  for (int rt = 0; rt < RT; ++rt) {
    V4 c = {(T)0, (T)0, (T)0, (T)0};
#pragma unroll
    for (int ks = 0; ks < KT; ++ks) {
      T av;
      if constexpr (PRE) av = at[rt][ks];
      else av = atile(rt, ks);
      if constexpr (F64) c = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bk[q][ks], c, 0, 0, 0);
      else c = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bk[q][ks], c, 0, 0, 0);
    }
    if constexpr (F64) {
      // rows rt*16 + kq + 4v: one store each
      if (valid[q]) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = rt * 16 + kq + 4 * v;
          if (r < ROWS) out[r] = c[v];
        }
      }
    } else {
      // the lane's 4 accumulator rows are consecutive in its pair's record:
      // one b128 store when records hold whole 4-row groups (and so stay
      // 16-byte aligned), else row by row
      const int r0 = rt * 16 + 4 * kq;
      if constexpr (ROWS % 4 == 0) {
        if (valid[q] && r0 < ROWS) *reinterpret_cast<V4*>(out + r0) = c;
      } else if (valid[q]) {
#pragma unroll
        for (int v = 0; v < 4; ++v)
          if (r0 + v < ROWS) out[r0 + v] = c[v];
      }
    }
  }
  }
}

// One Fourier-term sweep with launch shapes (G lanes per item in k_bvp, PG
// 16-pair groups per wavefront in k_sabdab); every shape gives the same bits.
template <int N, int S, int G, int PG>
int fourier_loop_shape(const RTWork& W, int nitems, int* const lists[2], int* counts, hipStream_t st) {
  const int nl = (int)(((size_t)nitems * W.km + 63) / 64);
  const int ns = (int)(((size_t)nitems * W.km + 16 * PG - 1) / (16 * PG));
  const int nb = (nitems + 64 / G - 1) / (64 / G) + LBINS;
  for (int M = 0; M <= W.nmoments; ++M) {
    const int* in = lists[M & 1];
    int* out = lists[(M + 1) & 1];
    prof_begin(PROF_SABDAB, st);
    if (W.precision == 1) hipLaunchKernelGGL((k_sabdab<N, S, float, PG>), dim3(ns), dim3(64), 0, st, W, M);
    else hipLaunchKernelGGL((k_sabdab<N, S, double, PG>), dim3(ns), dim3(64), 0, st, W, M);
    prof_end(PROF_SABDAB, st);
    prof_begin(PROF_LAYER, st);
    if (W.precision == 1) hipLaunchKernelGGL((k_layer<N, S, true>), dim3(nl), dim3(64), 0, st, W, M);
    else hipLaunchKernelGGL((k_layer<N, S, false>), dim3(nl), dim3(64), 0, st, W, M);
    prof_end(PROF_LAYER, st);
    prof_begin(PROF_BVP, st);
    hipLaunchKernelGGL((k_bvp<N, S, G>), dim3(nb), dim3(64), 0, st, W, M, in, counts + M * LBINS, out,
                       counts + (M + 1) * LBINS);
    prof_end(PROF_BVP, st);
    if (W.mid_event && M == std::min(W.mid_term, W.nmoments)) AA_CHECK(hipEventRecord(W.mid_event, st));
  }
  AA_CHECK(hipGetLastError());
  return 0;
}

// The launch shapes of W.shape_bvp / W.shape_sab (0: default, 1: the
// alternative the tuner may pick).
template <int N, int S>
int fourier_loop(const RTWork& W, int nitems, int* const lists[2], int* counts, hipStream_t st) {
  constexpr int n = N * S;
  constexpr int G0 = bvp_group(n), G1 = bvp_group_alt(n);
  constexpr int P0 = sab_groups(N, S), P1 = sab_groups_alt(N, S);
  const bool gb = W.shape_bvp == 1, gs = W.shape_sab == 1;
  if (!gb && !gs) return fourier_loop_shape<N, S, G0, P0>(W, nitems, lists, counts, st);
  if (gb && !gs) return fourier_loop_shape<N, S, G1, P0>(W, nitems, lists, counts, st);
  if (!gb && gs) return fourier_loop_shape<N, S, G0, P1>(W, nitems, lists, counts, st);
  return fourier_loop_shape<N, S, G1, P1>(W, nitems, lists, counts, st);
}

}  // namespace
}  // namespace aa
