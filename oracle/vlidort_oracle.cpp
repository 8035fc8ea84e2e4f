// ORACLE TEST INFRASTRUCTURE -- not part of the product.
//
// Serial FP64 C++ restatement of the AeroApps discrete-ordinate radiance path
// (VLIDORT 2.8.3 with the fixed AeroApps options: Lambertian surface, solar
// source, upwelling TOA output at one observation geometry, pseudo-spherical
// Chapman beam, delta-M scaling, Green's-function particular integral,
// first-order single-scatter + direct-bounce correction with outgoing
// sphericity (NFINE=3), double convergence test at accuracy 1e-4).  It is the
// checker for the HIP path; nothing in the product links it.
//
// Reference files (V283 = src/Components/rtms/RTSI/VLIDORT2p8p3, REG =
// V283/vlidort_main/regular, FOC = V283/vlidort_focode):
//   driver glue      V283/VLIDORT_ScatMod.F90:284-617 (vector), 975-1247 (scalar)
//                    src/Components/rtms/geo_vlidort/VLIDORT_LAMBERT.F90:34-342
//   fixed options    V283/VLIDORT_Mod.F90:65-383
//   master           REG/vlidort_masters_V2p8p3.f90:131-3100 (VLIDORT_MASTER),
//                    3107-5090 (VLIDORT_FOURIER)
//   input checks     REG/vlidort_inputs.f90:2439-3825, derive 3825-4606
//   geometry         REG/vlidort_geometry.f90:119-703 (Chapman factors)
//   setups           REG/vlidort_miscsetups.f90:370-1292, 1296-1614,
//                    2273-2932, 3187-3421
//   solutions        REG/vlidort_solutions.f90:150-1790, 2649-2932
//   BVP              REG/vlidort_bvproblem.f90:147-2311 (+ REG/lapack_tools.f90
//                    DGBTF2 :156, DGBTRS :7354)
//   intensity        REG/vlidort_intensity.f90:145-643, 1096-1496
//                    REG/vlidort_PostProcessing.f90:139-451
//   convergence      REG/vlidort_converge.f90:825-1231
//   Taylor series    REG/vlidort_Taylor.f90:119-210
//   ASYMTX/GETQUAD2  REG/vlidort_aux.f90:386-1000
//   first order      REG/vlidort_vfo_interface.f90:128-728, FOC/VFO_Master.f90:123,
//                    FOC/FO_WPgeometry_Routines.f90:776-991,1511-1971,2409-2859,
//                    FOC/FO_geometry_Generic.f90:100-330,
//                    FOC/FO_VectorSS_Spherfuncs.f90:93-300,
//                    FOC/FO_VectorSS_RTCalcs_I.f90:116-785
//
// Scope: NSTOKES 1 or 3 (real eigensolver branch only; NSTOKES=4 is not on the
// AeroApps path), NLAYERS >= 1 (one layer: the band LU covers the dense
// 2n x 2n system of the reference's DGETRF branch), one observation geometry per call,
// refraction off, plane-parallel optional.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>
#include <algorithm>

namespace {

constexpr double ZERO = 0.0, ONE = 1.0, TWO = 2.0, HALF = 0.5;
constexpr double DEG_TO_RAD = 1.7453292519943e-02;       // vlidort_pars.f90:397
constexpr double PIE = 180.0 * DEG_TO_RAD;
constexpr double PI4 = 4.0 * PIE;
constexpr double TAYLOR_SMALL = 1.0e-04;
constexpr double MAX_TAU_SPATH = 32.0, MAX_TAU_UPATH = 32.0, MAX_TAU_QPATH = 32.0;
constexpr double OMEGA_SMALLNUM = 1.0e-15;
constexpr int TAYLOR_ORDER = 3;                           // VLIDORT_Mod.F90:208 (TAYLOR_ORDER = 3)
constexpr int NFINE = 3;                                  // VLIDORT_Mod.F90:224
constexpr double EARTH_RADIUS = 6371.0;                   // VLIDORT_Mod.F90:238
constexpr double ASYMTX_TOL = 1.0e-12;                    // VLIDORT_Mod.F90:309
// VLIDORT_ACCURACY = 0.0001 assigned from a default-real literal (VLIDORT_Mod.F90:232)
const double VLIDORT_ACCURACY = (double)0.0001f;

// ---------------------------------------------------------------- helpers
struct A2 {  // 0-based 2-D array, first index fastest (Fortran order)
  int n0 = 0, n1 = 0; std::vector<double> d;
  void init(int a, int b) { n0 = a; n1 = b; d.assign((size_t)a * b, 0.0); }
  double& operator()(int i, int j) { return d[(size_t)j * n0 + i]; }
  double operator()(int i, int j) const { return d[(size_t)j * n0 + i]; }
};
struct A3 {
  int n0 = 0, n1 = 0, n2 = 0; std::vector<double> d;
  void init(int a, int b, int c) { n0 = a; n1 = b; n2 = c; d.assign((size_t)a * b * c, 0.0); }
  double& operator()(int i, int j, int k) { return d[((size_t)k * n1 + j) * n0 + i]; }
};
struct A4 {
  int n0 = 0, n1 = 0, n2 = 0, n3 = 0; std::vector<double> d;
  void init(int a, int b, int c, int e) { n0 = a; n1 = b; n2 = c; n3 = e; d.assign((size_t)a * b * c * e, 0.0); }
  double& operator()(int i, int j, int k, int l) { return d[(((size_t)l * n2 + k) * n1 + j) * n0 + i]; }
};

// Fortran SIGN(a,b)
inline double fsign(double a, double b) { return b >= 0.0 ? std::fabs(a) : -std::fabs(a); }

// ---------------------------------------------------------------- GETQUAD2
// Gauss-Legendre nodes/weights on [a,b] (REG/vlidort_aux.f90:960-1000;
// FOC/FO_geometry_Generic.f90:331-385 is the same routine).
void getquad2(double a, double b, int n, double* roots, double* wgts) {
  const double QUARTER = 0.25, QEPS = 1.0e-13;
  int n2 = (n + 1) / 2;
  double nr = (double)n;
  double midpt = HALF * (b + a), sfac = HALF * (b - a);
  for (int m = 1; m <= n2; ++m) {
    double mr = (double)m;
    double xx = PIE * (mr - QUARTER) / (nr + HALF);
    double sx = std::sin(xx);
    double x = (ONE - (nr - ONE) / (8.0 * nr * nr * nr) -
                ONE / (384.0 * nr * nr * nr * nr) * (39.0 - 28.0 / (sx * sx))) * std::cos(xx);
    double lp, lpm1, dlp;
    while (true) {
      lpm1 = ZERO; lp = ONE;
      for (int i = 1; i <= n; ++i) {
        double ir = (double)i, lpm2 = lpm1; lpm1 = lp;
        lp = ((TWO * ir - ONE) * x * lpm1 - (ir - ONE) * lpm2) / ir;
      }
      dlp = nr * (x * lp - lpm1) / (x * x - ONE);
      double xold = x; x = xold - lp / dlp;
      if (std::fabs(x - xold) <= QEPS) break;
    }
    roots[m - 1] = midpt - sfac * x;
    wgts[m - 1] = (TWO * sfac) / ((ONE - x * x) * dlp * dlp);
    int nm1 = n - m + 1;
    roots[nm1 - 1] = midpt + sfac * x;
    wgts[nm1 - 1] = wgts[m - 1];
  }
}

// ---------------------------------------------------------------- ASYMTX
// Real nonsymmetric eigensystem (balance, Householder Hessenberg, shifted
// double-QR, back substitution), REG/vlidort_aux.f90:386-959.  Arrays are
// 1-based views of column-major M x M storage with leading dimension M.
// Returns BAD_STATUS in *bad; IER (>0: eigenvalue IER did not converge).
int asymtx(double tol, double* aad_, int M, double* evec_, double* eval_, bool* bad) {
  auto AAD = [&](int i, int j) -> double& { return aad_[(j - 1) * M + (i - 1)]; };
  auto EVECD = [&](int i, int j) -> double& { return evec_[(j - 1) * M + (i - 1)]; };
  auto EVALD = [&](int i) -> double& { return eval_[i - 1]; };
  std::vector<double> wk(4 * M + 1);
  auto WKD = [&](int i) -> double& { return wk[i - 1]; };
  const double C1 = 0.4375, C2 = 0.5, C3 = 0.75, C4 = 0.95, C5 = 16.0, C6 = 256.0;
  int ier = 0; *bad = false;
  if (M < 1) { *bad = true; return 0; }
  if (M == 1) { EVALD(1) = AAD(1, 1); EVECD(1, 1) = 1.0; return 0; }
  if (M == 2) {
    double discri = (AAD(1, 1) - AAD(2, 2)) * (AAD(1, 1) - AAD(2, 2)) + 4.0 * AAD(1, 2) * AAD(2, 1);
    if (discri < ZERO) { *bad = true; return 0; }
    double sgn = ONE; if (AAD(1, 1) < AAD(2, 2)) sgn = -ONE;
    EVALD(1) = 0.5 * (AAD(1, 1) + AAD(2, 2) + sgn * std::sqrt(discri));
    EVALD(2) = 0.5 * (AAD(1, 1) + AAD(2, 2) - sgn * std::sqrt(discri));
    EVECD(1, 1) = ONE; EVECD(2, 2) = ONE;
    if (AAD(1, 1) == AAD(2, 2) && (AAD(2, 1) == ZERO || AAD(1, 2) == ZERO)) {
      double rnorm = std::fabs(AAD(1, 1)) + std::fabs(AAD(1, 2)) + std::fabs(AAD(2, 1)) + std::fabs(AAD(2, 2));
      double w = tol * rnorm;
      EVECD(2, 1) = AAD(2, 1) / w; EVECD(1, 2) = -AAD(1, 2) / w;
    } else {
      EVECD(2, 1) = AAD(2, 1) / (EVALD(1) - AAD(2, 2));
      EVECD(1, 2) = AAD(1, 2) / (EVALD(2) - AAD(1, 1));
    }
    return 0;
  }
  for (int i = 1; i <= M; ++i) { EVALD(i) = ZERO; for (int j = 1; j <= M; ++j) EVECD(i, j) = ZERO; EVECD(i, i) = ONE; }
  double rnorm = ZERO;
  int L = 1, K = M;
  // search for rows isolating an eigenvalue and push them down
L30:
  {
    int kkk = K;
    for (int j = kkk; j >= 1; --j) {
      double row = ZERO;
      for (int i = 1; i <= K; ++i) if (i != j) row += std::fabs(AAD(j, i));
      if (row == ZERO) {
        WKD(K) = j;
        if (j != K) {
          for (int i = 1; i <= K; ++i) std::swap(AAD(i, j), AAD(i, K));
          for (int i = L; i <= M; ++i) std::swap(AAD(j, i), AAD(K, i));
        }
        K = K - 1;
        goto L30;
      }
    }
  }
  // search for columns isolating an eigenvalue and push them left
L80:
  {
    int lll = L;
    for (int j = lll; j <= K; ++j) {
      double col = ZERO;
      for (int i = L; i <= K; ++i) if (i != j) col += std::fabs(AAD(i, j));
      if (col == ZERO) {
        WKD(L) = j;
        if (j != L) {
          for (int i = 1; i <= K; ++i) std::swap(AAD(i, j), AAD(i, L));
          for (int i = L; i <= M; ++i) std::swap(AAD(j, i), AAD(L, i));
        }
        L = L + 1;
        goto L80;
      }
    }
  }
  for (int i = L; i <= K; ++i) WKD(i) = ONE;
  // balance the submatrix in rows L through K
  {
    bool noconv;
    do {
      noconv = false;
      for (int i = L; i <= K; ++i) {
        double col = ZERO, row = ZERO;
        for (int j = L; j <= K; ++j) if (j != i) { col += std::fabs(AAD(j, i)); row += std::fabs(AAD(i, j)); }
        double f = ONE, g = row / C5, h = col + row;
        while (col < g) { f = f * C5; col = col * C6; }
        g = row * C5;
        while (col >= g) { f = f / C5; col = col / C6; }
        if ((col + row) / f < C4 * h) {
          WKD(i) = WKD(i) * f; noconv = true;
          for (int j = L; j <= M; ++j) AAD(i, j) = AAD(i, j) / f;
          for (int j = 1; j <= K; ++j) AAD(j, i) = AAD(j, i) * f;
        }
      }
    } while (noconv);
  }
  // Householder reduction to Hessenberg form
  if (!(K - 1 < L + 1)) {
    for (int n = L + 1; n <= K - 1; ++n) {
      double h = ZERO; WKD(n + M) = ZERO;
      double scale = ZERO;
      for (int i = n; i <= K; ++i) scale += std::fabs(AAD(i, n - 1));
      if (scale != ZERO) {
        for (int i = K; i >= n; --i) { WKD(i + M) = AAD(i, n - 1) / scale; h += WKD(i + M) * WKD(i + M); }
        double g = -fsign(std::sqrt(h), WKD(n + M));
        h = h - WKD(n + M) * g;
        WKD(n + M) = WKD(n + M) - g;
        for (int j = n; j <= M; ++j) {
          double f = ZERO;
          for (int i = K; i >= n; --i) f += WKD(i + M) * AAD(i, j);
          for (int i = n; i <= K; ++i) AAD(i, j) = AAD(i, j) - WKD(i + M) * f / h;
        }
        for (int i = 1; i <= K; ++i) {
          double f = ZERO;
          for (int j = K; j >= n; --j) f += WKD(j + M) * AAD(i, j);
          for (int j = n; j <= K; ++j) AAD(i, j) = AAD(i, j) - WKD(j + M) * f / h;
        }
        WKD(n + M) = scale * WKD(n + M);
        AAD(n, n - 1) = scale * g;
      }
    }
    for (int n = K - 2; n >= L; --n) {
      int n1 = n + 1, n2 = n + 2;
      double f = AAD(n1, n);
      if (f != ZERO) {
        f = f * WKD(n1 + M);
        for (int i = n2; i <= K; ++i) WKD(i + M) = AAD(i, n);
        if (n1 <= K) {
          for (int j = 1; j <= M; ++j) {
            double g = ZERO;
            for (int i = n1; i <= K; ++i) g += WKD(i + M) * EVECD(i, j);
            g = g / f;
            for (int i = n1; i <= K; ++i) EVECD(i, j) = EVECD(i, j) + g * WKD(i + M);
          }
        }
      }
    }
  }
  // QR iterations
  {
    int n = 1;
    for (int i = 1; i <= M; ++i) {
      for (int j = n; j <= M; ++j) rnorm += std::fabs(AAD(i, j));
      n = i;
      if (i < L || i > K) EVALD(i) = AAD(i, i);
    }
  }
  {
    int n = K;
    double t = ZERO;
    int in = 0, n1 = 0, n2 = 0, lb = 0, i = 0;
    double x, y, w, z, p = 0, q = 0, r = 0, s;
  L380:
    if (n < L) goto L530;
    in = 0; n1 = n - 1; n2 = n - 2;
  L390:
    for (i = L; i <= n; ++i) {
      lb = n + L - i;
      if (lb == L) break;
      s = std::fabs(AAD(lb - 1, lb - 1)) + std::fabs(AAD(lb, lb));
      if (s == ZERO) s = rnorm;
      if (std::fabs(AAD(lb, lb - 1)) <= tol * s) break;
    }
    x = AAD(n, n);
    if (lb == n) {            // one eigenvalue found
      AAD(n, n) = x + t; EVALD(n) = AAD(n, n); n = n1; goto L380;
    }
    y = AAD(n1, n1);
    w = AAD(n, n1) * AAD(n1, n);
    if (lb == n1) {           // two eigenvalues found
      p = (y - x) * C2; q = p * p + w;
      z = std::sqrt(std::fabs(q));
      AAD(n, n) = x + t; x = AAD(n, n);
      AAD(n1, n1) = y + t;
      z = p + fsign(z, p);
      EVALD(n1) = x + z; EVALD(n) = EVALD(n1);
      if (z != ZERO) EVALD(n) = x - w / z;
      x = AAD(n, n1);
      r = std::sqrt(x * x + z * z);
      p = x / r; q = z / r;
      for (int j = n1; j <= M; ++j) { z = AAD(n1, j); AAD(n1, j) = q * z + p * AAD(n, j); AAD(n, j) = q * AAD(n, j) - p * z; }
      for (int ii = 1; ii <= n; ++ii) { z = AAD(ii, n1); AAD(ii, n1) = q * z + p * AAD(ii, n); AAD(ii, n) = q * AAD(ii, n) - p * z; }
      for (int ii = L; ii <= K; ++ii) { z = EVECD(ii, n1); EVECD(ii, n1) = q * z + p * EVECD(ii, n); EVECD(ii, n) = q * EVECD(ii, n) - p * z; }
      n = n2; goto L380;
    }
    if (in == 30) { ier = n; goto L670; }
    if (in == 10 || in == 20) {   // exceptional shift
      t = t + x;
      for (int ii = L; ii <= n; ++ii) AAD(ii, ii) = AAD(ii, ii) - x;
      s = std::fabs(AAD(n, n1)) + std::fabs(AAD(n1, n2));
      x = C3 * s; y = x; w = -C1 * s * s;
    }
    in = in + 1;
    for (int j = lb; j <= n2; ++j) {
      i = n2 + lb - j;
      z = AAD(i, i);
      r = x - z; s = y - z;
      p = (r * s - w) / AAD(i + 1, i) + AAD(i, i + 1);
      q = AAD(i + 1, i + 1) - z - r - s;
      r = AAD(i + 2, i + 1);
      s = std::fabs(p) + std::fabs(q) + std::fabs(r);
      p = p / s; q = q / s; r = r / s;
      if (i == lb) break;
      double uu = std::fabs(AAD(i, i - 1)) * (std::fabs(q) + std::fabs(r));
      double vv = std::fabs(p) * (std::fabs(AAD(i - 1, i - 1)) + std::fabs(z) + std::fabs(AAD(i + 1, i + 1)));
      if (uu <= tol * vv) break;
    }
    AAD(i + 2, i) = ZERO;
    for (int j = i + 3; j <= n; ++j) { AAD(j, j - 2) = ZERO; AAD(j, j - 3) = ZERO; }
    for (int ka = i; ka <= n1; ++ka) {
      bool notlas = ka != n1;
      if (ka == i) {
        s = fsign(std::sqrt(p * p + q * q + r * r), p);
        if (lb != i) AAD(ka, ka - 1) = -AAD(ka, ka - 1);
      } else {
        p = AAD(ka, ka - 1); q = AAD(ka + 1, ka - 1); r = ZERO;
        if (notlas) r = AAD(ka + 2, ka - 1);
        x = std::fabs(p) + std::fabs(q) + std::fabs(r);
        if (x == ZERO) continue;
        p = p / x; q = q / x; r = r / x;
        s = fsign(std::sqrt(p * p + q * q + r * r), p);
        AAD(ka, ka - 1) = -s * x;
      }
      p = p + s; x = p / s; y = q / s; z = r / s; q = q / p; r = r / p;
      for (int j = ka; j <= M; ++j) {
        p = AAD(ka, j) + q * AAD(ka + 1, j);
        if (notlas) { p = p + r * AAD(ka + 2, j); AAD(ka + 2, j) = AAD(ka + 2, j) - p * z; }
        AAD(ka + 1, j) = AAD(ka + 1, j) - p * y;
        AAD(ka, j) = AAD(ka, j) - p * x;
      }
      int iimax = std::min(n, ka + 3);
      for (int ii = 1; ii <= iimax; ++ii) {
        p = x * AAD(ii, ka) + y * AAD(ii, ka + 1);
        if (notlas) { p = p + z * AAD(ii, ka + 2); AAD(ii, ka + 2) = AAD(ii, ka + 2) - p * r; }
        AAD(ii, ka + 1) = AAD(ii, ka + 1) - p * q;
        AAD(ii, ka) = AAD(ii, ka) - p;
      }
      for (int ii = L; ii <= K; ++ii) {
        p = x * EVECD(ii, ka) + y * EVECD(ii, ka + 1);
        if (notlas) { p = p + z * EVECD(ii, ka + 2); EVECD(ii, ka + 2) = EVECD(ii, ka + 2) - p * r; }
        EVECD(ii, ka + 1) = EVECD(ii, ka + 1) - p * q;
        EVECD(ii, ka) = EVECD(ii, ka) - p;
      }
    }
    goto L390;
  L530:
    // back-substitute for eigenvectors of the triangular form
    if (rnorm != ZERO) {
      for (int nn = M; nn >= 1; --nn) {
        int nn2 = nn;
        AAD(nn, nn) = ONE;
        for (int ii = nn - 1; ii >= 1; --ii) {
          double ww = AAD(ii, ii) - EVALD(nn);
          if (ww == ZERO) ww = tol * rnorm;
          double rr = AAD(ii, nn);
          for (int j = nn2; j <= nn - 1; ++j) rr = rr + AAD(ii, j) * AAD(j, nn);
          AAD(ii, nn) = -rr / ww;
          nn2 = ii;
        }
      }
      for (int ii = 1; ii <= M; ++ii)
        if (ii < L || ii > K) for (int j = ii; j <= M; ++j) EVECD(ii, j) = AAD(ii, j);
      if (K != 0) {
        for (int j = M; j >= L; --j)
          for (int ii = L; ii <= K; ++ii) {
            double zz = ZERO;
            int nmax = std::min(j, K);
            for (int nn = L; nn <= nmax; ++nn) zz = zz + EVECD(ii, nn) * AAD(nn, j);
            EVECD(ii, j) = zz;
          }
      }
    }
    for (int ii = L; ii <= K; ++ii) for (int j = 1; j <= M; ++j) EVECD(ii, j) = EVECD(ii, j) * WKD(ii);
    for (int ii = L - 1; ii >= 1; --ii) {
      int j = (int)WKD(ii);
      if (ii != j) for (int nn = 1; nn <= M; ++nn) std::swap(EVECD(ii, nn), EVECD(j, nn));
    }
    for (int ii = K + 1; ii <= M; ++ii) {
      int j = (int)WKD(ii);
      if (ii != j) for (int nn = 1; nn <= M; ++nn) std::swap(EVECD(ii, nn), EVECD(j, nn));
    }
  }
L670:
  return ier;
}

// ---------------------------------------------------------------- Taylor
// REG/vlidort_Taylor.f90:119-133
double taylor_series_1(int order, double eps, double delta, double udel, double sm) {
  double d[10]; int mterms = order + 1; d[0] = ONE;
  for (int m = 1; m <= mterms; ++m) d[m] = delta * d[m - 1] / (double)m;
  double mult = d[1], power = ONE;
  for (int m = 2; m <= mterms; ++m) { power = power * eps; mult = mult + power * d[m]; }
  return mult * udel * sm;
}
// REG/vlidort_Taylor.f90:137-210
double taylor_series_2(int order, double eps, double y, double delta, double fac1, double fac2, double sm) {
  double d[12], ac[12], cc[12], term1[12];
  int mterms = order + 1;
  if (std::fabs(y) < TAYLOR_SMALL) mterms = mterms + 1;
  d[0] = ONE;
  for (int m = 1; m <= mterms; ++m) d[m] = delta * d[m - 1] / (double)m;
  if (std::fabs(y) < TAYLOR_SMALL) {
    double power = ONE, power2 = ONE, mult = d[2];
    for (int m = 3; m <= mterms; ++m) {
      power = power * (eps - y);
      power2 = power - y * power2;
      mult = mult + d[m] * power2;
    }
    return mult * fac1 * sm;
  }
  double y1 = ONE / y;
  ac[0] = ONE;
  for (int m = 1; m <= mterms; ++m) ac[m] = y1 * ac[m - 1];
  cc[0] = ONE;
  for (int m = 1; m <= mterms; ++m) { cc[m] = ZERO; for (int j = 0; j <= m; ++j) cc[m] = cc[m] + ac[j] * d[m - j]; }
  for (int m = 0; m <= mterms; ++m) term1[m] = fac1 * ac[m] - fac2 * cc[m];
  double power = ONE, mult = term1[1];
  for (int m = 2; m <= mterms; ++m) { power = eps * power; mult = mult + power * term1[m]; }
  return mult * sm * y1;
}

// ---------------------------------------------------------------- FO geometry helpers
// FOC/FO_geometry_Generic.f90:105-160
void find_sun(bool donadir, bool overhead, double radius, const double* sdir, double cumangle, double theta_boa_r,
              double& theta, double& stheta, double& ctheta, bool& dirsun) {
  if (donadir) { dirsun = true; theta = theta_boa_r; ctheta = std::cos(theta_boa_r); stheta = std::sin(theta_boa_r); return; }
  if (overhead) { dirsun = true; ctheta = std::cos(cumangle); stheta = std::sin(cumangle); theta = cumangle; return; }
  double px0 = -radius * std::sin(cumangle), px2 = radius * std::cos(cumangle);
  double b = px0 * sdir[0] + ZERO * sdir[1] + px2 * sdir[2];
  ctheta = -b / radius;
  dirsun = (ctheta >= ZERO);
  stheta = std::sqrt(ONE - ctheta * ctheta);
  theta = std::acos(ctheta);
}
// FOC/FO_geometry_Generic.f90:164-213.  radii[0..L], sunpaths 1-based [1..n]
void find_sunpaths_d(bool zerosunboa, double radstart, const double* radii, double thstart, double sthstart, int n,
                     double* sunpaths) {
  int n1 = n - 1;
  if (zerosunboa) {
    sunpaths[n] = radii[n1] - radstart;
    for (int k = n1; k >= 1; --k) sunpaths[k] = radii[k - 1] - radii[k];
    return;
  }
  double sth0 = sthstart, th0 = thstart;
  double sth1 = sth0 * radstart / radii[n1];
  double th1 = std::asin(sth1);
  double ks1 = th0 - th1;
  sunpaths[n] = std::sin(ks1) * radstart / sth1;
  sth0 = sth1; th0 = th1;
  for (int k = n1; k >= 1; --k) {
    sth1 = sth0 * radii[k] / radii[k - 1];
    th1 = std::asin(sth1);
    ks1 = th0 - th1;
    sunpaths[k] = std::sin(ks1) * radii[k] / sth1;
    sth0 = sth1; th0 = th1;
  }
}
// FOC/FO_geometry_Generic.f90:217-290
void find_sunpaths_t(double pie, double radstart, const double* radii, double thstart, double sthstart, int n, int L,
                     double* sunpaths, int& nt) {
  int n1 = n - 1;
  nt = n;
  double tanr = sthstart * radstart;
  int k = n1; bool trawl = true;
  while (k >= n1 && trawl && k <= L + 1) { trawl = (radii[k] > tanr); k = k + 1; }
  nt = k - 1;
  double sth0, th0, sth1, th1, ks1;
  if (nt > n) {
    th0 = pie - thstart; sth0 = sthstart;
    sth1 = sth0 * radstart / radii[n];
    th1 = std::asin(sth1); ks1 = th0 - th1;
    sunpaths[n] = TWO * std::sin(ks1) * radstart / sth1;
    sth0 = sth1; th0 = th1;
    for (int kk = n + 1; kk <= nt - 1; ++kk) {
      sth1 = sth0 * radii[kk - 1] / radii[kk];
      th1 = std::asin(sth1); ks1 = th0 - th1;
      sunpaths[kk] = TWO * std::sin(ks1) * radii[kk] / sth0;
      sth0 = sth1; th0 = th1;
    }
    sth1 = ONE; ks1 = 0.5 * pie - th0;
    sunpaths[nt] = TWO * std::sin(ks1) * radii[nt - 1];
  } else if (nt == n) {
    sunpaths[n] = -TWO * radstart * std::cos(thstart);
  }
  th0 = pie - thstart; sth0 = sthstart;
  sth1 = sth0 * radstart / radii[n1];
  th1 = std::asin(sth1); ks1 = th0 - th1;
  sunpaths[n] = sunpaths[n] + std::sin(ks1) * radstart / sth1;
  sth0 = sth1; th0 = th1;
  for (int kk = n1; kk >= 1; --kk) {
    sth1 = sth0 * radii[kk] / radii[kk - 1];
    th1 = std::asin(sth1); ks1 = th0 - th1;
    sunpaths[kk] = std::sin(ks1) * radii[kk] / sth1;
    sth0 = sth1; th0 = th1;
  }
}

struct Column {
  int L, N, S, ngreek;
  const double* deltau;   // [L]
  const double* omega;    // [L]
  const double* greek;    // greek(l, n, k) = greek[(k*L + n)*(ngreek+1) + l], k in 0..15 (Fortran (0:ng, L, 16))
  const double* height;   // [L+1] km, TOA first
  double sza, vza, raa, albedo, flux_factor;
  bool plane_parallel, focorr_nadir;
};
struct Result { double I, Q, U; int fourier_saved; int status; };

inline double G_IN(const Column& c, int l, int n, int k) { return c.greek[((size_t)k * c.L + n) * (c.ngreek + 1) + l]; }

// ---------------------------------------------------------------- first-order (single scatter + direct bounce)
// Upwelling TOA FO radiance for one observation geometry: VFO_MASTER_INTERFACE
// (REG/vlidort_vfo_interface.f90:128-728) -> FO_SSWPGeometry_Master (obsgeom,
// enhanced PS or regular PS / plane parallel) -> FO_VectorSS_spherfuncs ->
// SSV_Integral_I_UP.  ss[o] = FO_STOKES_ATMOS, db[o] = FO_STOKES_SURF.
void first_order(const Column& c, const std::vector<double>& deltau_scaled, const std::vector<double>& truncfac,
                 double ss[3], double db[3]) {
  const int L = c.L, S = c.S;
  const double vsign = 1.0;
  const double dtr = DEG_TO_RAD;
  const bool do_planpar = c.plane_parallel;
  const bool do_enhanced = !do_planpar && !c.focorr_nadir;
  double sza = c.sza, vza = c.vza, raa = c.raa;
  if (raa < ZERO) raa = -raa;
  if (raa > 360.0) raa = 360.0 - raa - 360.0;
  std::vector<double> radii(L + 2, 0.0), losw(L + 1, 0.0), alpha(L + 1, 0.0), sina(L + 1, 0.0), cosa(L + 1, 0.0);
  for (int n = 0; n <= L; ++n) radii[n] = EARTH_RADIUS + c.height[n];
  double mu0, mu1, cosscat, raycon = 0.0;
  // sunpaths(n, 1..n) and fine-layer sunpaths; ntraverse
  std::vector<double> sunp((size_t)(L + 1) * (L + 2), 0.0);
  std::vector<int> ntrav(L + 1, 0);
  std::vector<double> sunpf((size_t)(L + 1) * NFINE * (L + 2), 0.0);
  std::vector<int> ntravf((size_t)(L + 1) * NFINE, 0);
  std::vector<double> xfine((size_t)(L + 1) * NFINE, 0.0), wfine((size_t)(L + 1) * NFINE, 0.0);
  auto SP = [&](int n, int k) -> double& { return sunp[(size_t)n * (L + 2) + k]; };
  auto SPF = [&](int n, int j, int k) -> double& { return sunpf[((size_t)n * NFINE + j) * (L + 2) + k]; };
  bool donadir = false;
  if (do_planpar) {
    // ObsGeom_PlanPar_WP (FOC/FO_WPgeometry_Routines.f90:126-311)
    double alpha_r = vza * dtr, theta_r = sza * dtr;
    double calpha = (vza == 90.0) ? ZERO : std::cos(alpha_r), salpha = (vza == 90.0) ? ONE : std::sin(alpha_r);
    double stheta = std::sin(theta_r), ctheta = std::cos(theta_r);
    double cphi = std::cos(raa * dtr);
    mu0 = ctheta; mu1 = calpha;
    bool overhead = (sza == ZERO);
    double term1, term2;
    if (overhead) { term1 = ZERO; term2 = calpha; cosscat = -vsign * term2; if (term2 == ZERO) cosscat = term2; }
    else { term1 = salpha * stheta * cphi; term2 = calpha * ctheta; cosscat = -vsign * term2 + term1; }
    double utheta = ONE / ctheta;
    for (int n = 1; n <= L; ++n) {
      ntrav[n] = n;
      double dh = c.height[n - 1] - c.height[n];
      losw[n] = dh / calpha;
      for (int k = 1; k <= n; ++k) SP(n, k) = (c.height[k - 1] - c.height[k]) * utheta;
    }
  } else if (!do_enhanced) {
    // Obsgeom_RegularPS_WP (FOC/FO_WPgeometry_Routines.f90:776-991)
    double alpha_r = vza * dtr, theta_r = sza * dtr;
    double calpha = (vza == 90.0) ? ZERO : std::cos(alpha_r), salpha = (vza == 90.0) ? ONE : std::sin(alpha_r);
    double ctheta, stheta;
    if (sza == 90.0) { ctheta = ZERO; stheta = ONE; } else { stheta = std::sin(theta_r); ctheta = std::cos(theta_r); }
    double cphi = std::cos(raa * dtr);
    mu0 = ctheta; mu1 = calpha;
    bool overhead = (sza == ZERO);
    raycon = salpha * radii[L];
    double term1, term2;
    if (overhead) { term1 = ZERO; term2 = calpha; cosscat = -vsign * term2; if (term2 == ZERO) cosscat = term2; }
    else { term1 = salpha * stheta * cphi; term2 = calpha * ctheta; cosscat = -vsign * term2 + term1; }
    std::vector<double> tmp(L + 2, 0.0);
    for (int n = 1; n <= L; ++n) {
      ntrav[n] = n;
      losw[n] = (c.height[n - 1] - c.height[n]) / calpha;
      std::fill(tmp.begin(), tmp.end(), 0.0);
      find_sunpaths_d(overhead, radii[n], radii.data(), theta_r, stheta, n, tmp.data());
      for (int k = 1; k <= n; ++k) SP(n, k) = tmp[k];
    }
  } else {
    // LosOut_EnhancedPS_Initial_WP (FOC/FO_WPgeometry_Routines.f90:1511-1696)
    if (vza == ZERO) {
      donadir = true;
      for (int n = L; n >= 1; --n) losw[n] = radii[n - 1] - radii[n];
    } else {
      double alpha_boa_r = vza * dtr, salpha_boa, calpha1;
      if (vza == 90.0) { salpha_boa = ONE; calpha1 = ZERO; } else { salpha_boa = std::sin(alpha_boa_r); calpha1 = std::cos(alpha_boa_r); }
      cosa[L] = calpha1; sina[L] = salpha_boa; alpha[L] = alpha_boa_r;
      raycon = salpha_boa * radii[L];
      for (int n = L - 1; n >= 0; --n) {
        int n1 = n + 1;
        sina[n] = raycon / radii[n]; alpha[n] = std::asin(sina[n]);
        double calpha = std::cos(alpha[n]); cosa[n] = calpha;
        losw[n1] = radii[n] * calpha - radii[n1] * calpha1;
        calpha1 = calpha;
      }
    }
    // LosOut_EnhancedPS_Quadrature_WP (upwelling), :1700-1971
    std::vector<double> radiif((size_t)(L + 1) * NFINE, 0.0), alphaf((size_t)(L + 1) * NFINE, 0.0);
    double tfine[NFINE], afine[NFINE];
    for (int n = L; n >= 1; --n) {
      if (donadir) {
        double difh = radii[n - 1] - radii[n];
        getquad2(ZERO, difh, NFINE, tfine, afine);
        for (int j = 0; j < NFINE; ++j) {
          radiif[n * NFINE + j] = radii[n] + tfine[j];
          xfine[n * NFINE + j] = tfine[j]; wfine[n * NFINE + j] = afine[j];
        }
      } else {
        double rcn = radii[n] * cosa[n];
        getquad2(ZERO, losw[n], NFINE, tfine, afine);
        for (int j = 0; j < NFINE; ++j) {
          double tanfine = raycon / (rcn + tfine[j]);
          double af = std::atan(tanfine);
          double sf = std::sin(af);
          alphaf[n * NFINE + j] = af;
          radiif[n * NFINE + j] = raycon / sf;
          xfine[n * NFINE + j] = tfine[j]; wfine[n * NFINE + j] = afine[j];
        }
      }
    }
    // SolarIn_EnhancedPS_ObsGeom_SunPaths_WP, :2409-2859
    for (int n = 1; n <= L; ++n) { ntrav[n] = n; for (int j = 0; j < NFINE; ++j) ntravf[n * NFINE + j] = n; }
    bool overhead = (sza == ZERO);
    double calpha_boa, salpha_boa;
    if (vza == 90.0) { calpha_boa = ZERO; salpha_boa = ONE; }
    else { salpha_boa = std::sin(alpha[L]); calpha_boa = std::cos(alpha[L]); }
    mu1 = calpha_boa;
    double theta_boa_r = sza * dtr, ctheta_boa, stheta_boa;
    if (sza == 90.0) { ctheta_boa = ZERO; stheta_boa = ONE; } else { stheta_boa = std::sin(theta_boa_r); ctheta_boa = std::cos(theta_boa_r); }
    mu0 = ctheta_boa;
    double phi_boa_r = raa * dtr, cphi_boa = std::cos(phi_boa_r), sphi_boa = std::sin(phi_boa_r);
    double sdir[3];
    if (overhead) { sdir[0] = sdir[1] = sdir[2] = 0.0; }
    else { sdir[0] = -stheta_boa * cphi_boa * vsign; sdir[1] = -stheta_boa * sphi_boa; sdir[2] = -ctheta_boa; }
    double term1, term2;
    if (overhead) { term1 = ZERO; term2 = calpha_boa; cosscat = -vsign * term2; if (term2 == ZERO) cosscat = term2; }
    else { term1 = salpha_boa * stheta_boa * cphi_boa; term2 = calpha_boa * ctheta_boa; cosscat = -vsign * term2 + term1; }
    std::vector<double> tmp(L + 2, 0.0);
    for (int n = L; n >= 1; --n) {
      bool zerosunboa = overhead && (n == L || donadir);
      double radstart = radii[n], theta_n, stheta, ctheta; bool directsun;
      if (n == L) { theta_n = theta_boa_r; stheta = stheta_boa; ctheta = ctheta_boa; directsun = true; }
      else {
        double cum = alpha[L] - alpha[n];
        find_sun(donadir, overhead, radii[n], sdir, cum, theta_boa_r, theta_n, stheta, ctheta, directsun);
      }
      double thetaf[NFINE], sthetaf[NFINE], cthetaf[NFINE]; bool dsf[NFINE];
      for (int j = 0; j < NFINE; ++j) {
        double cum = alpha[L] - alphaf[n * NFINE + j];
        find_sun(donadir, overhead, radiif[n * NFINE + j], sdir, cum, theta_boa_r, thetaf[j], sthetaf[j], cthetaf[j], dsf[j]);
      }
      std::fill(tmp.begin(), tmp.end(), 0.0);
      if (directsun) find_sunpaths_d(zerosunboa, radstart, radii.data(), theta_n, stheta, n, tmp.data());
      else find_sunpaths_t(PIE, radstart, radii.data(), theta_n, stheta, n, L, tmp.data(), ntrav[n]);
      for (int k = 1; k <= L; ++k) SP(n, k) = tmp[k];
      for (int j = 0; j < NFINE; ++j) {
        std::fill(tmp.begin(), tmp.end(), 0.0);
        if (dsf[j]) find_sunpaths_d(zerosunboa, radiif[n * NFINE + j], radii.data(), thetaf[j], sthetaf[j], n, tmp.data());
        else find_sunpaths_t(PIE, radiif[n * NFINE + j], radii.data(), thetaf[j], sthetaf[j], n, L, tmp.data(), ntravf[n * NFINE + j]);
        for (int k = 1; k <= L; ++k) SPF(n, j, k) = tmp[k];
      }
    }
  }
  // spherical functions (FOC/FO_VectorSS_Spherfuncs.f90:93-300), sunlight, obsgeom
  const int nmom = c.ngreek;
  std::vector<double> gs1(nmom + 1), gs2(nmom + 1, 0.0);
  double rot3 = 0.0, rot4 = 0.0;
  {
    const double SMALL = 1.0e-12;
    double mus = cosscat;
    std::vector<double> gh1(nmom + 1, 0.0), gh2(nmom + 1, 0.0), gh3(nmom + 1, 0.0), gh4(nmom + 1, 0.0);
    double sql41 = 0.0;
    for (int l = 2; l <= nmom; ++l) {
      double dl = l, dl1 = l - 1, dl21 = 2 * l - 1;
      gh1[l] = dl21 / dl; gh2[l] = dl1 / dl;
      if (S > 1) {
        if (l == 2) sql41 = ZERO;
        else {
          double sql4 = sql41, dl41 = dl * dl - 4.0;
          sql41 = std::sqrt(dl41);
          gh3[l] = dl21 / sql41; gh4[l] = sql4 / sql41;
        }
      }
    }
    double qroot6 = -0.25 * std::sqrt(6.0);
    double theta_r = dtr * sza, ctheta = std::cos(theta_r), stheta = std::sin(theta_r);
    double alpha_r = dtr * vza, calpha = std::cos(alpha_r), salpha = std::sin(alpha_r);
    double phi_r = dtr * raa, cphi = std::cos(phi_r);
    double help = ONE - mus * mus, sinscat = (help <= ZERO) ? SMALL : std::sqrt(help);
    double csig1, csig2;
    if (std::fabs(sinscat) <= SMALL) { csig1 = ZERO; csig2 = ZERO; }
    else {
      csig1 = (stheta == ZERO) ? -cphi : (vsign * calpha + ctheta * mus) / sinscat / stheta;
      csig2 = (salpha == ZERO) ? -cphi : (-ctheta - vsign * calpha * mus) / sinscat / salpha;
    }
    csig1 = std::min(ONE, std::max(-ONE, csig1));
    csig2 = std::min(ONE, std::max(-ONE, csig2));
    double csig2_2 = TWO * csig2;
    double ssig2 = (std::fabs(csig2 - ONE) < SMALL) ? ZERO : std::sqrt(ONE - csig2 * csig2);
    double c2 = csig2_2 * csig2 - ONE;
    double s2 = (raa <= 180.0) ? csig2_2 * ssig2 : -csig2_2 * ssig2;
    rot3 = c2; rot4 = s2;
    if (S == 1) {
      gs1[0] = ONE; if (nmom >= 1) gs1[1] = mus;
      for (int l = 2; l <= nmom; ++l) gs1[l] = gh1[l] * gs1[l - 1] * mus - gh2[l] * gs1[l - 2];
    } else {
      double p00[2], p02[2]; int lnew = 0, lold = 1;
      for (int l = 0; l <= nmom; ++l) {
        if (l == 0) { p00[lold] = ONE; p00[lnew] = ZERO; p02[lold] = ZERO; p02[lnew] = ZERO; }
        else if (l == 1) p00[lold] = mus;
        else p00[lold] = gh1[l] * mus * p00[lnew] - gh2[l] * p00[lold];
        gs1[l] = p00[lold];
        if (l == 2) {
          double mum1 = ONE - mus, mup1 = ONE + mus;
          p02[lold] = qroot6 * mum1 * mup1; p02[lnew] = ZERO;
          gs2[l] = p02[lold];
        } else if (l > 2) {
          p02[lold] = gh3[l] * mus * p02[lnew] - gh4[l] * p02[lold];
          gs2[l] = p02[lold];
        }
        std::swap(lnew, lold);
      }
    }
  }
  // SSV_Integral_I_UP (FOC/FO_VectorSS_RTCalcs_I.f90:116-785), TOA level only
  const double cutoff = 88.0;
  std::vector<double> ext(L + 1, 0.0), tms(L + 1, 0.0);
  for (int n = 1; n <= L; ++n) {
    ext[n] = deltau_scaled[n - 1] / (c.height[n - 1] - c.height[n]);
    double help = ONE - truncfac[n - 1] * c.omega[n - 1];
    tms[n] = c.omega[n - 1] / help;
  }
  std::vector<double> exact((size_t)(L + 1) * 3, 0.0);
  for (int n = 1; n <= L; ++n) {
    double f1 = 0.0, f2 = 0.0;
    for (int l = 0; l <= nmom; ++l) {
      f1 = f1 + gs1[l] * G_IN(c, l, n - 1, 0);
      if (S > 1) f2 = f2 + gs2[l] * G_IN(c, l, n - 1, 1);
    }
    if (S == 1) exact[n * 3] = f1 * tms[n];
    else {
      exact[n * 3 + 0] = tms[n] * (+f1);
      exact[n * 3 + 1] = tms[n] * (-f2 * rot3);
      exact[n * 3 + 2] = tms[n] * (+f2 * rot4);
    }
  }
  std::vector<double> atten(L + 1, 0.0), suntau(L + 1, 0.0), attenf((size_t)(L + 1) * NFINE, 0.0);
  for (int n = 0; n <= L; ++n) {
    int nt = ntrav[n];
    double sumd = 0.0;
    for (int k = 1; k <= nt; ++k) sumd = sumd + ext[k] * SP(n, k);
    suntau[n] = sumd;
    if (sumd < cutoff) atten[n] = std::exp(-sumd);
  }
  if (do_enhanced) {
    for (int n = 1; n <= L; ++n)
      for (int j = 0; j < NFINE; ++j) {
        int nt = ntravf[n * NFINE + j];
        double sumd = 0.0;
        for (int k = 1; k <= nt; ++k) sumd = sumd + ext[k] * SPF(n, j, k);
        if (sumd < cutoff) attenf[n * NFINE + j] = std::exp(-sumd);
      }
  }
  std::vector<double> lostrans(L + 1, 0.0), mult(L + 1, 0.0);
  if (!do_enhanced) {
    for (int n = L; n >= 1; --n) {
      if (mu1 > ZERO) {
        double lostau = deltau_scaled[n - 1] / mu1;
        if (lostau < cutoff) lostrans[n] = std::exp(-lostau);
        double factor1 = atten[n - 1] - atten[n] * lostrans[n];
        double factor2 = ONE + (suntau[n] - suntau[n - 1]) / lostau;
        mult[n] = factor1 / factor2;
      }
    }
  } else {
    for (int n = L; n >= 1; --n) {
      double kn = ext[n], path_up = losw[n];
      double lostau = kn * path_up;
      if (lostau < cutoff) lostrans[n] = std::exp(-lostau);
      double sum = 0.0;
      for (int j = 0; j < NFINE; ++j) {
        double dj = losw[n] - xfine[n * NFINE + j];
        double tran = std::exp(-kn * dj);
        sum = sum + attenf[n * NFINE + j] * tran * wfine[n * NFINE + j];
      }
      mult[n] = sum * kn;
    }
  }
  const double fvec1 = 1.0;  // FLUXVEC(1)
  double cums[3] = {0, 0, 0};
  double m4 = 4.0 * mu0;
  double cdb = m4 * c.albedo * atten[L] * fvec1;
  double cdbv[3] = {cdb, 0.0, 0.0};
  for (int n = L; n >= 1; --n) {
    for (int o = 0; o < S; ++o) {
      double src = exact[n * 3 + o] * fvec1 * mult[n];
      cdbv[o] = lostrans[n] * cdbv[o];
      cums[o] = lostrans[n] * cums[o] + src;
    }
  }
  const double flux = c.flux_factor / PI4;  // SS_FLUX_MULTIPLIER (masters:1665)
  for (int o = 0; o < S; ++o) { ss[o] = flux * cums[o]; db[o] = flux * cdbv[o]; }
}

// ---------------------------------------------------------------- the path
Result run_column(const Column& c) {
  Result res{0, 0, 0, 0, 0};
  const int L = c.L, N = c.N, S = c.S;
  const int n = S * N;                 // NSTKS_NSTRMS
  const int n2 = 2 * n;                // NSTKS_NSTRMS_2
  const int NS2 = 2 * N;               // NSTREAMS_2
  // ---- input checks (REG/vlidort_inputs.f90 VLIDORT_CHECK_INPUT) that can fire
  bool bad = false;
  if (c.sza < ZERO || c.sza >= 90.0) bad = true;
  if (c.raa > 360.0 || c.raa < ZERO) bad = true;
  if (c.vza > 90.0 || c.vza < ZERO) bad = true;
  for (int i = 1; i <= L; ++i) if (c.height[i - 1] <= c.height[i]) bad = true;
  for (int l = 0; l < L; ++l) {
    if (c.omega[l] > ONE - OMEGA_SMALLNUM || c.omega[l] < OMEGA_SMALLNUM) bad = true;
    if (G_IN(c, 0, l, 0) != ONE) bad = true;
  }
  if (bad) { res.status = 3; return res; }
  // ---- derived inputs (VLIDORT_DERIVE_INPUT)
  int ngreek = c.ngreek;
  const int NMOMENTS = std::min(2 * N - 1, ngreek);
  std::vector<double> qs(N), qw(N), qsw(N), qhw(N);
  if (N == 1) { qs[0] = std::sqrt(ONE / 3.0); qw[0] = ONE; }
  else getquad2(ZERO, ONE, N, qs.data(), qw.data());
  for (int i = 0; i < N; ++i) { qsw[i] = qs[i] * qw[i]; qhw[i] = HALF * qw[i]; }
  const double csza = std::cos(c.sza * DEG_TO_RAD);
  const double ustream = std::cos(DEG_TO_RAD * c.vza);
  const double usec = ONE / ustream;
  // layer-scattering flags (DERIVE, :4136-4160)
  std::vector<char> layscat((size_t)(NMOMENTS + 1) * L, 1);
  for (int m = 3; m <= NMOMENTS; ++m) {
    int qc = NMOMENTS - m + 1;
    for (int nl = 0; nl < L; ++nl) {
      int q = 0;
      for (int l = m; l <= NMOMENTS; ++l) if (G_IN(c, l, nl, 0) == ZERO) q++;
      layscat[(size_t)m * L + nl] = (q < qc);
    }
  }
  // ---- Chapman factors (VLIDORT_CHAPMAN, non-refractive)
  A2 chap; chap.init(L, L);  // chap(n,k) 0-based, k<=n
  {
    std::vector<double> radii(L + 1), delz(L + 1);
    radii[0] = c.height[0] + EARTH_RADIUS;
    for (int nl = 1; nl <= L; ++nl) { radii[nl] = c.height[nl] + EARTH_RADIUS; delz[nl] = c.height[nl - 1] - c.height[nl]; }
    double th_toa = c.sza * DEG_TO_RAD, mu_toa = std::cos(th_toa), sn_toa = std::sqrt(ONE - mu_toa * mu_toa);
    if (c.plane_parallel) {
      for (int nl = 1; nl <= L; ++nl) for (int k = 1; k <= nl; ++k) chap(nl - 1, k - 1) = ONE / mu_toa;
    } else if (sn_toa == ZERO) {
      for (int nl = 1; nl <= L; ++nl) for (int k = 1; k <= nl; ++k) chap(nl - 1, k - 1) = ONE;
    } else {
      for (int nl = 1; nl <= L; ++nl) {
        double sinth1 = sn_toa * radii[nl] / radii[0];
        double sth1 = std::asin(sinth1);
        double re_upper = radii[0];
        for (int k = 1; k <= nl; ++k) {
          double re_lower = re_upper - delz[k];
          double sinth2 = re_upper * sinth1 / re_lower;
          double sth2 = std::asin(sinth2);
          double phi = sth2 - sth1;
          double sinphi = std::sin(phi);
          double dist = re_upper * sinphi / sinth2;
          chap(nl - 1, k - 1) = dist / delz[k];
          re_upper = re_lower; sinth1 = sinth2; sth1 = sth2;
        }
      }
    }
  }
  // ---- DELTAMSCALE (REG/vlidort_miscsetups.f90:370-758)
  std::vector<double> dtau(L), omega(L), fac1(L), trunc(L), taugrid(L + 1);
  // greek scaled: gm(l, n, k) for l<=NMOMENTS, k in 0..15
  std::vector<double> gm((size_t)(NMOMENTS + 1) * L * 16, 0.0);
  auto GM = [&](int l, int nl, int k) -> double& { return gm[((size_t)k * L + nl) * (NMOMENTS + 1) + l]; };
  {
    const int KT1[4] = {0, 5, 10, 15}, KT2[4] = {1, 4, 11, 14};
    taugrid[0] = ZERO;
    int nm1 = NMOMENTS + 1;
    double dnm1 = (double)(2 * nm1 + 1);
    for (int nl = 0; nl < L; ++nl) {
      double fdel = G_IN(c, nm1, nl, 0) / dnm1;
      double fac2 = ONE - fdel;
      fac1[nl] = ONE - fdel * c.omega[nl];
      trunc[nl] = fdel;
      for (int l = 0; l <= NMOMENTS; ++l) {
        double dnl1 = (double)(2 * l + 1), fdnl1 = fdel * dnl1;
        if (S > 1) {
          for (int k = 0; k < 4; ++k) {
            GM(l, nl, KT1[k]) = (G_IN(c, l, nl, KT1[k]) - fdnl1) / fac2;
            GM(l, nl, KT2[k]) = G_IN(c, l, nl, KT2[k]) / fac2;
          }
        } else {
          GM(l, nl, 0) = (G_IN(c, l, nl, 0) - fdnl1) / fac2;
        }
      }
      GM(0, nl, 0) = ONE;
      dtau[nl] = c.deltau[nl] * fac1[nl];
      omega[nl] = c.omega[nl] * fac2 / fac1[nl];
      taugrid[nl + 1] = taugrid[nl] + dtau[nl];
    }
    for (int nl = 0; nl < L; ++nl) {
      if (omega[nl] > 0.999999999) omega[nl] = 0.999999999;
      else if (omega[nl] < 1.0e-9) omega[nl] = 1.0e-9;
      double g1 = GM(1, nl, 0) / 3.0;
      if (g1 > 0.999999999) g1 = 0.999999999;
      else if (g1 < -0.999999999) g1 = -0.999999999;
      else if (g1 >= 0.0 && g1 < 1.0e-9) g1 = 1.0e-9;
      else if (g1 < 0.0 && g1 > -1.0e-9) g1 = -1.0e-9;
      GM(1, nl, 0) = 3.0 * g1;
    }
  }
  // DELTAU_SLANT (scaled) from Chapman factors
  A2 dslant; dslant.init(L, L);
  for (int nl = 0; nl < L; ++nl) for (int k = 0; k <= nl; ++k) dslant(nl, k) = c.deltau[k] * chap(nl, k) * fac1[k];
  // ---- SSALBINIT: omega_greek(l, n, o1, o2)
  auto MI = [](int o1, int o2) { return 4 * o1 + o2; };   // MUELLER_INDEX, 0-based
  std::vector<double> og((size_t)(NMOMENTS + 1) * L * S * S, 0.0);
  auto OG = [&](int l, int nl, int o1, int o2) -> double& { return og[(((size_t)o2 * S + o1) * L + nl) * (NMOMENTS + 1) + l]; };
  for (int nl = 0; nl < L; ++nl) for (int l = 0; l <= NMOMENTS; ++l)
    for (int o1 = 0; o1 < S; ++o1) for (int o2 = 0; o2 < S; ++o2) OG(l, nl, o1, o2) = omega[nl] * GM(l, nl, MI(o1, o2));
  // ---- QSPREP (:813-1010)
  int beam_cutoff = L;
  std::vector<double> avsec(L, 0.0), itrans(L, 0.0);
  double tau_solar, trans_solar_beam = 0.0; bool do_reflected_directbeam = false;
  if (c.plane_parallel) {
    double sec0 = ONE / csza;
    for (int nl = 1; nl <= L; ++nl) {
      double tsl = taugrid[nl] * sec0;
      if (nl <= beam_cutoff) {
        if (tsl > MAX_TAU_SPATH) beam_cutoff = nl;
        avsec[nl - 1] = sec0; itrans[nl - 1] = std::exp(-taugrid[nl - 1] * sec0);
      }
    }
    tau_solar = taugrid[L] * sec0;
  } else {
    std::vector<double> tauslant(L + 1, 0.0);
    for (int nl = 1; nl <= L; ++nl) { double tau = ZERO; for (int k = 1; k <= nl; ++k) tau += dslant(nl - 1, k - 1); tauslant[nl] = tau; }
    tau_solar = tauslant[L];
    double s_t_1 = ZERO, s_t_0 = ONE;
    for (int nl = 1; nl <= L; ++nl) {
      if (nl <= beam_cutoff) {
        if (tauslant[nl] > MAX_TAU_SPATH) beam_cutoff = nl;
        else s_t_1 = std::exp(-tauslant[nl]);
        avsec[nl - 1] = (tauslant[nl] - tauslant[nl - 1]) / dtau[nl - 1];
        itrans[nl - 1] = s_t_0;
        s_t_0 = s_t_1;
      }
    }
  }
  if (tau_solar < MAX_TAU_SPATH) { trans_solar_beam = std::exp(-tau_solar); do_reflected_directbeam = true; }
  // ---- PREPTRANS (:1014-1292)
  std::vector<double> tdmubar(L, 0.0), tduserm(L, 0.0), itransu(L, 0.0);
  for (int nl = 1; nl <= L; ++nl) {
    if (nl > beam_cutoff) tdmubar[nl - 1] = ZERO;
    else { double sp = dtau[nl - 1] * avsec[nl - 1]; tdmubar[nl - 1] = (sp > MAX_TAU_SPATH) ? ZERO : std::exp(-sp); }
    itransu[nl - 1] = itrans[nl - 1] * usec;
    double sp = dtau[nl - 1] * usec;
    tduserm[nl - 1] = (sp > MAX_TAU_UPATH) ? ZERO : std::exp(-sp);
  }
  // ---- EMULT_MASTER_OBSGEO (:3187-3421), upwelling
  std::vector<double> sigp(L), emult_up(L, 0.0);
  for (int nl = 0; nl < L; ++nl) {
    sigp[nl] = avsec[nl] + usec;
    if (nl + 1 > beam_cutoff) emult_up[nl] = ZERO;
    else {
      double wdel = tdmubar[nl], wudel = wdel * tduserm[nl];
      double su = (ONE - wudel) / sigp[nl];
      emult_up[nl] = itransu[nl] * su;
    }
  }
  // ---- FO single scatter + direct bounce
  double ss[3] = {0, 0, 0}, db[3] = {0, 0, 0};
  first_order(c, dtau, trunc, ss, db);

  // ---- Fourier loop (VLIDORT_MASTER :2304-2700)
  const int N_FOURIERS = NMOMENTS;
  double stokes[3] = {0, 0, 0};
  int testconv = 0, fourier_saved = 0;
  bool local_iteration = true;
  int fourier = -1;
  // PIMATRIX persistent state (SAVE / inout between Fourier terms)
  const int NSA = N + 2;   // quad streams, user stream, solar cosine
  std::vector<double> XM(NSA);
  for (int i = 0; i < N; ++i) XM[i] = qs[i];
  XM[N] = ustream; XM[N + 1] = csza;
  std::vector<double> pimm11(NSA, 0.0), pimmkm(NSA, 0.0);
  std::vector<double> df_l(NMOMENTS + 2), df_lp1(NMOMENTS + 2), df_2lp1(NMOMENTS + 2), df_lsqm4(NMOMENTS + 2, 0.0),
      df_rt(NMOMENTS + 2, 0.0), zhelp(NMOMENTS + 2, 0.0);
  std::vector<double> umxsq(NSA), rt_umxsq(NSA), upxsq_d(NSA), pleg20(NSA), m2x_d(NSA), x_rt(NSA);
  // BVP work
  const int NT = n2 * L;
  const int KL = (L == 1) ? 2 * n - 1 : 3 * n - 1, KU = KL;
  const int LDAB = 2 * KL + KU + 1;
  std::vector<double> ab((size_t)LDAB * NT), colv(NT);
  std::vector<int> ipiv(NT);

  while (local_iteration && fourier < N_FOURIERS) {
    fourier++;
    const int M = fourier;
    const bool do_include_surface = (M == 0) && !(c.albedo == ZERO);
    const double surface_factor = (M == 0) ? TWO : ONE;
    const double delta_factor = (M == 0) ? ONE : TWO;
    const double flux_multiplier = delta_factor;
    const bool do_localbeam = do_include_surface ? do_reflected_directbeam : false;
    // DIRECTRADIANCE (:1296-1614) -> reflected direct beam at the quadrature streams
    double rf_direct = 0.0;
    if (do_include_surface && do_reflected_directbeam) {
      double x0_flux = 4.0 * csza / delta_factor;
      x0_flux = c.flux_factor * x0_flux / PI4;
      double attn = x0_flux * trans_solar_beam;
      rf_direct = attn * c.albedo;  // Lambertian, M=0
    }
    // PIMATRIX_SETUP_OMP (:2273-2932)
    // pi(l, i, s1, s2) for i in 0..NSA-1
    std::vector<double> pi((size_t)(NMOMENTS + 1) * NSA * 16, 0.0);
    auto PI = [&](int l, int i, int s1, int s2) -> double& { return pi[(((size_t)s2 * 4 + s1) * NSA + i) * (NMOMENTS + 1) + l]; };
    std::vector<double> phi(NMOMENTS + 1), pinorm(NMOMENTS + 1);
    {
      double rc20 = 0.25 * std::sqrt(6.0), rc21 = 2.0 * rc20;
      double fac = ONE;
      for (int l = 1; l <= 2 * M; ++l) fac = fac * (double)l;
      phi[M] = ONE / fac;
      for (int l = M + 1; l <= NMOMENTS; ++l) { double f = (double)(l - M) / (double)(l + M); phi[l] = phi[l - 1] * f; }
      for (int l = M; l <= NMOMENTS; ++l) pinorm[l] = std::sqrt(phi[l]);
      if (M == 0) {
        for (int l = 0; l <= NMOMENTS; ++l) { df_l[l] = l; df_lp1[l] = l + 1; df_2lp1[l] = 2 * l + 1; }
        df_lsqm4[2] = ZERO;
        for (int l = 3; l <= NMOMENTS; ++l) df_lsqm4[l] = std::sqrt((double)(l * l - 4));
        for (int l = 2; l <= NMOMENTS; ++l) { df_rt[l] = std::sqrt((double)((l + 3) * (l - 1))); zhelp[l] = TWO * df_2lp1[l] / df_lp1[l] / df_l[l]; }
        for (int i = 0; i < NSA; ++i) {
          double xsq = XM[i] * XM[i];
          pleg20[i] = HALF * (3.0 * xsq - ONE);
          umxsq[i] = ONE - xsq;
          if (XM[i] == ONE) { rt_umxsq[i] = ZERO; x_rt[i] = ZERO; upxsq_d[i] = ZERO; m2x_d[i] = ZERO; }
          else {
            rt_umxsq[i] = std::sqrt(umxsq[i]); x_rt[i] = rt_umxsq[i] * XM[i];
            upxsq_d[i] = (ONE + xsq) / umxsq[i]; m2x_d[i] = -TWO * XM[i] / umxsq[i];
          }
        }
      }
      std::vector<double> xlm((NMOMENTS + 2) * 4, 0.0), ylm((NMOMENTS + 2) * 4, 0.0), zlm((NMOMENTS + 2) * 16, 0.0);
      auto XLM = [&](int l, int s) -> double& { return xlm[l * 4 + s]; };
      auto YLM = [&](int l, int s) -> double& { return ylm[l * 4 + s]; };
      auto ZLM = [&](int l, int s1, int s2) -> double& { return zlm[l * 16 + s1 * 4 + s2]; };
      if (M == 0) {
        for (int l = 2; l <= NMOMENTS; ++l) {
          XLM(l, 0) = df_lp1[l]; XLM(l, 1) = df_rt[l]; XLM(l, 2) = XLM(l, 1); XLM(l, 3) = XLM(l, 0);
          for (int s = 0; s < S; ++s) XLM(l, s) = ONE / XLM(l, s);
        }
        for (int l = 2; l <= NMOMENTS; ++l) { YLM(l, 0) = df_l[l]; YLM(l, 1) = df_lsqm4[l]; YLM(l, 2) = YLM(l, 1); YLM(l, 3) = YLM(l, 0); }
      } else {
        for (int l = M; l <= NMOMENTS; ++l) {
          double dlp1mm = (double)(l + 1 - M);
          XLM(l, 0) = ONE / dlp1mm;
          if (l == 1) XLM(l, 1) = ZERO;
          else { XLM(l, 1) = df_rt[l] * dlp1mm / df_lp1[l]; XLM(l, 1) = ONE / XLM(l, 1); }
          XLM(l, 2) = XLM(l, 1); XLM(l, 3) = XLM(l, 0);
        }
        for (int s = 0; s < S; ++s) YLM(M, s) = ZERO;
        for (int l = M + 1; l <= NMOMENTS; ++l) {
          double dlpm = (double)(l + M);
          YLM(l, 0) = dlpm; YLM(l, 1) = dlpm * df_lsqm4[l] / df_l[l]; YLM(l, 2) = YLM(l, 1); YLM(l, 3) = YLM(l, 0);
        }
        for (int l = 2; l <= NMOMENTS; ++l) {
          double zv = (double)M * zhelp[l];
          for (int s1 = 0; s1 < S; ++s1) for (int s2 = 0; s2 < S; ++s2) ZLM(l, s1, s2) = ZERO;
          if (S > 2) { ZLM(l, 1, 2) = zv; ZLM(l, 2, 1) = zv; }
        }
      }
      if (M == 0) {
        for (int i = 0; i < NSA; ++i) { PI(0, i, 0, 0) += ONE; PI(0, i, 3, 3) += ONE; }
        if (NMOMENTS >= 1) for (int i = 0; i < NSA; ++i) { PI(1, i, 0, 0) += XM[i]; PI(1, i, 3, 3) += XM[i]; }
        if (NMOMENTS >= 2) for (int i = 0; i < NSA; ++i) {
          double pl20 = pleg20[i], rl20 = rc20 * umxsq[i];
          PI(2, i, 0, 0) += pl20; PI(2, i, 1, 1) += rl20; PI(2, i, 2, 2) = PI(2, i, 1, 1); PI(2, i, 3, 3) = PI(2, i, 0, 0);
        }
        for (int l = 2; l <= NMOMENTS - 1; ++l)
          for (int i = 0; i < NSA; ++i) for (int s1 = 0; s1 < S; ++s1) for (int s2 = 0; s2 < S; ++s2) {
            double hx = df_2lp1[l] * XM[i] * PI(l, i, s1, s2);
            double hy = YLM(l, s1) * PI(l - 1, i, s1, s2);
            PI(l + 1, i, s1, s2) = (hx - hy) * XLM(l, s1);
          }
      } else if (M == 1) {
        for (int i = 0; i < NSA; ++i) { PI(1, i, 0, 0) += rt_umxsq[i]; PI(1, i, 3, 3) += rt_umxsq[i]; }
        if (NMOMENTS >= 2) for (int i = 0; i < NSA; ++i) {
          double pl21 = 3.0 * x_rt[i], rl21 = -rc21 * x_rt[i], tl21 = +rc21 * rt_umxsq[i];
          PI(2, i, 0, 0) += pl21; PI(2, i, 1, 1) += rl21; PI(2, i, 1, 2) += tl21;
          PI(2, i, 2, 2) = PI(2, i, 1, 1); PI(2, i, 3, 3) = PI(2, i, 0, 0); PI(2, i, 2, 1) = PI(2, i, 1, 2);
        }
        for (int l = 2; l <= NMOMENTS - 1; ++l)
          for (int i = 0; i < NSA; ++i) for (int s1 = 0; s1 < S; ++s1) for (int s2 = 0; s2 < S; ++s2) {
            double hx = df_2lp1[l] * XM[i] * PI(l, i, s1, s2);
            double hy = YLM(l, s1) * PI(l - 1, i, s1, s2);
            double hz = ZERO;
            for (int s3 = 0; s3 < S; ++s3) hz += ZLM(l, s1, s3) * PI(l, i, s3, s2);
            PI(l + 1, i, s1, s2) = (hx - hy + hz) * XLM(l, s1);
          }
      } else {
        if (M == 2) {
          for (int i = 0; i < NSA; ++i) { pimm11[i] = 3.0 * umxsq[i]; pimmkm[i] = (XM[i] == ONE) ? rc21 : rc21 * umxsq[i]; }
        } else {
          double h1 = df_2lp1[M - 1];
          double h2 = df_lp1[M - 1] * h1 / df_rt[M - 1];
          for (int i = 0; i < NSA; ++i) {
            pimm11[i] = h1 * rt_umxsq[i] * pimm11[i];
            pimmkm[i] = (XM[i] == ONE) ? ZERO : h2 * rt_umxsq[i] * pimmkm[i];
          }
        }
        for (int i = 0; i < NSA; ++i) {
          PI(M, i, 0, 0) += pimm11[i];
          if (XM[i] == ONE) { PI(M, i, 1, 1) += pimmkm[i] * TWO; PI(M, i, 1, 2) -= pimmkm[i] * TWO; }
          else { PI(M, i, 1, 1) += pimmkm[i] * upxsq_d[i]; PI(M, i, 1, 2) += pimmkm[i] * m2x_d[i]; }
          PI(M, i, 2, 2) = PI(M, i, 1, 1); PI(M, i, 3, 3) = PI(M, i, 0, 0); PI(M, i, 2, 1) = PI(M, i, 1, 2);
        }
        if (M < NMOMENTS)
          for (int i = 0; i < NSA; ++i) for (int s1 = 0; s1 < S; ++s1) for (int s2 = 0; s2 < S; ++s2) {
            double hx = df_2lp1[M] * XM[i] * PI(M, i, s1, s2);
            double hz = ZERO;
            for (int s3 = 0; s3 < S; ++s3) hz += ZLM(M, s1, s3) * PI(M, i, s3, s2);
            PI(M + 1, i, s1, s2) = (hx + hz) * XLM(M, s1);
          }
        for (int l = M + 1; l <= NMOMENTS - 1; ++l)
          for (int i = 0; i < NSA; ++i) for (int s1 = 0; s1 < S; ++s1) for (int s2 = 0; s2 < S; ++s2) {
            double hx = df_2lp1[l] * XM[i] * PI(l, i, s1, s2);
            double hy = YLM(l, s1) * PI(l - 1, i, s1, s2);
            double hz = ZERO;
            for (int s3 = 0; s3 < S; ++s3) hz += ZLM(l, s1, s3) * PI(l, i, s3, s2);
            PI(l + 1, i, s1, s2) = (hx - hy + hz) * XLM(l, s1);
          }
      }
    }
    // normalised PI products (DMAT = diag(1,1,-1,-1); PISIGN)
    const double DM[4] = {1, 1, -1, -1};
    auto pisign = [&](int l) { return ((l - M) % 2 == 0) ? ONE : -ONE; };
    // xqp(l,i,s1,s2), xqm_pre, xqm_post, xqm for quad streams; user: xup, xum, xum_post, xup_pre; x0p
    const int LM = NMOMENTS + 1;
    std::vector<double> xqp((size_t)LM * N * 16, 0), xqm_pre((size_t)LM * N * 16, 0), xqm_post((size_t)LM * N * 16, 0),
        xqm((size_t)LM * N * 16, 0);
    std::vector<double> xup((size_t)LM * 16, 0), xum((size_t)LM * 16, 0), xum_post((size_t)LM * 16, 0), xup_pre((size_t)LM * 16, 0),
        x0p((size_t)LM * 16, 0);
    auto Q4 = [&](std::vector<double>& v, int l, int i, int s1, int s2) -> double& { return v[(((size_t)s2 * 4 + s1) * N + i) * LM + l]; };
    auto U4 = [&](std::vector<double>& v, int l, int s1, int s2) -> double& { return v[((size_t)s2 * 4 + s1) * LM + l]; };
    for (int l = M; l <= NMOMENTS; ++l) {
      for (int i = 0; i < N; ++i) {
        for (int s1 = 0; s1 < S; ++s1) for (int s2 = 0; s2 < S; ++s2) Q4(xqp, l, i, s1, s2) = PI(l, i, s1, s2) * pinorm[l];
        for (int s1 = 0; s1 < S; ++s1) for (int s2 = 0; s2 < S; ++s2) {
          double h1 = DM[s1] * Q4(xqp, l, i, s1, s2);
          double h2 = Q4(xqp, l, i, s1, s2) * DM[s2];
          Q4(xqm_pre, l, i, s1, s2) = h1 * pisign(l);
          Q4(xqm_post, l, i, s1, s2) = h2 * pisign(l);
          Q4(xqm, l, i, s1, s2) = DM[s1] * Q4(xqp, l, i, s1, s2) * DM[s2] * pisign(l);
        }
      }
      for (int s1 = 0; s1 < S; ++s1) for (int s2 = 0; s2 < S; ++s2) U4(xup, l, s1, s2) = PI(l, N, s1, s2) * pinorm[l];
      for (int s1 = 0; s1 < S; ++s1) for (int s2 = 0; s2 < S; ++s2) {
        U4(xup_pre, l, s1, s2) = DM[s1] * U4(xup, l, s1, s2);
        U4(xum_post, l, s1, s2) = DM[s2] * U4(xup, l, s1, s2) * pisign(l);
        U4(xum, l, s1, s2) = DM[s1] * U4(xup, l, s1, s2) * DM[s2] * pisign(l);
        U4(x0p, l, s1, s2) = PI(l, N + 1, s1, s2) * pinorm[l];
      }
    }
    // ---- homogeneous solutions per layer
    // X+/X- : xpos(i, o, k, nl) i in 0..2N-1, o < S, k < n
    std::vector<double> xpos((size_t)NS2 * S * n * L, 0.0), xneg((size_t)NS2 * S * n * L, 0.0);
    auto XP = [&](int i, int o, int k, int nl) -> double& { return xpos[(((size_t)nl * n + k) * S + o) * NS2 + i]; };
    auto XN = [&](int i, int o, int k, int nl) -> double& { return xneg[(((size_t)nl * n + k) * S + o) * NS2 + i]; };
    std::vector<double> keig((size_t)n * L), tdel((size_t)n * L), normsv((size_t)n * L);
    std::vector<double> uhom_updn((size_t)S * n * L), uhom_upup((size_t)S * n * L), hmult1((size_t)n * L), hmult2((size_t)n * L);
    std::vector<double> sab((size_t)n * n), dab((size_t)n * n);  // SAB(I,J,O1,O2) as matrix rows I*S+O1, cols J*S+O2
    for (int nl = 0; nl < L; ++nl) {
      // VLIDORT_QHOM_SOLUTION (REG/vlidort_solutions.f90:150-1047)
      std::vector<double> h2p((size_t)S * LM * S * N, 0.0), h2m((size_t)S * LM * S * N, 0.0);
      auto H = [&](std::vector<double>& v, int o3, int l, int o2, int j) -> double& { return v[(((size_t)j * S + o2) * LM + l) * S + o3]; };
      for (int j = 0; j < N; ++j) for (int o2 = 0; o2 < S; ++o2) for (int l = M; l <= NMOMENTS; ++l) for (int o3 = 0; o3 < S; ++o3) {
        double sp = 0.0, sm = 0.0;
        for (int o = 0; o < S; ++o) { sp += OG(l, nl, o3, o) * Q4(xqp, l, j, o, o2); sm += OG(l, nl, o3, o) * Q4(xqm_pre, l, j, o, o2); }
        H(h2p, o3, l, o2, j) = sp; H(h2m, o3, l, o2, j) = sm;
      }
      for (int i = 0; i < N; ++i) {
        double xinv = -ONE / qs[i];
        for (int j = 0; j < N; ++j) {
          double fac = xinv * qhw[j];
          for (int o1 = 0; o1 < S; ++o1) for (int o2 = 0; o2 < S; ++o2) {
            double dp = ZERO, dm = ZERO;
            for (int l = M; l <= NMOMENTS; ++l) {
              double sp = 0.0, sm = 0.0;
              for (int o = 0; o < S; ++o) { sp += Q4(xqp, l, i, o1, o) * H(h2p, o, l, o2, j); sm += Q4(xqp, l, i, o1, o) * H(h2m, o, l, o2, j); }
              dp += sp; dm += sm;
            }
            sab[(size_t)(j * S + o2) * n + (i * S + o1)] = fac * (dp + dm);
            dab[(size_t)(j * S + o2) * n + (i * S + o1)] = fac * (dp - dm);
          }
        }
        for (int o1 = 0; o1 < S; ++o1) {
          sab[(size_t)(i * S + o1) * n + (i * S + o1)] -= xinv;
          dab[(size_t)(i * S + o1) * n + (i * S + o1)] -= xinv;
        }
      }
      std::vector<double> eig((size_t)n * n), eigsave((size_t)n * n), evec((size_t)n * n), eval(n);
      for (int r = 0; r < n; ++r) for (int cc = 0; cc < n; ++cc) {
        double s = 0.0;
        // sum over (J, O) in Fortran column-major order of the 2-D slice DAB(I,1:N,O1,1:S): J fastest
        for (int o = 0; o < S; ++o) for (int j = 0; j < N; ++j) s += dab[(size_t)(j * S + o) * n + r] * sab[(size_t)cc * n + (j * S + o)];
        eig[(size_t)cc * n + r] = s;
      }
      eigsave = eig;
      bool fail = false;
      int ier = asymtx(ASYMTX_TOL, eig.data(), n, evec.data(), eval.data(), &fail);
      if (fail) { res.status = 4; return res; }
      if (ier > 0) {
        eig = eigsave; std::fill(evec.begin(), evec.end(), 0.0);
        ier = asymtx(1.0e-10, eig.data(), n, evec.data(), eval.data(), &fail);
        if (ier > 0) {
          eig = eigsave; std::fill(evec.begin(), evec.end(), 0.0);
          ier = asymtx(1.0e-8, eig.data(), n, evec.data(), eval.data(), &fail);
          if (ier > 0) { res.status = 4; return res; }
        }
      }
      for (int aa = 0; aa < n; ++aa) {
        double s = 0.0; for (int r = 0; r < n; ++r) s += evec[(size_t)aa * n + r] * evec[(size_t)aa * n + r];
        double nr = std::sqrt(s);
        for (int r = 0; r < n; ++r) evec[(size_t)aa * n + r] /= nr;
      }
      for (int k = 0; k < n; ++k) {
        double kv = std::sqrt(eval[k]);
        keig[(size_t)nl * n + k] = kv;
        double sc_r = ONE / kv;
        double s = 0.0; for (int r = 0; r < n; ++r) s += evec[(size_t)k * n + r] * evec[(size_t)k * n + r];
        double norm_r = std::sqrt(s);
        std::vector<double> sum(n), dif(n);
        for (int r = 0; r < n; ++r) sum[r] = evec[(size_t)k * n + r] / norm_r;
        for (int i = 0; i < N; ++i) for (int o1 = 0; o1 < S; ++o1) {
          double h = 0.0;
          for (int o = 0; o < S; ++o) for (int j = 0; j < N; ++j) h += sab[(size_t)(j * S + o) * n + (i * S + o1)] * sum[j * S + o];
          dif[i * S + o1] = h * sc_r;
        }
        for (int i = 0; i < N; ++i) for (int o = 0; o < S; ++o) {
          double fpd = HALF * (sum[i * S + o] + dif[i * S + o]);
          double fnd = HALF * (sum[i * S + o] - dif[i * S + o]);
          XP(i, o, k, nl) = fpd; XN(i, o, k, nl) = fnd;
          if (o < 2) { XP(i + N, o, k, nl) = fnd; XN(i + N, o, k, nl) = fpd; }
          else { XP(i + N, o, k, nl) = -fnd; XN(i + N, o, k, nl) = -fpd; }
        }
        double help = keig[(size_t)nl * n + k] * dtau[nl];
        tdel[(size_t)nl * n + k] = (help > MAX_TAU_QPATH) ? ZERO : std::exp(-help);
      }
      // VLIDORT_UHOM_SOLUTION (:1167-1505), upwelling, one user stream
      bool lsc = (M < 3) ? true : layscat[(size_t)M * L + nl];
      for (int k = 0; k < n; ++k) {
        double rho_p = usec + keig[(size_t)nl * n + k], rho_m = usec - keig[(size_t)nl * n + k];
        double zeta_p = ONE / rho_p;
        bool hsingo = std::fabs(rho_m) < TAYLOR_SMALL;
        double zeta_m = hsingo ? rho_m : ONE / rho_m;
        if (!lsc) {
          for (int o = 0; o < S; ++o) { uhom_updn[((size_t)nl * n + k) * S + o] = 0; uhom_upup[((size_t)nl * n + k) * S + o] = 0; }
        } else {
          std::vector<double> gaux((size_t)LM * S, 0.0);
          for (int l = M; l <= NMOMENTS; ++l) {
            double hs[3];
            for (int o1 = 0; o1 < S; ++o1) {
              double sps = ZERO;
              for (int j = 0; j < N; ++j) {
                double a5 = qhw[j], sp = 0.0, sm = 0.0;
                for (int o = 0; o < S; ++o) { sp += Q4(xqp, l, j, o1, o) * a5 * XP(j, o, k, nl); sm += Q4(xqm_pre, l, j, o1, o) * a5 * XN(j, o, k, nl); }
                sps = sps + sp + sm;
              }
              hs[o1] = sps;
            }
            for (int o1 = 0; o1 < S; ++o1) { double s = 0.0; for (int o = 0; o < S; ++o) s += OG(l, nl, o1, o) * hs[o]; gaux[l * S + o1] = s; }
          }
          for (int o1 = 0; o1 < S; ++o1) {
            double sdn = 0.0, sup = 0.0;
            // sum over (l, o) with l fastest
            for (int o = 0; o < S; ++o) for (int l = M; l <= NMOMENTS; ++l) {
              sdn += U4(xum, l, o1, o) * gaux[l * S + o];
              sup += U4(xup_pre, l, o1, o) * gaux[l * S + o];
            }
            uhom_updn[((size_t)nl * n + k) * S + o1] = sdn;
            uhom_upup[((size_t)nl * n + k) * S + o1] = sup;
          }
        }
        // WHOLELAYER_HMULT (:1651-1790)
        double udel = tduserm[nl], sm = usec;
        double zdel = tdel[(size_t)nl * n + k], zudel = zdel * udel;
        double th2 = ONE - zudel, th1 = zdel - udel;
        if (hsingo) hmult1[(size_t)nl * n + k] = taylor_series_1(TAYLOR_ORDER, zeta_m, dtau[nl], udel, sm);
        else hmult1[(size_t)nl * n + k] = sm * th1 * zeta_m;
        hmult2[(size_t)nl * n + k] = sm * th2 * zeta_p;
      }
    }
    // VLIDORT_QHOM_NORMS (:1051-1163)
    for (int nl = 0; nl < L; ++nl) for (int k = 0; k < n; ++k) {
      double norm = ZERO;
      for (int j = 0; j < N; ++j) {
        double t1 = 0.0, t2 = 0.0;
        for (int o = 0; o < S; ++o) { t1 += XP(j, o, k, nl) * XP(j, o, k, nl); t2 += XN(j, o, k, nl) * XN(j, o, k, nl); }
        norm = norm + qsw[j] * (t1 - t2);
      }
      normsv[(size_t)nl * n + k] = norm;
    }
    // ---- Green's-function beam solutions (VLIDORT_GBEAM_SOLUTION, :2649-2932)
    std::vector<double> wupper((size_t)NS2 * S * L, 0.0), wlower((size_t)NS2 * S * L, 0.0);
    auto WU = [&](int i, int o, int nl) -> double& { return wupper[((size_t)nl * S + o) * NS2 + i]; };
    auto WL = [&](int i, int o, int nl) -> double& { return wlower[((size_t)nl * S + o) * NS2 + i]; };
    std::vector<double> aterm((size_t)n * L, 0), bterm((size_t)n * L, 0), gamp((size_t)n * L, 0), gamm((size_t)n * L, 0);
    const double F1 = c.flux_factor / PI4;
    for (int nl = 0; nl < L; ++nl) {
      bool lsc = (M < 3) ? true : layscat[(size_t)M * L + nl];
      if (nl + 1 > beam_cutoff || !lsc) continue;   // WUPPER/WLOWER zero
      double secbar = avsec[nl], cst = itrans[nl], wdel = tdmubar[nl];
      std::vector<double> cfunc(n), dfunc(n);
      for (int k = 0; k < n; ++k) {
        double zdel = tdel[(size_t)nl * n + k], zwdel = zdel * wdel;
        gamp[(size_t)nl * n + k] = secbar + keig[(size_t)nl * n + k];
        gamm[(size_t)nl * n + k] = secbar - keig[(size_t)nl * n + k];
        double eps = gamm[(size_t)nl * n + k];
        if (std::fabs(eps) < TAYLOR_SMALL) cfunc[k] = taylor_series_1(TAYLOR_ORDER, eps, dtau[nl], wdel, cst);
        else cfunc[k] = cst * (zdel - wdel) / gamm[(size_t)nl * n + k];
        dfunc[k] = cst * (ONE - zwdel) / gamp[(size_t)nl * n + k];
      }
      std::vector<double> hq((size_t)LM * S, 0.0);
      for (int o1 = 0; o1 < S; ++o1) for (int l = M; l <= NMOMENTS; ++l) {
        double gsum = ZERO;
        for (int o2 = 0; o2 < S; ++o2) {
          double s = 0.0;
          for (int o = 0; o < S; ++o) s += U4(x0p, l, o2, o) * (o == 0 ? ONE : ZERO);   // DFLUX = (1,0,0,0)
          gsum = gsum + OG(l, nl, o1, o2) * s;
        }
        hq[l * S + o1] = gsum;
      }
      std::vector<double> dpi((size_t)N * S), dmi((size_t)N * S);
      for (int o1 = 0; o1 < S; ++o1) for (int i = 0; i < N; ++i) {
        double sp = 0.0, sm = 0.0;
        for (int o = 0; o < S; ++o) for (int l = M; l <= NMOMENTS; ++l) {
          sp += Q4(xqp, l, i, o1, o) * hq[l * S + o];
          sm += Q4(xqm_post, l, i, o1, o) * hq[l * S + o];
        }
        dpi[i * S + o1] = F1 * sp; dmi[i * S + o1] = F1 * sm;
      }
      std::vector<double> gup(n), gdn(n);
      for (int k = 0; k < n; ++k) {
        double a = 0.0, b = 0.0;
        for (int i = 0; i < N; ++i) {
          double h1 = 0.0, h2 = 0.0;
          for (int o = 0; o < S; ++o) h1 += dpi[i * S + o] * XP(i, o, k, nl);
          double h1b = 0.0; for (int o = 0; o < S; ++o) h1b += dmi[i * S + o] * XN(i, o, k, nl);
          h1 = h1 + h1b;
          for (int o = 0; o < S; ++o) h2 += dmi[i * S + o] * XP(i, o, k, nl);
          double h2b = 0.0; for (int o = 0; o < S; ++o) h2b += dpi[i * S + o] * XN(i, o, k, nl);
          h2 = h2 + h2b;
          a += qw[i] * h1; b += qw[i] * h2;
        }
        aterm[(size_t)nl * n + k] = a / normsv[(size_t)nl * n + k];
        bterm[(size_t)nl * n + k] = b / normsv[(size_t)nl * n + k];
        gdn[k] = cfunc[k] * aterm[(size_t)nl * n + k];
        gup[k] = dfunc[k] * bterm[(size_t)nl * n + k];
      }
      for (int o1 = 0; o1 < S; ++o1) for (int i = 0; i < N; ++i) {
        double wu = 0.0, wu1 = 0.0, wl1 = 0.0, wl = 0.0;
        for (int k = 0; k < n; ++k) {
          wu += gup[k] * XN(i, o1, k, nl); wu1 += gup[k] * XP(i, o1, k, nl);
          wl1 += gdn[k] * XN(i, o1, k, nl); wl += gdn[k] * XP(i, o1, k, nl);
        }
        WU(i, o1, nl) = wu; WU(i + N, o1, nl) = wu1; WL(i + N, o1, nl) = wl1; WL(i, o1, nl) = wl;
        if (o1 > 1) { WL(i + N, o1, nl) = -WL(i + N, o1, nl); WU(i + N, o1, nl) = -WU(i + N, o1, nl); }
      }
    }
    // ---- BVP (BVP_MATRIXSETUP_MASTER + BVP_SOLUTION_MASTER), band storage
    std::vector<double> lcon((size_t)n * L), mcon((size_t)n * L);
    {   // NLAYERS = 1: the band covers the whole 2n x 2n matrix (KL = KU = 2n-1), and
        // DGBTF2 picks the pivots of DGETRF's partial pivoting
      std::fill(ab.begin(), ab.end(), 0.0);
      const int KV = KU + KL;
      auto AB = [&](int r, int cidx) -> double& {  // 1-based matrix (r,c) -> band storage
        return ab[(size_t)(cidx - 1) * LDAB + (KV + r - cidx)];
      };
      // surface reflection of homogeneous solutions (BVP_SURFACE_SETUP_HOM)
      std::vector<double> r2p(n, 0.0), r2m(n, 0.0);
      if (do_include_surface) {
        double kmult = surface_factor * c.albedo;
        for (int k = 0; k < n; ++k) {
          double sp = 0.0, sm = 0.0;
          for (int j = 0; j < N; ++j) { sp += qsw[j] * XP(j, 0, k, L - 1); sm += qsw[j] * XN(j, 0, k, L - 1); }
          r2p[k] = sp * kmult; r2m[k] = sm * kmult;
        }
      }
      // TOA rows
      for (int i = 0; i < N; ++i) for (int o1 = 0; o1 < S; ++o1) {
        int irow = S * i + o1 + 1;
        for (int ep = 1; ep <= n; ++ep) {
          AB(irow, ep) = XP(i, o1, ep - 1, 0);
          AB(irow, ep + n) = XN(i, o1, ep - 1, 0) * tdel[ep - 1];
        }
      }
      int c0 = -n;
      for (int nl = 2; nl <= L; ++nl) {
        int n1 = nl - 1;
        c0 = c0 + n2;
        int ce = c0 - n;
        for (int i = 0; i < NS2; ++i) for (int o1 = 0; o1 < S; ++o1) {
          int irow = S * i + o1 + 1, cm = c0 + irow;
          for (int ep = 1; ep <= n; ++ep) {
            int cep = ce + ep, cem = cep + n;
            AB(cm, cep) = tdel[(size_t)(n1 - 1) * n + ep - 1] * XP(i, o1, ep - 1, n1 - 1);
            AB(cm, cem) = XN(i, o1, ep - 1, n1 - 1);
          }
          for (int ep = 1; ep <= n; ++ep) {
            int cep1 = ce + ep + n2, cem1 = cep1 + n;
            AB(cm, cep1) = -XP(i, o1, ep - 1, nl - 1);
            AB(cm, cem1) = -tdel[(size_t)(nl - 1) * n + ep - 1] * XN(i, o1, ep - 1, nl - 1);
          }
        }
      }
      c0 = c0 + n2;
      {
        int ce = c0 - n;
        for (int i = 0; i < N; ++i) for (int o1 = 0; o1 < S; ++o1) {
          int irow = S * i + o1 + 1, cp = c0 + irow;
          for (int ep = 1; ep <= n; ++ep) {
            double xpnet = XP(i + N, o1, ep - 1, L - 1) - (o1 == 0 ? r2p[ep - 1] : 0.0);
            double xmnet = XN(i + N, o1, ep - 1, L - 1) - (o1 == 0 ? r2m[ep - 1] : 0.0);
            AB(cp, ce + ep) = tdel[(size_t)(L - 1) * n + ep - 1] * xpnet;
            AB(cp, ce + ep + n) = xmnet;
          }
        }
      }
      // DGBTF2 (REG/lapack_tools.f90:156-357)
      {
        int ju = 1;
        for (int j = 1; j <= NT; ++j) {
          if (j + KV <= NT) for (int i = 1; i <= KL; ++i) ab[(size_t)(j + KV - 1) * LDAB + (i - 1)] = 0.0;
          int km = std::min(KL, NT - j);
          // IDAMAX over AB(KV+1 .. KV+1+km, j)
          int jp = 1; double dmax = std::fabs(ab[(size_t)(j - 1) * LDAB + KV]);
          for (int t = 2; t <= km + 1; ++t) { double v = std::fabs(ab[(size_t)(j - 1) * LDAB + KV + t - 1]); if (v > dmax) { dmax = v; jp = t; } }
          ipiv[j - 1] = jp + j - 1;
          if (ab[(size_t)(j - 1) * LDAB + KV + jp - 1] != ZERO) {
            ju = std::max(ju, std::min(j + KU + jp - 1, NT));
            if (jp != 1) {
              // DSWAP(JU-J+1, AB(KV+JP,J), LDAB-1, AB(KV+1,J), LDAB-1)
              for (int t = 0; t <= ju - j; ++t) {
                size_t a1 = (size_t)(j - 1 + t) * LDAB + (KV + jp - 1 - t);
                size_t a2 = (size_t)(j - 1 + t) * LDAB + (KV - t);
                std::swap(ab[a1], ab[a2]);
              }
            }
            if (km > 0) {
              double rp = ONE / ab[(size_t)(j - 1) * LDAB + KV];
              for (int t = 1; t <= km; ++t) ab[(size_t)(j - 1) * LDAB + KV + t] *= rp;
              if (ju > j) {
                // DGER(KM, JU-J, -ONE, AB(KV+2,J),1, AB(KV,J+1),LDAB-1, AB(KV+1,J+1),LDAB-1)
                for (int cc = 1; cc <= ju - j; ++cc) {
                  double y = ab[(size_t)(j - 1 + cc) * LDAB + (KV - cc)];
                  if (y != 0.0) {
                    double tmp = -ONE * y;
                    for (int t = 1; t <= km; ++t)
                      ab[(size_t)(j - 1 + cc) * LDAB + (KV - cc + t)] += ab[(size_t)(j - 1) * LDAB + KV + t] * tmp;
                  }
                }
              }
            }
          } else { res.status = 4; return res; }
        }
      }
      // RHS (BVP_SURFACE_SETUP_BEAM + BVP_COLUMN_SETUP)
      std::fill(colv.begin(), colv.end(), 0.0);
      for (int i = 0; i < N; ++i) for (int o1 = 0; o1 < S; ++o1) colv[S * i + o1] = -WU(i, o1, 0);
      for (int lay = 2; lay <= L; ++lay) {
        int cc0 = (lay - 1) * n2 - n;
        for (int i = 0; i < NS2; ++i) for (int o1 = 0; o1 < S; ++o1)
          colv[cc0 + S * i + o1] = WU(i, o1, lay - 1) - WL(i, o1, lay - 2);
      }
      {
        int cc0 = (L - 1) * n2 + n;
        double r2beam = 0.0;
        if (do_include_surface) {
          double rb = 0.0; for (int j = 0; j < N; ++j) rb += qsw[j] * WL(j, 0, L - 1);
          r2beam = rb * surface_factor * c.albedo;
        }
        for (int i = 0; i < N; ++i) for (int o1 = 0; o1 < S; ++o1) {
          double v = -WL(i + N, o1, L - 1);
          if (do_include_surface) v += (o1 == 0 ? r2beam : 0.0);
          colv[cc0 + S * i + o1] = v;
        }
        if (do_localbeam && do_include_surface)
          for (int i = 0; i < N; ++i) colv[cc0 + S * i + 0] += rf_direct;
      }
      // DGBTRS (REG/lapack_tools.f90:7354), 'N', one RHS
      {
        const int KD = KU + KL + 1;
        if (KL > 0) {
          for (int j = 1; j <= NT - 1; ++j) {
            int lm = std::min(KL, NT - j);
            int l = ipiv[j - 1];
            if (l != j) std::swap(colv[l - 1], colv[j - 1]);
            double bj = colv[j - 1];
            for (int t = 1; t <= lm; ++t) colv[j - 1 + t] += -bj * ab[(size_t)(j - 1) * LDAB + (KD - 1 + t)];
          }
        }
        // DTBSV upper, no-transpose, non-unit, bandwidth KL+KU
        for (int j = NT; j >= 1; --j) {
          if (colv[j - 1] != ZERO) {
            int l = KD - j;   // 1-based offset: AB(L+I, J)
            colv[j - 1] = colv[j - 1] / ab[(size_t)(j - 1) * LDAB + (KD - 1)];
            double temp = colv[j - 1];
            for (int i = j - 1; i >= std::max(1, j - (KL + KU)); --i)
              colv[i - 1] = colv[i - 1] - temp * ab[(size_t)(j - 1) * LDAB + (l + i - 1)];
          }
        }
      }
      for (int lay = 1; lay <= L; ++lay) {
        int cc0 = (lay - 1) * n2;
        for (int k = 0; k < n; ++k) { lcon[(size_t)(lay - 1) * n + k] = colv[cc0 + k]; mcon[(size_t)(lay - 1) * n + k] = colv[cc0 + k + n]; }
      }
    }
    // ---- VLIDORT_UPUSER_INTENSITY -> TOA (REG/vlidort_intensity.f90:145-643)
    double cum[3] = {0, 0, 0};
    if (do_include_surface && M == 0) {   // GET_BOA_SOURCE Lambertian
      double kmult = surface_factor * c.albedo;
      double refl = 0.0;
      for (int i = 0; i < N; ++i) {
        double spar = WL(i, 0, L - 1), shom = ZERO;
        for (int k = 0; k < n; ++k) {
          double lxr = lcon[(size_t)(L - 1) * n + k] * XP(i, 0, k, L - 1);
          double mxr = mcon[(size_t)(L - 1) * n + k] * XN(i, 0, k, L - 1);
          shom = shom + lxr * tdel[(size_t)(L - 1) * n + k] + mxr;
        }
        refl += qsw[i] * (spar + shom);
      }
      cum[0] = kmult * refl;
    }
    for (int nl = L; nl >= 1; --nl) {
      // WHOLELAYER_STERM_UP (REG/vlidort_PostProcessing.f90:139-451)
      double src[3] = {0, 0, 0};
      bool lsc = (M < 3) ? true : layscat[(size_t)M * L + nl - 1];
      if (lsc) {
        for (int o = 0; o < S; ++o) {
          double h = ZERO;
          for (int k = 0; k < n; ++k) {
            size_t ik = (size_t)(nl - 1) * n + k;
            double luxr = lcon[ik] * uhom_updn[ik * S + o];
            double muxr = mcon[ik] * uhom_upup[ik * S + o];
            h = h + luxr * hmult2[ik] + muxr * hmult1[ik];
          }
          src[o] = h;
        }
        if (nl <= beam_cutoff) {
          double wdel = tdmubar[nl - 1], it = itrans[nl - 1], itw = -it * wdel;
          double spar[3] = {0, 0, 0};
          for (int k = 0; k < n; ++k) {
            size_t ik = (size_t)(nl - 1) * n + k;
            double sd;
            if (std::fabs(gamm[ik]) < TAYLOR_SMALL) {
              double mult = taylor_series_2(TAYLOR_ORDER, gamm[ik], sigp[nl - 1], dtau[nl - 1], ONE, wdel * tduserm[nl - 1], ONE);
              sd = itransu[nl - 1] * mult;
            } else sd = (it * hmult2[ik] - emult_up[nl - 1]) / gamm[ik];
            double su = (itw * hmult1[ik] + emult_up[nl - 1]) / gamp[ik];
            for (int o = 0; o < S; ++o)
              spar[o] = spar[o] + uhom_updn[ik * S + o] * sd * aterm[ik] + uhom_upup[ik * S + o] * su * bterm[ik];
          }
          for (int o = 0; o < S; ++o) src[o] += spar[o];
        }
      }
      for (int o = 0; o < S; ++o) cum[o] = src[o] + tduserm[nl - 1] * cum[o];
    }
    double stokes_f[3] = {0, 0, 0};
    for (int o = 0; o < S; ++o) stokes_f[o] = flux_multiplier * cum[o];
    // ---- VLIDORT_CONVERGE_OBSGEO (REG/vlidort_converge.f90:825-1231)
    if (M == 0) {
      for (int o = 0; o < S; ++o) stokes[o] = stokes_f[o];
      for (int o = 0; o < S; ++o) stokes[o] = stokes[o] + ss[o];
      for (int o = 0; o < S; ++o) stokes[o] = stokes[o] + db[o];
    } else {
      double dfc = (double)M;
      double azm = c.raa * dfc * DEG_TO_RAD;
      double azmfac[3] = {std::cos(azm), std::cos(azm), std::sin(azm)};
      double told[3], tazm[3], tnew[3];
      int count = 0;
      for (int o = 0; o < S; ++o) { told[o] = stokes[o]; tazm[o] = azmfac[o] * stokes_f[o]; tnew[o] = told[o] + tazm[o]; }
      if (tazm[0] != ZERO) { double accur = std::fabs(tazm[0] / tnew[0]); if (accur < VLIDORT_ACCURACY) count++; }
      else count++;
      for (int o = 0; o < S; ++o) stokes[o] = tnew[o];
      if (count == 1) {
        testconv++;
        if (testconv == 2) local_iteration = false;
        if (!local_iteration) fourier_saved = M;
      } else { testconv = 0; fourier_saved = 2 * N - 1; }
    }
  }
  res.I = stokes[0]; res.Q = (S > 1) ? stokes[1] : 0.0; res.U = (S > 2) ? stokes[2] : 0.0;
  res.fourier_saved = fourier_saved;
  return res;
}

bool is_missing(double x, double missing) { return std::fabs(x / missing - 1) < 0.001; }

}  // namespace

// ------------------------------------------------------------------ C-ABI of the oracle
// Same argument layout as VLIDORT_Scalar_Lambert / VLIDORT_Vector_Lambert
// (src/Components/rtms/geo_vlidort/VLIDORT_LAMBERT.F90:34-342), Fortran column-major arrays,
// plus nfourier(nobs,nch) = Fourier terms executed.
extern "C" {

static void lambert_common(bool vector, int km, int nch, int nobs, const double* channels, int nstreams, int plane_parallel,
                           int nmom, int npol, const double* rot, const double* depol, const double* alpha,
                           const double* tau, const double* ssa, const double* pmom, const double* pe, const double* he,
                           const double* te, const double* albedo, const double* sza, const double* raa,
                           const double* vza, const double* flux_factor, double missing, double* radiance,
                           double* reflectance, double* qout, double* uout, int* nfourier, int* rc) {
  (void)channels; (void)pe; (void)te;
  *rc = 0;
  if (nstreams > 25) { *rc = 2; return; }
  if (km > 72) { *rc = 3; return; }
  const int S = vector ? 3 : 1;
  const float pi_f = 4.0f * std::atan(1.0f);
  const double pi = (double)pi_f;                     // VLIDORT_ScatMod.F90:1036 pi = 4.*atan(1.0)
  bool latch_nadir = false;                           // SZA==0 latch (VLIDORT_ScatMod.F90:1123-1126)
  std::vector<double> greek((size_t)(nmom + 1) * km * 16, 0.0), dtau(km), om(km), hgt(km + 1);
  for (int j = 0; j < nobs; ++j) {
    for (int i = 0; i < nch; ++i) nfourier[i * nobs + j] = 0;
    if (is_missing(sza[j], missing) || is_missing(vza[j], missing) || is_missing(raa[j], missing)) {
      for (int i = 0; i < nch; ++i) {
        radiance[i * nobs + j] = missing; reflectance[i * nobs + j] = missing;
        if (vector) { qout[i * nobs + j] = missing; uout[i * nobs + j] = missing; }
      }
      continue;
    }
    for (int i = 0; i < nch; ++i) {
      size_t oi = (size_t)i * nobs + j;
      if (is_missing(albedo[oi], missing)) {
        radiance[oi] = missing; reflectance[oi] = missing;
        if (vector) { qout[oi] = missing; uout[oi] = missing; }
        continue;
      }
      if (sza[j] == 0.0) latch_nadir = true;
      const double d = depol[i];
      const double raysmom2 = (1.0 - d) / (2.0 + d);
      const double sqrt6f = (double)std::sqrt(6.0f);
      const double gammamom2 = -(sqrt6f * (1 - d) / (2 + d));
      const double alphamom2 = 6 * (1 - d) / (2 + d);
      const double deltamom1 = 3 * (1 - 2 * d) / (2 + d);
      std::fill(greek.begin(), greek.end(), 0.0);
      auto GR = [&](int l, int n, int k) -> double& { return greek[((size_t)k * km + n) * (nmom + 1) + l]; };
      auto PM = [&](int n, int l, int p) -> double { // pmom(km,nch,nobs,nMom,nPol): (n, i, j, l, p)
        return pmom[((((size_t)p * nmom + l) * nobs + j) * nch + i) * km + n];
      };
      for (int n = 0; n < km; ++n) {
        size_t ti = ((size_t)j * nch + i) * km + n;   // tau(km,nch,nobs)
        double ray_l = rot[((size_t)i * nobs + j) * km + n];     // ROT(km,nobs,nch)
        double alpha_l = vector ? alpha[((size_t)i * nobs + j) * km + n] : 0.0;
        double tau_l = tau[ti], ssa_l = ssa[ti];
        double tau_ext = alpha_l + ray_l + tau_l;
        double tau_scat = ray_l + ssa_l * tau_l;
        double ssa_tot = tau_scat / tau_ext;
        if (ssa_tot > (double)0.99999f) ssa_tot = (double)0.99999f;
        dtau[n] = tau_ext; om[n] = ssa_tot;
        double aerswt = ssa_l * tau_l / tau_scat;
        double rayswt = ray_l / tau_scat;
        if (!vector) {
          GR(0, n, 0) = 1.0;
          double a1 = (1 <= nmom - 1) ? PM(n, 1, 0) : 0.0;
          double a2 = (2 <= nmom - 1) ? PM(n, 2, 0) : 0.0;
          GR(1, n, 0) = a1 * aerswt;
          GR(2, n, 0) = raysmom2 * rayswt + a2 * aerswt;
          for (int l = 3; l <= nmom; ++l) GR(l, n, 0) = ((l <= nmom - 1) ? PM(n, l, 0) : 0.0) * aerswt;
        } else {
          // aerosol Greek map (VLIDORT_ScatMod.F90:521-530), Rayleigh (:456-485)
          double ray[3][16] = {{0}};
          ray[0][0] = 1.0; ray[2][0] = raysmom2;
          ray[2][1] = gammamom2; ray[2][4] = gammamom2; ray[2][5] = alphamom2; ray[1][15] = deltamom1;
          auto aer = [&](int l, int k) -> double {
            if (l > nmom - 1) return 0.0;
            switch (k) {
              case 0: return PM(n, l, 0);
              case 1: return PM(n, l, 1);
              case 4: return PM(n, l, 1);
              case 10: return PM(n, l, 2);
              case 11: return PM(n, l, 3);
              case 14: return -PM(n, l, 3);
              case 5: return PM(n, l, 4);
              case 15: return PM(n, l, 5);
              default: return 0.0;
            }
          };
          for (int k = 0; k < 16; ++k) {
            for (int l = 0; l <= 2; ++l) GR(l, n, k) = ray[l][k] * rayswt + aer(l, k) * aerswt;
            for (int l = 3; l <= nmom; ++l) GR(l, n, k) = aer(l, k) * aerswt;
          }
          GR(0, n, 0) = 1.0;
        }
      }
      for (int n = 0; n <= km; ++n) hgt[n] = he[(size_t)j * (km + 1) + n] * (double)1.e-3f;
      Column c;
      c.L = km; c.N = nstreams; c.S = S; c.ngreek = nmom;
      c.deltau = dtau.data(); c.omega = om.data(); c.greek = greek.data(); c.height = hgt.data();
      c.sza = sza[j]; c.vza = vza[j]; c.raa = raa[j]; c.albedo = albedo[oi];
      c.flux_factor = vector ? flux_factor[(size_t)j * nch + i] : 1.0;
      c.plane_parallel = plane_parallel != 0; c.focorr_nadir = latch_nadir;
      Result r = run_column(c);
      if (r.status != 0) {
        radiance[oi] = missing; reflectance[oi] = missing;
        if (vector) { qout[oi] = missing; uout[oi] = missing; }
        continue;
      }
      radiance[oi] = r.I;
      reflectance[oi] = (pi * r.I) / (std::cos(sza[j] * pi / (double)180.0f) * c.flux_factor);
      if (vector) { qout[oi] = r.Q; uout[oi] = r.U; }
      nfourier[oi] = r.fourier_saved + 1;
    }
  }
}

void oracle_scalar_lambert(int km, int nch, int nobs, const double* channels, int nstreams, int plane_parallel, int nmom,
                           int npol, const double* rot, const double* depol, const double* tau, const double* ssa,
                           const double* g, const double* pmom, const double* pe, const double* he, const double* te,
                           const double* albedo, const double* sza, const double* raa, const double* vza, double missing,
                           double* radiance, double* reflectance, int* nfourier, int* rc) {
  (void)g;
  lambert_common(false, km, nch, nobs, channels, nstreams, plane_parallel, nmom, npol, rot, depol, nullptr, tau, ssa, pmom,
                 pe, he, te, albedo, sza, raa, vza, nullptr, missing, radiance, reflectance, nullptr, nullptr, nfourier, rc);
}

void oracle_vector_lambert(int km, int nch, int nobs, const double* channels, int nstreams, int plane_parallel, int nmom,
                           int npol, const double* rot, const double* depol, const double* alpha, const double* tau,
                           const double* ssa, const double* pmom, const double* pe, const double* he, const double* te,
                           const double* albedo, const double* sza, const double* raa, const double* vza,
                           const double* flux_factor, double missing, double* radiance, double* reflectance, double* q,
                           double* u, int* nfourier, int* rc) {
  lambert_common(true, km, nch, nobs, channels, nstreams, plane_parallel, nmom, npol, rot, depol, alpha, tau, ssa, pmom,
                 pe, he, te, albedo, sza, raa, vza, flux_factor, missing, radiance, reflectance, q, u, nfourier, rc);
}

}  // extern "C"
