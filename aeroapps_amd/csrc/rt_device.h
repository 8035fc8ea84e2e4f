// Device building blocks of the discrete-ordinate radiance kernels.
//
// Everything here is FP64 and follows the reference routines' steps, so
// results agree with the reference to rounding level (the Fourier-term
// kernels that include this header are compiled with FMA contraction, the
// column set-up without).  Array indices are compile-time wherever possible
// so that the per-thread matrices live in VGPRs.
#pragma once
#include "aa_common.h"

namespace aa {

// Fortran SIGN(a,b)
__device__ __forceinline__ double fsign(double a, double b) { return b >= 0.0 ? fabs(a) : -fabs(a); }

// Taylor_series_1 (REG/vlidort_Taylor.f90:119-133)
__device__ __forceinline__ double taylor1(double eps, double delta, double udel, double sm) {
  double d[TAYLOR_ORDER + 2];
  d[0] = 1.0;
#pragma unroll
  for (int m = 1; m <= TAYLOR_ORDER + 1; ++m) d[m] = delta * d[m - 1] / (double)m;
  double mult = d[1], power = 1.0;
#pragma unroll
  for (int m = 2; m <= TAYLOR_ORDER + 1; ++m) { power = power * eps; mult = mult + power * d[m]; }
  return mult * udel * sm;
}

// Taylor_series_2 (REG/vlidort_Taylor.f90:137-210)
__device__ __forceinline__ double taylor2(double eps, double y, double delta, double fac1, double fac2, double sm) {
  constexpr int MT = TAYLOR_ORDER + 2;
  double d[MT + 1], ac[MT + 1], cc[MT + 1];
  const bool small = fabs(y) < TAYLOR_SMALL;
  const int mterms = small ? TAYLOR_ORDER + 2 : TAYLOR_ORDER + 1;
  d[0] = 1.0;
#pragma unroll
  for (int m = 1; m <= MT; ++m) d[m] = (m <= mterms) ? delta * d[m - 1] / (double)m : 0.0;
  if (small) {
    double power = 1.0, power2 = 1.0, mult = d[2];
#pragma unroll
    for (int m = 3; m <= MT; ++m) {
      if (m > mterms) break;
      power = power * (eps - y);
      power2 = power - y * power2;
      mult = mult + d[m] * power2;
    }
    return mult * fac1 * sm;
  }
  const double y1 = 1.0 / y;
  ac[0] = 1.0;
#pragma unroll
  for (int m = 1; m <= MT; ++m) ac[m] = y1 * ac[m - 1];
  cc[0] = 1.0;
#pragma unroll
  for (int m = 1; m <= MT; ++m) {
    double c = 0.0;
#pragma unroll
    for (int j = 0; j <= m; ++j) c = c + ac[j] * d[m - j];
    cc[m] = c;
  }
  double power = 1.0, mult = fac1 * ac[1] - fac2 * cc[1];
#pragma unroll
  for (int m = 2; m <= MT; ++m) {
    if (m > mterms) break;
    power = eps * power;
    mult = mult + power * (fac1 * ac[m] - fac2 * cc[m]);
  }
  return mult * sm * y1;
}

// ---------------------------------------------------------------------------
// ASYMTX (REG/vlidort_aux.f90:386-959) for an M x M matrix held in registers.
//
// Same steps as the reference routine (a Householder reflection's f / h is
// formed once per row or column and the balancing scale's exact reciprocal
// once per row, where the reference divides per element: rounding level);
// the control flow is restated with
// compile-time array indices: loops whose bounds depend on the deflation state
// run over all static positions under a predicate.  The row/column isolation
// prepass of the reference only triggers for matrices with an all-zero
// off-diagonal row or column; such matrices return *fail = 2 (the caller
// reports the element as a calculation failure).  A(i,j) = a[i][j] 0-based.
// Returns IER (> 0: eigenvalue IER did not converge in 30 iterations).
template <int M>
__device__ int asymtx_dev(double tol, double (&a)[M][M], double (&evec)[M][M], double (&eval)[M], int* fail) {
  constexpr double C1 = 0.4375, C2 = 0.5, C3 = 0.75, C4 = 0.95, C5 = 16.0, C6 = 256.0;
  *fail = 0;
  double wk[M];       // WKD(1..M): balancing scale factors
  double hv[M];       // WKD(M+1..2M): Householder vector
  if constexpr (M == 1) {
    eval[0] = a[0][0]; evec[0][0] = 1.0; return 0;
  } else if constexpr (M == 2) {
    double discri = (a[0][0] - a[1][1]) * (a[0][0] - a[1][1]) + 4.0 * a[0][1] * a[1][0];
    if (discri < 0.0) { *fail = 1; return 0; }
    double sgn = 1.0; if (a[0][0] < a[1][1]) sgn = -1.0;
    eval[0] = 0.5 * (a[0][0] + a[1][1] + sgn * sqrt(discri));
    eval[1] = 0.5 * (a[0][0] + a[1][1] - sgn * sqrt(discri));
    evec[0][0] = 1.0; evec[1][1] = 1.0;
    if (a[0][0] == a[1][1] && (a[1][0] == 0.0 || a[0][1] == 0.0)) {
      double rnorm = fabs(a[0][0]) + fabs(a[0][1]) + fabs(a[1][0]) + fabs(a[1][1]);
      double w = tol * rnorm;
      evec[1][0] = a[1][0] / w; evec[0][1] = -a[0][1] / w;
    } else {
      evec[1][0] = a[1][0] / (eval[0] - a[1][1]);
      evec[0][1] = a[0][1] / (eval[1] - a[0][0]);
    }
    return 0;
  } else {
#pragma unroll
    for (int i = 0; i < M; ++i) {
      eval[i] = 0.0;
#pragma unroll
      for (int j = 0; j < M; ++j) evec[i][j] = (i == j) ? 1.0 : 0.0;
    }
    // Isolation prepass.  A diagonal matrix isolates every row: the reference
    // then returns the diagonal as eigenvalues and the identity as
    // eigenvectors.  Partial isolation (some but not all rows) is reported as
    // *fail = 2.
    {
      bool iso = false, alldiag = true;
#pragma unroll
      for (int j = 0; j < M; ++j) {
        double row = 0.0, col = 0.0;
#pragma unroll
        for (int i = 0; i < M; ++i) if (i != j) { row += fabs(a[j][i]); col += fabs(a[i][j]); }
        if (row == 0.0 || col == 0.0) iso = true;
        if (row != 0.0 || col != 0.0) alldiag = false;
      }
      if (alldiag) {
#pragma unroll
        for (int i = 0; i < M; ++i) eval[i] = a[i][i];
        return 0;
      }
      if (iso) { *fail = 2; return 0; }
    }
    // balance (L=1, K=M)
#pragma unroll
    for (int i = 0; i < M; ++i) wk[i] = 1.0;
    for (int guard = 0; guard < 1000; ++guard) {
      bool noconv = false;
#pragma unroll
      for (int i = 0; i < M; ++i) {
        double col = 0.0, row = 0.0;
#pragma unroll
        for (int j = 0; j < M; ++j) if (j != i) { col += fabs(a[j][i]); row += fabs(a[i][j]); }
        double f = 1.0, g = row / C5, h = col + row;
        // (the caps only matter for non-finite input, which the reference would not finish either)
        for (int it = 0; it < 256 && col < g; ++it) { f = f * C5; col = col * C6; }
        g = row * C5;
        // f is a power of 16: dividing by it is multiplying by its exact reciprocal
        for (int it = 0; it < 256 && col >= g; ++it) { f = f * (1.0 / C5); col = col * (1.0 / C6); }
        const double rf = 1.0 / f;
        if ((col + row) * rf < C4 * h) {
          wk[i] = wk[i] * f; noconv = true;
#pragma unroll
          for (int j = 0; j < M; ++j) a[i][j] = a[i][j] * rf;
#pragma unroll
          for (int j = 0; j < M; ++j) a[j][i] = a[j][i] * f;
        }
      }
      if (!noconv) break;
    }
    // Householder reduction to upper Hessenberg form (n = L+1 .. K-1, 1-based)
#pragma unroll
    for (int i = 0; i < M; ++i) hv[i] = 0.0;
#pragma unroll
    for (int n = 1; n <= M - 2; ++n) {       // 0-based n == 1-based n-... (1-based n = n+1)
      double h = 0.0;
      hv[n] = 0.0;
      double scale = 0.0;
#pragma unroll
      for (int i = n; i < M; ++i) scale += fabs(a[i][n - 1]);
      if (scale != 0.0) {
#pragma unroll
        for (int i = M - 1; i >= n; --i) { hv[i] = a[i][n - 1] / scale; h += hv[i] * hv[i]; }
        double g = -fsign(sqrt(h), hv[n]);
        h = h - hv[n] * g;
        hv[n] = hv[n] - g;
        // (hv f / h as hv (f / h): one division per row or column of the
        // reflection instead of one per element; rounding level)
#pragma unroll
        for (int j = n; j < M; ++j) {
          double f = 0.0;
#pragma unroll
          for (int i = M - 1; i >= n; --i) f += hv[i] * a[i][j];
          const double fh = f / h;
#pragma unroll
          for (int i = n; i < M; ++i) a[i][j] = a[i][j] - hv[i] * fh;
        }
#pragma unroll
        for (int i = 0; i < M; ++i) {
          double f = 0.0;
#pragma unroll
          for (int j = M - 1; j >= n; --j) f += hv[j] * a[i][j];
          const double fh = f / h;
#pragma unroll
          for (int j = n; j < M; ++j) a[i][j] = a[i][j] - hv[j] * fh;
        }
        hv[n] = scale * hv[n];
        a[n][n - 1] = scale * g;
      }
    }
    // accumulate the transformations (n = K-2 .. L, 1-based)
#pragma unroll
    for (int n = M - 3; n >= 0; --n) {
      const int n1 = n + 1;
      double f = a[n1][n];
      if (f != 0.0) {
        f = f * hv[n1];
        double hw[M];
#pragma unroll
        for (int i = 0; i < M; ++i) hw[i] = hv[i];
#pragma unroll
        for (int i = n + 2; i < M; ++i) hw[i] = a[i][n];
#pragma unroll
        for (int j = 0; j < M; ++j) {
          double g = 0.0;
#pragma unroll
          for (int i = n1; i < M; ++i) g += hw[i] * evec[i][j];
          g = g / f;
#pragma unroll
          for (int i = n1; i < M; ++i) evec[i][j] = evec[i][j] + g * hw[i];
        }
#pragma unroll
        for (int i = n + 2; i < M; ++i) hv[i] = hw[i];
      }
    }
    // norm of the Hessenberg matrix
    double rnorm = 0.0;
    {
      int nn = 0;
#pragma unroll
      for (int i = 0; i < M; ++i) {
#pragma unroll
        for (int j = 0; j < M; ++j) if (j >= nn) rnorm += fabs(a[i][j]);
        nn = i;
      }
    }
    // shifted double-QR iterations; nA = active end (0-based), t = shift sum
    int nA = M - 1;
    double t = 0.0;
    int ier = 0;
    int in = 0;
    bool newn = true;
    for (int sweep = 0; sweep < 64 * M && nA >= 0; ++sweep) {
      if (newn) { in = 0; newn = false; }
      // deflation point lb: largest c <= nA with negligible subdiagonal
      int lb = 0;
#pragma unroll
      for (int c = 1; c < M; ++c) {
        if (c <= nA) {
          double s = fabs(a[c - 1][c - 1]) + fabs(a[c][c]);
          if (s == 0.0) s = rnorm;
          if (fabs(a[c][c - 1]) <= tol * s) lb = c;
        }
      }
      double x = 0.0, y = 0.0, w = 0.0;
#pragma unroll
      for (int c = 0; c < M; ++c) if (c == nA) x = a[c][c];
      if (lb == nA) {   // one eigenvalue
#pragma unroll
        for (int c = 0; c < M; ++c) if (c == nA) { a[c][c] = x + t; eval[c] = a[c][c]; }
        nA = nA - 1; newn = true;
        continue;
      }
#pragma unroll
      for (int c = 1; c < M; ++c) if (c == nA) { y = a[c - 1][c - 1]; w = a[c][c - 1] * a[c - 1][c]; }
      if (lb == nA - 1) {   // two eigenvalues
#pragma unroll
        for (int c = 1; c < M; ++c) {
          if (c != nA) continue;
          const int n1 = c - 1;
          double p = (y - x) * C2, q = p * p + w;
          double z = sqrt(fabs(q));
          a[c][c] = x + t; x = a[c][c];
          a[n1][n1] = y + t;
          z = p + fsign(z, p);
          eval[n1] = x + z; eval[c] = eval[n1];
          if (z != 0.0) eval[c] = x - w / z;
          x = a[c][n1];
          double r = sqrt(x * x + z * z);
          p = x / r; q = z / r;
#pragma unroll
          for (int j = n1; j < M; ++j) { z = a[n1][j]; a[n1][j] = q * z + p * a[c][j]; a[c][j] = q * a[c][j] - p * z; }
#pragma unroll
          for (int i = 0; i <= c; ++i) { z = a[i][n1]; a[i][n1] = q * z + p * a[i][c]; a[i][c] = q * a[i][c] - p * z; }
#pragma unroll
          for (int i = 0; i < M; ++i) { z = evec[i][n1]; evec[i][n1] = q * z + p * evec[i][c]; evec[i][c] = q * evec[i][c] - p * z; }
        }
        nA = nA - 2; newn = true;
        continue;
      }
      if (in == 30) { ier = nA + 1; break; }
      if (in == 10 || in == 20) {   // exceptional shift
        t = t + x;
#pragma unroll
        for (int i = 0; i < M; ++i) if (i <= nA) a[i][i] = a[i][i] - x;
        double s = 0.0;
#pragma unroll
        for (int c = 2; c < M; ++c) if (c == nA) s = fabs(a[c][c - 1]) + fabs(a[c - 1][c - 2]);
        x = C3 * s; y = x; w = -C1 * s * s;
      }
      in = in + 1;
      // start of the double-shift sweep: scan i = nA-2 down to lb
      int i0 = lb;
      double p = 0.0, q = 0.0, r = 0.0;
      bool found = false;
#pragma unroll
      for (int i = M - 3; i >= 0; --i) {
        if (found || i > nA - 2 || i < lb) continue;
        double z = a[i][i];
        double rr = x - z, ss = y - z;
        double pp = (rr * ss - w) / a[i + 1][i] + a[i][i + 1];
        double qq = a[i + 1][i + 1] - z - rr - ss;
        double r3 = a[i + 2][i + 1];
        double s = fabs(pp) + fabs(qq) + fabs(r3);
        pp = pp / s; qq = qq / s; r3 = r3 / s;
        bool stop = (i == lb);
        if (!stop) {
          double uu = fabs(a[i][i - 1 < 0 ? 0 : i - 1]) * (fabs(qq) + fabs(r3));
          double vv = fabs(pp) * (fabs(a[i - 1 < 0 ? 0 : i - 1][i - 1 < 0 ? 0 : i - 1]) + fabs(z) + fabs(a[i + 1][i + 1]));
          if (uu <= tol * vv) stop = true;
        }
        if (stop) { found = true; i0 = i; p = pp; q = qq; r = r3; }
      }
#pragma unroll
      for (int c = 0; c < M - 2; ++c) if (c == i0) a[c + 2][c] = 0.0;
#pragma unroll
      for (int j = 3; j < M; ++j) if (j >= i0 + 3 && j <= nA) { a[j][j - 2] = 0.0; a[j][j - 3] = 0.0; }
      double s = 0.0;
#pragma unroll
      for (int ka = 0; ka < M - 1; ++ka) {
        if (ka < i0 || ka > nA - 1) continue;
        const bool notlas = (ka != nA - 1);
        if (ka == i0) {
          s = fsign(sqrt(p * p + q * q + r * r), p);
          if (lb != i0 && ka >= 1) a[ka][ka - 1] = -a[ka][ka - 1];
        } else {
          p = a[ka][ka - 1]; q = a[ka + 1][ka - 1]; r = 0.0;
          if (notlas && ka + 2 < M) r = a[ka + 2][ka - 1];
          x = fabs(p) + fabs(q) + fabs(r);
          if (x == 0.0) continue;
          p = p / x; q = q / x; r = r / x;
          s = fsign(sqrt(p * p + q * q + r * r), p);
          a[ka][ka - 1] = -s * x;
        }
        p = p + s; x = p / s; y = q / s; double z = r / s; q = q / p; r = r / p;
#pragma unroll
        for (int j = ka; j < M; ++j) {
          double pp = a[ka][j] + q * a[ka + 1][j];
          if (notlas && ka + 2 < M) { pp = pp + r * a[ka + 2][j]; a[ka + 2][j] = a[ka + 2][j] - pp * z; }
          a[ka + 1][j] = a[ka + 1][j] - pp * y;
          a[ka][j] = a[ka][j] - pp * x;
        }
#pragma unroll
        for (int ii = 0; ii < M; ++ii) {
          if (ii > nA || ii > ka + 3) continue;
          double pp = x * a[ii][ka] + y * a[ii][ka + 1];
          if (notlas && ka + 2 < M) { pp = pp + z * a[ii][ka + 2]; a[ii][ka + 2] = a[ii][ka + 2] - pp * r; }
          a[ii][ka + 1] = a[ii][ka + 1] - pp * q;
          a[ii][ka] = a[ii][ka] - pp;
        }
#pragma unroll
        for (int ii = 0; ii < M; ++ii) {
          double pp = x * evec[ii][ka] + y * evec[ii][ka + 1];
          if (notlas && ka + 2 < M) { pp = pp + z * evec[ii][ka + 2]; evec[ii][ka + 2] = evec[ii][ka + 2] - pp * r; }
          evec[ii][ka + 1] = evec[ii][ka + 1] - pp * q;
          evec[ii][ka] = evec[ii][ka] - pp;
        }
      }
    }
    if (!ier && nA >= 0) ier = nA + 1;
    if (ier) return ier;
    // eigenvectors of the triangular form by back substitution
    if (rnorm != 0.0) {
#pragma unroll
      for (int nn = M - 1; nn >= 0; --nn) {
        a[nn][nn] = 1.0;
#pragma unroll
        for (int ii = nn - 1; ii >= 0; --ii) {
          double ww = a[ii][ii] - eval[nn];
          if (ww == 0.0) ww = tol * rnorm;
          double rr = a[ii][nn];
#pragma unroll
          for (int j = ii + 1; j <= nn - 1; ++j) rr = rr + a[ii][j] * a[j][nn];
          a[ii][nn] = -rr / ww;
        }
      }
#pragma unroll
      for (int j = M - 1; j >= 0; --j)
#pragma unroll
        for (int ii = 0; ii < M; ++ii) {
          double zz = 0.0;
#pragma unroll
          for (int nn = 0; nn <= j; ++nn) zz = zz + evec[ii][nn] * a[nn][j];
          evec[ii][j] = zz;
        }
    }
#pragma unroll
    for (int ii = 0; ii < M; ++ii)
#pragma unroll
      for (int j = 0; j < M; ++j) evec[ii][j] = evec[ii][j] * wk[ii];
    return 0;
  }
}


// ---------------------------------------------------------------------------
// Gauss-Legendre nodes/weights on [a,b] (REG/vlidort_aux.f90:960-1000,
// FOC/FO_geometry_Generic.f90:331-385).
__host__ __device__ inline void getquad2(double a, double b, int n, double* roots, double* wgts) {
  const double QUARTER = 0.25, QEPS = 1.0e-13;
  const int n2 = (n + 1) / 2;
  const double nr = (double)n;
  const double midpt = 0.5 * (b + a), sfac = 0.5 * (b - a);
  for (int m = 1; m <= n2; ++m) {
    double mr = (double)m;
    double xx = PIE * (mr - QUARTER) / (nr + 0.5);
    double sx = sin(xx);
    double x = (1.0 - (nr - 1.0) / (8.0 * nr * nr * nr) -
                1.0 / (384.0 * nr * nr * nr * nr) * (39.0 - 28.0 / (sx * sx))) * cos(xx);
    double lp = 1.0, lpm1 = 0.0, dlp = 1.0;
    for (int it = 0; it < 100; ++it) {
      lpm1 = 0.0; lp = 1.0;
      for (int i = 1; i <= n; ++i) {
        double ir = (double)i, lpm2 = lpm1; lpm1 = lp;
        lp = ((2.0 * ir - 1.0) * x * lpm1 - (ir - 1.0) * lpm2) / ir;
      }
      dlp = nr * (x * lp - lpm1) / (x * x - 1.0);
      double xold = x; x = xold - lp / dlp;
      if (fabs(x - xold) <= QEPS) break;
    }
    roots[m - 1] = midpt - sfac * x;
    wgts[m - 1] = (2.0 * sfac) / ((1.0 - x * x) * dlp * dlp);
    roots[n - m] = midpt + sfac * x;
    wgts[n - m] = wgts[m - 1];
  }
}

// ---------------------------------------------------------------------------
// PI matrices of one cosine x at Fourier term M, PIMATRIX_SETUP_OMP
// (REG/vlidort_miscsetups.f90:2273-2932).  Output pi[l][s1][s2] scaled by the
// normalisation sqrt((l-M)!/(l+M)!) for l in [M, NMOM]; zero elsewhere.  The
// reference keeps PIMM_11/PIMM_KM between Fourier terms (it is called for
// M = 0, 1, 2, ... in order); here that chain is replayed from M = 2 so each
// call is self-contained and performs the same multiplications.
template <int S, int LM>
__host__ __device__ inline void pimatrix_x(int M, int NMOM, double x, double (&pi)[LM][S][S]) {
  const double rc20 = 0.25 * sqrt(6.0), rc21 = 2.0 * rc20;
#pragma unroll
  for (int l = 0; l < LM; ++l)
#pragma unroll
    for (int a = 0; a < S; ++a)
#pragma unroll
      for (int b = 0; b < S; ++b) pi[l][a][b] = 0.0;
  const double xsq = x * x;
  const double pleg20 = 0.5 * (3.0 * xsq - 1.0);
  const double umxsq = 1.0 - xsq;
  double rt_umxsq, x_rt, upxsq_d, m2x_d;
  if (x == 1.0) { rt_umxsq = 0.0; x_rt = 0.0; upxsq_d = 0.0; m2x_d = 0.0; }
  else { rt_umxsq = sqrt(umxsq); x_rt = rt_umxsq * x; upxsq_d = (1.0 + xsq) / umxsq; m2x_d = -2.0 * x / umxsq; }
  auto df_rt = [](int l) { return l >= 2 ? sqrt((double)((l + 3) * (l - 1))) : 0.0; };
  auto df_lsqm4 = [](int l) { return l >= 3 ? sqrt((double)(l * l - 4)) : 0.0; };
  auto zhelp = [](int l) { return 2.0 * (double)(2 * l + 1) / (double)(l + 1) / (double)l; };
  // XLM / YLM / ZLM coefficients of the recurrence at degree l
  auto xlm = [&](int l, int s) -> double {
    if (M == 0) return (s == 0 || s == 3) ? 1.0 / (double)(l + 1) : 1.0 / df_rt(l);
    double dlp1mm = (double)(l + 1 - M);
    if (s == 0 || s == 3) return 1.0 / dlp1mm;
    if (l == 1) return 0.0;
    double t = df_rt(l) * dlp1mm / (double)(l + 1);
    return 1.0 / t;
  };
  auto ylm = [&](int l, int s) -> double {
    if (M == 0) return (s == 0 || s == 3) ? (double)l : df_lsqm4(l);
    if (l == M) return 0.0;
    double dlpm = (double)(l + M);
    return (s == 0 || s == 3) ? dlpm : dlpm * df_lsqm4(l) / (double)l;
  };
  auto zlm = [&](int l, int s1, int s3) -> double {
    if (M == 0 || S <= 2) return 0.0;
    return ((s1 == 1 && s3 == 2) || (s1 == 2 && s3 == 1)) ? (double)M * zhelp(l) : 0.0;
  };
  int lstart;   // first degree l whose PI(l+1) comes from the three-term recurrence
  if (M == 0) {
    pi[0][0][0] += 1.0;
    if (NMOM >= 1) pi[1][0][0] += x;
    if (NMOM >= 2 && LM > 2) {
      pi[2][0][0] += pleg20;
      if (S > 1) { pi[2 % LM][1 % S][1 % S] += rc20 * umxsq; }
      if (S > 2) pi[2 % LM][2 % S][2 % S] = pi[2 % LM][1 % S][1 % S];
    }
    lstart = 2;
  } else if (M == 1) {
    pi[1][0][0] += rt_umxsq;
    if (NMOM >= 2 && LM > 2) {
      pi[2 % LM][0][0] += 3.0 * x_rt;
      if (S > 1) pi[2 % LM][1 % S][1 % S] += -rc21 * x_rt;
      if (S > 2) {
        pi[2 % LM][1 % S][2 % S] += rc21 * rt_umxsq;
        pi[2 % LM][2 % S][2 % S] = pi[2 % LM][1 % S][1 % S];
        pi[2 % LM][2 % S][1 % S] = pi[2 % LM][1 % S][2 % S];
      }
    }
    lstart = 2;
  } else {
    double pimm11 = 3.0 * umxsq;
    double pimmkm = (x == 1.0) ? rc21 : rc21 * umxsq;
    for (int m = 3; m <= M; ++m) {
      double h1 = (double)(2 * (m - 1) + 1);
      double h2 = (double)m * h1 / df_rt(m - 1);
      pimm11 = h1 * rt_umxsq * pimm11;
      pimmkm = (x == 1.0) ? 0.0 : h2 * rt_umxsq * pimmkm;
    }
    double p11 = 0.0, p22 = 0.0, p23 = 0.0;
    p11 += pimm11;
    if (x == 1.0) { p22 += pimmkm * 2.0; p23 -= pimmkm * 2.0; }
    else { p22 += pimmkm * upxsq_d; p23 += pimmkm * m2x_d; }
#pragma unroll
    for (int l = 0; l < LM; ++l) {
      if (l != M) continue;
      pi[l][0][0] = p11;
      if (S > 1) pi[l][1 % S][1 % S] = p22;
      if (S > 2) { pi[l][1 % S][2 % S] = p23; pi[l][2 % S][2 % S] = p22; pi[l][2 % S][1 % S] = p23; }
    }
    // first step PI(M+1) has no YLM term
#pragma unroll
    for (int l = 0; l < LM - 1; ++l) {
      if (l != M || M >= NMOM) continue;
#pragma unroll
      for (int s1 = 0; s1 < S; ++s1)
#pragma unroll
        for (int s2 = 0; s2 < S; ++s2) {
          double hx = (double)(2 * M + 1) * x * pi[l][s1][s2];
          double hz = 0.0;
#pragma unroll
          for (int s3 = 0; s3 < S; ++s3) hz += zlm(M, s1, s3) * pi[l][s3][s2];
          pi[l + 1][s1][s2] = (hx + hz) * xlm(M, s1);
        }
    }
    lstart = M + 1;
  }
#pragma unroll
  for (int l = 1; l < LM - 1; ++l) {
    if (l < lstart || l > NMOM - 1) continue;
#pragma unroll
    for (int s1 = 0; s1 < S; ++s1)
#pragma unroll
      for (int s2 = 0; s2 < S; ++s2) {
        double hx = (double)(2 * l + 1) * x * pi[l][s1][s2];
        double hy = ylm(l, s1) * pi[l - 1][s1][s2];
        if (M == 0) {
          pi[l + 1][s1][s2] = (hx - hy) * xlm(l, s1);
        } else {
          double hz = 0.0;
#pragma unroll
          for (int s3 = 0; s3 < S; ++s3) hz += zlm(l, s1, s3) * pi[l][s3][s2];
          pi[l + 1][s1][s2] = (hx - hy + hz) * xlm(l, s1);
        }
      }
  }
  // normalisation
  double fac = 1.0;
  for (int l = 1; l <= 2 * M; ++l) fac = fac * (double)l;
  double phi = 1.0 / fac;
#pragma unroll
  for (int l = 0; l < LM; ++l) {
    if (l < M || l > NMOM) {
#pragma unroll
      for (int a = 0; a < S; ++a)
#pragma unroll
        for (int b = 0; b < S; ++b) pi[l][a][b] = 0.0;
      continue;
    }
    if (l > M) { double f = (double)(l - M) / (double)(l + M); phi = phi * f; }
    double pn = sqrt(phi);
#pragma unroll
    for (int a = 0; a < S; ++a)
#pragma unroll
      for (int b = 0; b < S; ++b) pi[l][a][b] = pi[l][a][b] * pn;
  }
}

}  // namespace aa
