// MODIS land-surface BRDF kernels (Lambertian + Ross-thick + Li-sparse) as the
// VBRDF supplement evaluates them, shared by the supplement kernel (brdf.hip)
// and the SZA==0 latch (rt_setup.hip), which both need the exact direct bounce.
//
// Reference (VB = src/Components/rtms/RTSI/VLIDORT2p8p3/vsup/vbrdf):
//   ROSSTHICK_VFUNCTION  VB/vbrdf_sup_kernels.f90:287-332
//   LISPARSE_VFUNCTION   VB/vbrdf_sup_kernels.f90:478-582
//   direct bounce        VB/vbrdf_sup_masters.f90 (DBOUNCE_BRDFUNC accumulated
//                        over the kernels in order, factor * kernel)
// Arguments: (xi, sxi) incident cosine/sine, (xj, sxj) reflected, phi relative
// azimuth in radians with cphi = cos(phi), sphi = sin(phi).  The arithmetic is
// the reference's, operation for operation, with contraction off.
#pragma once
#include "aa_common.h"

namespace aa {

__host__ __device__ inline double ross_thick(double xi, double sxi, double xj, double sxj, double cphi) {
#pragma clang fp contract(off)
  const double pio2 = 0.5 * PIE, pio4 = 0.25 * PIE;
  const double ckphi = -cphi;
  const double ds1 = xi * xj, ds2 = sxi * sxj, ds3 = xi + xj;
  double cksi = ds1 + ds2 * ckphi;
  if (cksi > 1.0) cksi = 1.0;
  const double sksi = sqrt(1.0 - cksi * cksi);
  const double ksi = acos(cksi);
  const double func = ((pio2 - ksi) * cksi + sksi) / ds3;
  return func - pio4;
}

// Li-sparse angle transform of one direction: tan of the equivalent angle
// for the crown shape b/r, its square, cosine and sine.
struct LiAng {
  double t, t2, x, sx;
};
__host__ __device__ inline LiAng li_angle(double br, double x, double sx) {
#pragma clang fp contract(off)
  LiAng a;
  const double tx = sx / x;
  a.t = br * tx;
  a.t2 = a.t * a.t;
  const double ang = atan(a.t);
  a.x = cos(ang);
  a.sx = sin(ang);
  return a;
}

__host__ __device__ inline double li_sparse(double hb, const LiAng& in, const LiAng& rf, double cphi, double sphi) {
#pragma clang fp contract(off)
  const double ckphi = -cphi;
  const double cksi = in.x * rf.x + in.sx * rf.sx * ckphi;
  const double p = (1.0 + cksi) / rf.x / in.x;
  const double a = 1.0 / in.x, b = 1.0 / rf.x;
  const double r = a + b;
  const double dt1 = rf.t2 + in.t2;
  const double dt2 = in.t * rf.t;
  const double dt2sq = dt2 * dt2;
  const double delta = sqrt(dt1 - 2.0 * dt2 * ckphi);
  const double dsq = delta * delta;
  const double h = sqrt(dsq + sphi * sphi * dt2sq);
  const double cost = hb * h / r;
  double q;
  if (cost > 1.0) {
    q = 1.0;
  } else {
    const double t = acos(cost);
    const double sint = sqrt(1.0 - cost * cost);
    const double sintcost = sint * cost;
    q = 1.0 - ((t - sintcost) / PIE);
  }
  const double qr = q * r;
  return 0.5 * p - qr;
}

// Exact direct-bounce BRDF of the MODIS kernel set for one geometry (degrees):
// 0 + fiso*1, then + fvol*Ross-thick, then + fgeo*Li-sparse(h/b, b/r).
__host__ __device__ inline double modis_db(double fiso, double fgeo, double fvol, double hb, double br, double sza,
                                           double vza, double raa) {
#pragma clang fp contract(off)
  const double mu0 = cos(sza * DEG_TO_RAD);
  const double s0 = sqrt(1.0 - mu0 * mu0);
  const double muv = cos(vza * DEG_TO_RAD);
  const double sv = sqrt(1.0 - muv * muv);
  const double phi = raa * DEG_TO_RAD;
  const double cphi = cos(phi), sphi = sin(phi);
  double db = fiso * 1.0;
  db = db + fvol * ross_thick(mu0, s0, muv, sv, cphi);
  db = db + fgeo * li_sparse(hb, li_angle(br, mu0, s0), li_angle(br, muv, sv), cphi, sphi);
  return db;
}

// ---- Cox-Munk glint (COXMUNK_VFUNCTION, VB/vbrdf_sup_kernels.f90:1018-1141)
// with the supplement's erfc (VB/vbrdf_sup_aux.f90:266-286).  pars: sigma^2,
// m^2; the shadowing factor is split into its per-direction terms, which the
// reference computes with the same operations in every call.
__host__ __device__ inline double derfc_e(double x) {
#pragma clang fp contract(off)
  const double z = fabs(x);
  const double t = 1.0 / (1.0 + 0.5 * z);
  double r = t * exp(-z * z - 1.26551223 +
                     t * (1.00002368 + t * (.37409196 + t * (.09678418 + t * (-.18628806 + t * (.27886807 + t *
                     (-1.13520398 + t * (1.48851587 + t * (-.82215223 + t * .17087277)))))))));
  if (x < 0.0) r = 2.0 - r;
  return r;
}

// Shadowing term of one direction cosine x.
__host__ __device__ inline double cm_shadow(double sigma2, double x) {
#pragma clang fp contract(off)
  const double s1 = sqrt(sigma2 / PIE);
  const double s3 = 1.0 / (sqrt(sigma2));
  const double s2 = s3 * s3;
  const double xx = x * x;
  const double dcot = x / sqrt(1.0 - xx);
  const double t1 = exp(-dcot * dcot * s2);
  const double t2 = derfc_e(dcot * s3);
  return 0.5 * (s1 * t1 / dcot - t2);
}

// Kernel value with the shadowing terms shi (incident) and shr (reflected).
__host__ __device__ inline double cox_munk(double sigma2, double m2, double xi, double sxi, double xj, double sxj,
                                           double ckphi, double shi, double shr) {
#pragma clang fp contract(off)
  const double ckphi_neg = -ckphi;
  double z = xi * xj + sxi * sxj * ckphi_neg;
  if (z > 1.0) z = 1.0;
  const double z1 = acos(z);
  const double z2 = cos(z1 * 0.5);
  const double z2_sq_m1 = z2 * z2 + -1.0;
  const double h1 = m2 * z2;
  const double h2 = sqrt(m2 + z2_sq_m1);
  const double rp = (h1 - h2) / (h1 + h2);
  const double rl = (z2 - h2) / (z2 + h2);
  const double xmp = 0.5 * (rp * rp + rl * rl);
  double a = 2.0 * z2;
  double b = (xi + xj) / a;
  if (b > 1.0) b = 1.0;
  a = 0.5 * PIE - asin(b);
  const double ta = tan(a);
  const double argument = ta * ta / sigma2;
  double k = 0.0;
  if (argument < 88.0) {
    const double prob = exp(-argument);
    const double fac1 = prob / sigma2;
    const double fac2 = 0.25 / xi / pow(b, 4.0);   // B ** FOUR with a real exponent
    k = xmp * fac1 * fac2 / xj;
  }
  const double shadow = 1.0 / (1.0 + shi + shr);
  return k * shadow;
}

// sigma^2 of VLIDORT_CoxMunk_SingleGeom (V283/VLIDORT_SurfaceMod.F90:880-886):
// 0.003 + 0.00512 |w10| with default-real literals, |w10| capped at 50 m/s.
__host__ __device__ inline double cm_sigma2(double u10m, double v10m) {
#pragma clang fp contract(off)
  double w = sqrt(u10m * u10m + v10m * v10m);
  if (w < 0.0) w = 0.0;
  if (w > 50.0) w = 50.0;
  return (double)0.003f + (double)0.00512f * w;
}

// Exact direct bounce of the Cox-Munk surface (one kernel, factor 1).
__host__ __device__ inline double cm_db(double sigma2, double m2, double sza, double vza, double raa) {
#pragma clang fp contract(off)
  const double mu0 = cos(sza * DEG_TO_RAD);
  const double s0 = sqrt(1.0 - mu0 * mu0);
  const double muv = cos(vza * DEG_TO_RAD);
  const double sv = sqrt(1.0 - muv * muv);
  const double phi = raa * DEG_TO_RAD;
  const double k = cox_munk(sigma2, m2, mu0, s0, muv, sv, cos(phi), cm_shadow(sigma2, mu0), cm_shadow(sigma2, muv));
  return 0.0 + 1.0 * k;
}

// ---- BPDF-NDVI polarised land kernel (BPDFNDVI_VFUNCTION,
// VB/vbrdf_sup_kernels.f90:3017-3198) with the vector Fresnel reflection
// amplitudes of a plane facet (FRESNEL_VECTOR, :3575-3702).  pars: the
// refractive index of the facets, NDVI (outside [-1, 1] taken as 0) and the
// scaling factor C.  Writes the 3 x 3 block k[3 o1 + o2] of the reflection
// matrix (full = false: the 2 x 2 block only, as the white-sky sums ask for
// the (1,1) element).  Direction cosines within 1e-9 of 1 are moved to
// 0.999999999999 (their sines kept), as the reference does to its arguments.
__host__ __device__ inline void bpdf_ndvi(double rn, double ndvi, double cfac, bool full, double xi, double sxi,
                                          double xj, double sxj, double cphi, double skphi, double* k) {
#pragma clang fp contract(off)
  if (fabs(xi - 1.0) < 1e-9) xi = 0.999999999999;
  if (fabs(xj - 1.0) < 1e-9) xj = 0.999999999999;
  double z = xi * xj - sxi * sxj * cphi;
  if (z > 1.0) z = 1.0;
  const double zh = cos(acos(z) * 0.5);       // cosine of half the scattering angle
  if (ndvi > 1.0 || ndvi < -1.0) ndvi = 0.0;
  const double fp0 = 0.25 * exp(-ndvi) / (xi + xj);
  const double sg = sqrt(1.0 - zh * zh);
  const double att = exp(-sg / zh);
  double h = cfac * fp0 * att;
  // Fresnel amplitudes: incident direction (sxi, 0, -xi) at azimuth 0,
  // reflected (sxj cphi, sxj skphi, xj); the facet normal bisects them
  const double vi1 = sxi * 1.0, vi2 = sxi * 0.0, vi3 = -xi;
  const double vr1 = sxj * cphi, vr2 = sxj * skphi, vr3 = xj;
  const double u1 = vi1 - vr1, u2 = vi2 - vr2, u3 = vi3 - vr3;
  const double un = sqrt(1.0 / (u1 * u1 + u2 * u2 + u3 * u3));
  const double cinc = un * (u1 * vi1 + u2 * vi2 + u3 * vi3);   // cosine of the local incidence angle
  const double ctr = sqrt(1.0 - (1.0 - cinc * cinc) * 1.0 * 1.0 / (rn * rn));
  double c1 = 1.0 * cinc, c2 = rn * ctr;
  const double rper = (c1 - c2) / (c1 + c2);
  c1 = rn * cinc; c2 = 1.0 * ctr;
  const double rpar = (c1 - c2) / (c1 + c2);
  const double ti1 = -xi * 1.0, ti2 = -xi * 0.0, ti3 = -sxi;
  const double tr1 = xj * cphi, tr2 = xj * skphi, tr3 = -sxj;
  const double pi1 = -0.0, pi2 = 1.0, pi3 = 0.0;
  const double pr1 = -skphi, pr2 = cphi, pr3 = 0.0;
  const double pikr = pi1 * vr1 + pi2 * vr2 + pi3 * vr3;
  const double prki = pr1 * vi1 + pr2 * vi2 + pr3 * vi3;
  const double tikr = ti1 * vr1 + ti2 * vr2 + ti3 * vr3;
  const double trki = tr1 * vi1 + tr2 * vi2 + tr3 * vi3;
  const double e1 = pikr * prki, e2 = tikr * trki, e3 = tikr * prki, e4 = pikr * trki;
  double f11 = e1 * rper + e2 * rpar;
  const double f12 = -(e3 * rper) + e4 * rpar;
  const double f21 = -(e4 * rper) + e3 * rpar;
  double f22 = e2 * rper + e1 * rpar;
  const double vp1 = vi2 * vr3 - vi3 * vr2, vp2 = vi3 * vr1 - vi1 * vr3, vp3 = vi1 * vr2 - vi2 * vr1;
  double dmod = vp1 * vp1 + vp2 * vp2 + vp3 * vp3;
  dmod = dmod * dmod;
  if (dmod == 0.0) { f11 = rpar; f22 = rper; dmod = 1.0; }
  h = h / dmod;
  double a11 = fabs(f11), a12 = fabs(f12), a21 = fabs(f21), a22 = fabs(f22);
  a11 = a11 * a11; a12 = a12 * a12; a21 = a21 * a21; a22 = a22 * a22;
  const double hh = 0.5 * h;
  k[0] = (a11 + a12 + a21 + a22) * hh;
  k[1] = (a11 - a12 + a21 - a22) * hh;
  k[3] = (a11 - a22 + a12 - a21) * hh;
  k[4] = (a11 - a12 - a21 + a22) * hh;
  if (!full) return;
  const double c1112 = f11 * f12, c1121 = f11 * f21, c1122 = f11 * f22;
  const double c1221 = f12 * f21, c1222 = f12 * f22, c2122 = f21 * f22;
  k[2] = -((c1112 + c2122) * h);   // sine Fourier components
  k[5] = -((c1112 - c2122) * h);   // sine
  k[6] = -((c1121 + c1222) * h);   // sine
  k[7] = -((c1121 - c1222) * h);   // sine
  k[8] = (c1122 + c1221) * h;      // cosine
}

}  // namespace aa
