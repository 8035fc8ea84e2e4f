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

}  // namespace aa
