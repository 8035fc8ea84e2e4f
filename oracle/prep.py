"""Oracle (test infrastructure only): CPU restatement of the input
preparation around the radiance path, used by tests/ as the checker for the
GPU kernels in aeroapps_amd/csrc/prep.hip.  Nothing in the product imports it.

* rot_calc: VLIDORT_ROT_CALC (GEO/VLIDORT_ROT.F90:19-85) -> VLIDORT_Rayleigh,
  coef_depol, F_air (V283/VLIDORT_ScatMod.F90:143-281).  float64 with the
  reference's default-real (float32) literals rounded to float32 first, in the
  reference's operation order.  Pinned by the published Rayleigh optical depth
  and depolarisation of Bodhaine et al. (1999) (tests/test_prep.py).
* edge_vars: getEdgeVars (CB/MieObs_py.F90:27-83), float32.  Pinned by the
  hydrostatic and closed-form checks in tests/test_prep.py.
* calc_qm: calc_qm (GEO/geo_vlidort.F90:2124-2144, grav
  GEO/geo_vlidort_pars.F90:47), float32.

Paths: CB = src/Shared/GMAO_Shared/Chem_Base,
GEO = src/Components/rtms/geo_vlidort, V283 = src/Components/rtms/RTSI/VLIDORT2p8p3.
"""
from __future__ import annotations

import math

import numpy as np

f32 = np.float32


def F32(x) -> float:
    """A default-real literal as the reference's real*8 arithmetic sees it."""
    return float(f32(x))


def f_air(xx2: float, c: float) -> float:
    """King factor of air (V283/VLIDORT_ScatMod.F90:262-281)."""
    depol1 = F32(1.034) + F32(3.17e-4) * xx2
    depol2 = F32(1.096) + F32(1.385e-3) * xx2 + F32(1.448e-4) * xx2 * xx2
    depol3 = 1.0
    depol4 = F32(1.15)
    den0 = float((f32(78.084) + f32(20.946)) + f32(0.934))   # default-real sum
    return (F32(78.084) * depol1 + F32(20.946) * depol2 + F32(0.934) * depol3 + c * F32(100.0) * depol4) / \
        (den0 + c * F32(100.0))


def co2_ppv() -> float:
    return float(f32(360.0) * f32(1e-6))


def coef_depol(x: float) -> float:
    """Depolarisation ratio, x = wavelength in micron (:245-260)."""
    xx = 1.0 / x
    xx2 = xx * xx
    c = co2_ppv()
    return 6.0 * (f_air(xx2, c) - 1.0) / (7.0 * f_air(xx2, c) + 3.0)


def ns_std() -> float:
    av, ts = F32(6.0221367e23), F32(288.15)
    return (av / F32(22.4141)) * (F32(273.15) / ts) * float(f32(1.0) / f32(1000))


def ray_xsec(x: float) -> float:
    """Rayleigh cross section per molecule [cm^2], x in micron (A, :195-240)."""
    xx = 1.0 / x
    xx2 = xx * xx
    cm = x / F32(10000.0)
    pi = float(f32(4.0) * f32(math.atan(1.0)))
    pi3 = pi * pi * pi
    ri = F32(8060.51) + F32(2480990.0) / (F32(132.274) - xx2) + F32(17455.7) / (F32(39.32957) - xx2)
    ri = ri * F32(1e-8) + 1.0
    c = co2_ppv()
    ri = (ri - 1.0) * (1.0 + F32(0.54) * (c - F32(0.0003))) + 1.0
    ri2 = ri * ri
    ns = ns_std()
    return 24.0 * pi3 * f_air(xx2, c) * ((ri2 - 1.0) * (ri2 - 1.0)) / \
        (ns * ns * cm * cm * cm * cm * (ri2 + 2.0) * (ri2 + 2.0))


def rot_calc(channels, pe, he, te):
    """ROT[km, nobs, nch], depol[nch] from pe, he, te [km+1, nobs] (VLIDORT_Rayleigh, :143-190)."""
    pe, he, te = (np.asarray(a, dtype=np.float64) for a in (pe, he, te))
    km1, nobs = pe.shape
    km = km1 - 1
    nch = len(channels)
    rot = np.zeros((km, nobs, nch))
    depol = np.zeros(nch)
    ps, ts = F32(1013.25), F32(288.15)
    for i, lam_nm in enumerate(channels):
        w = float(lam_nm) * F32(1.e-3)
        a = ray_xsec(w)
        ns = ns_std()
        for j in range(nobs):
            vol = [a * ns * ((pe[k, j] * F32(1.e-2)) / ps) * (ts / te[k, j]) * F32(1.e5) for k in range(km1)]
            for k in range(km):
                difz = he[k, j] * F32(1.e-3) - he[k + 1, j] * F32(1.e-3)
                rot[k, j, i] = difz * (vol[k] + vol[k + 1]) / 2.0
        depol[i] = coef_depol(w)
    return rot, depol


def edge_vars(airdens, delp, ptop=1.0):
    """(pe, ze, te)[km+1, nobs] from airdens, delp [km, nobs] (CB/MieObs_py.F90:27-83), float32."""
    ad = np.asarray(airdens, dtype=f32)
    dp = np.asarray(delp, dtype=f32)
    km, nobs = ad.shape
    grav, rgas, kappa = f32(9.80616), f32(287.0), f32(2.0) / f32(7.0)
    pe = np.zeros((km + 1, nobs), dtype=f32)
    ze = np.zeros_like(pe)
    te = np.zeros_like(pe)
    lg = lambda x: f32(math.log(float(x)))
    for n in range(nobs):
        pe[0, n] = f32(ptop)
        for k in range(km):
            pe[k + 1, n] = pe[k, n] + dp[k, n]
        ze[km, n] = f32(0.0)
        for k in range(km - 1, -1, -1):
            ze[k, n] = ze[k + 1, n] + dp[k, n] / (ad[k, n] * grav)
        pm = np.array([(pe[k, n] + pe[k + 1, n]) / f32(2.0) for k in range(km)], dtype=f32)
        tm = np.array([pm[k] / (ad[k, n] * rgas) for k in range(km)], dtype=f32)
        te[0, n] = tm[0]
        for k in range(1, km):
            alpha = lg(pe[k, n] / pm[k - 1]) / lg(pm[k] / pm[k - 1])
            te[k, n] = tm[k - 1] + alpha * (tm[k] - tm[k - 1])
        te[km, n] = tm[km - 1] * f32(float(pe[km, n] / pm[km - 1]) ** float(kappa))
    return pe, ze, te


def calc_qm(q, delp):
    """qm[km, nq, nobs] = q * delp / 9.81 in float32."""
    q = np.asarray(q, dtype=f32)
    dp = np.asarray(delp, dtype=f32)
    return (q * dp[:, None, :]) / f32(9.81)
