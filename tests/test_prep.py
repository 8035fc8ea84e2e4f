"""Oracle of the input preparation (oracle/prep.py) pinned by published and
closed-form results:

* Rayleigh cross section: Bodhaine et al. (1999), J. Atmos. Oceanic Technol.
  16, 1854-1861, eq. (29) fit for 360 ppm CO2 (accuracy of the fit ~0.01%
  over 0.25-1 um);
* the layer optical thickness integrates the molecular density over height:
  for an isothermal exponential profile the column is Ns*(Ts/T)*H exactly;
* the spectral slope of the Rayleigh optical thickness (about lambda^-4.05);
* getEdgeVars: log-pressure interpolation is exact for T linear in ln p,
  heights are the hydrostatic sums from the surface, the bottom edge follows
  the dry adiabat from the bottom layer's mean.
"""
import math

import numpy as np
import pytest

from oracle import prep as P


def bodhaine_xsec(lam_um):
    l2 = lam_um * lam_um
    il2 = 1.0 / l2
    return (1.0455996 - 341.29061 * il2 - 0.90230850 * l2) / (1.0 + 0.0027059889 * il2 - 85.968563 * l2) * 1e-28


@pytest.mark.parametrize("lam_nm", [354.0, 388.0, 440.0, 550.0, 670.0, 870.0, 1000.0])
def test_rayleigh_cross_section_matches_published_fit(lam_nm):
    s = P.ray_xsec(lam_nm * P.F32(1e-3))
    assert abs(s / bodhaine_xsec(lam_nm * 1e-3) - 1.0) < 2e-3


def test_column_optical_depth_is_density_integral():
    # isothermal exponential atmosphere, 4000 layers to 200 km
    T, H = 250.0, 7.3          # K, km
    z = np.linspace(200.0, 0.0, 4001)
    p = 101325.0 * np.exp(-z / H)       # Pa
    pe, he, te = p[:, None], (z * 1000.0)[:, None], np.full((4001, 1), T)
    rot, _ = P.rot_calc([550.0], pe, he, te)
    lam = 550.0 * P.F32(1e-3)
    ns = P.ns_std()
    ts, ps = P.F32(288.15), P.F32(1013.25)
    # density at the surface (p = 1013.25 hPa): Ns * (p/Ps) * (Ts/T); column = n0 * H
    n0 = ns * ((101325.0 * P.F32(1e-2)) / ps) * (ts / T)
    expect = P.ray_xsec(lam) * n0 * H * 1e5 * (1.0 - math.exp(-200.0 / H))
    assert abs(rot.sum() / expect - 1.0) < 1e-4


def test_spectral_slope():
    z = np.linspace(60.0, 0.0, 61)
    p = 101325.0 * np.exp(-z / 7.5)
    pe, he, te = p[:, None], (z * 1000.0)[:, None], np.full((61, 1), 260.0)
    rot, depol = P.rot_calc([400.0, 800.0], pe, he, te)
    slope = math.log(rot[:, 0, 0].sum() / rot[:, 0, 1].sum()) / math.log(2.0)
    assert 4.0 < slope < 4.15
    # King-factor depolarisation of air: a few percent, decreasing with wavelength
    assert 0.02 < depol[1] < depol[0] < 0.035


def _log_linear_profile(km, nobs, seed=3):
    rng = np.random.default_rng(seed)
    a = rng.uniform(150.0, 200.0, nobs)
    b = rng.uniform(3.0, 8.0, nobs)       # T = a + b ln p
    delp = rng.uniform(200.0, 2500.0, (km, nobs)).astype(np.float32)
    ptop = 1.0
    pe = ptop + np.vstack([np.zeros((1, nobs)), np.cumsum(delp.astype(np.float64), axis=0)])
    pm = 0.5 * (pe[:-1] + pe[1:])
    tm = a + b * np.log(pm)
    airdens = (pm / (287.0 * tm)).astype(np.float32)
    return airdens, delp, ptop, a, b, pe


def test_edge_vars_closed_forms():
    km, nobs = 72, 5
    ad, dp, ptop, a, b, pe_exact = _log_linear_profile(km, nobs)
    pe, ze, te = P.edge_vars(ad, dp, ptop)
    assert np.allclose(pe, pe_exact, rtol=1e-5)
    # heights: hydrostatic sums from the surface
    dz = dp.astype(np.float64) / (ad.astype(np.float64) * 9.80616)
    z_exact = np.vstack([np.cumsum(dz[::-1], axis=0)[::-1], np.zeros((1, nobs))])
    assert np.allclose(ze, z_exact, rtol=2e-5, atol=1e-3)
    # interior edges: exact for T linear in ln p
    t_exact = a + b * np.log(pe_exact)
    assert np.allclose(te[1:km], t_exact[1:km], rtol=2e-5)
    # top: isothermal at the top layer's mean; bottom: dry adiabat from the bottom layer
    pm = 0.5 * (pe_exact[:-1] + pe_exact[1:])
    tm = pm / (ad.astype(np.float64) * 287.0)
    assert np.allclose(te[0], tm[0], rtol=2e-6)
    assert np.allclose(te[km], tm[km - 1] * (pe_exact[km] / pm[km - 1]) ** (2.0 / 7.0), rtol=2e-5)


def test_calc_qm():
    rng = np.random.default_rng(5)
    q = rng.uniform(0, 1e-8, (7, 3, 4)).astype(np.float32)
    dp = rng.uniform(100, 2000, (7, 4)).astype(np.float32)
    qm = P.calc_qm(q, dp)
    assert qm.dtype == np.float32
    assert np.allclose(qm, q.astype(np.float64) * dp[:, None, :] / 9.81, rtol=1e-6)
