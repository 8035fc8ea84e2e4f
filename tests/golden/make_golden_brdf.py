"""Writes the MODIS land-surface (BRDF) golden fixtures into tests/golden/brdf/.

* selftest.json: the Ross-thick and Li-sparse blocks of the reference's own
  BRDF self-tester results (vlidort_s_test/saved_results/gfortran/
  results_brdf_self_tester.res) with the tester's inputs
  (vlidort_s_test/VBRDF_ReadInput_self_test.cfg and GET_VBRDF_PLUS in
  2p8p3_brdf_self_tester.f90): one solar beam at 35 deg, view 10 deg,
  relative azimuths 0/79/180 deg, one half-space stream, 100 BRDF azimuth
  streams, kernel factor 0.99, Li-sparse parameters (1, 2).
* supplement.npz: the BRDF supplement of the compiled reference
  (oracle/_ref, VBRDF_MAINMASTER under the VLIDORT_LANDMODIS_SingleGeom set-up)
  for a set of geometries, kernel weights and stream counts, including
  white-sky-albedo failures and a negative direct bounce.
* modis_*.npz: radiances of the reference's own VLIDORT_Scalar_LandMODIS
  driver (oracle/_ref/libvlidort_drv.so via oracle/ref_calls.f90) on
  synthetic columns (inputs stored with the outputs); the oracle's
  statement-by-statement harness must reproduce them bit for bit, and its
  Fourier-term counts are stored with them.
* supplement_cx.npz, cx_*.npz: the same for the Cox-Munk ocean surface
  (VLIDORT_CoxMunk_SingleGeom supplement, VLIDORT_Scalar_CX radiances).
* vcx_*.npz: the vector Cox-Munk driver (VLIDORT_Vector_CX_SingleGeom:
  NSTOKES = 3, I, Q, U and the direct-bounce BRDF outputs).
* supplement_bpdf.npz, vbpdf_*.npz: the polarised land surfaces
  (VLIDORT_LAMBERTIAN_BPDF / VLIDORT_LANDMODIS_BPDF supplements with NSTOKES
  = 3; VLIDORT_Vector_Lambert_BPDF and VLIDORT_Vector_LandMODIS_BPDF
  radiances, I, Q, U and the BR, BR_Q, BR_U outputs).

Run here (the reference is only available in this container):
    python tests/golden/make_golden_brdf.py          # everything
    python tests/golden/make_golden_brdf.py vcx      # the vector Cox-Munk cases only
    python tests/golden/make_golden_brdf.py vbpdf    # the polarised land cases and their supplement
    python tests/golden/make_golden_brdf.py cx_n20   # named cases only
"""
import json
import os
import re
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from aeroapps_amd import synthetic as S  # noqa: E402
import oracle  # noqa: E402
from oracle import chem_mie as CM  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "brdf")
SELFTEST = ("/root/reference/src/Components/rtms/RTSI/VLIDORT2p8p3/vlidort_s_test/saved_results/"
            "gfortran/results_brdf_self_tester.res")


def selftest_vectors():
    text = open(SELFTEST).read()
    blocks = re.split(r"\*{10,}", text)
    out = {}
    for b in blocks:
        m = re.search(r"VBRDF Name: (\S+)\s+VBRDF Index:\s+(\d+)", b)
        if not m or m.group(1) not in ("Ross-thick", "Li-sparse"):
            continue
        name = m.group(1)
        exact = [float(v) for v in re.findall(r"^\s+1\s+1\s+\d\s+(\S+)\s*$", b, re.M)]
        four = {}
        for fm in re.finditer(r"Fourier component #\s+(\d+)(.*?)(?=Fourier component|\Z)", b, re.S):
            vals = {}
            for key in ("Stream-Stream", "Solars-Stream", "Stream-Users", "Solars-Users"):
                vals[key] = float(re.search(key + r"\s+1\s+1\s+(\S+)", fm.group(2)).group(1))
            four[int(fm.group(1))] = vals
        out[name] = dict(index=int(m.group(2)), exact=exact, fourier=four)
    inputs = dict(nstreams=1, sza=35.0, vza=10.0, raa=[0.0, 79.0, 180.0], nstreams_brdf=100, factor=0.99,
                  lisparse_pars=[1.0, 2.0, 0.0])
    return dict(inputs=inputs, kernels=out)


SUP_CASES = [
    # (nstreams, sza, vza, raa, fiso, fgeo, fvol, p1, p2)
    (6, 35.0, 10.0, 79.0, 0.10, 0.03, 0.05, 2.0, 1.0),
    (6, 60.0, 45.0, 150.0, 0.20, 0.08, 0.10, 2.0, 1.0),
    (6, 20.0, 0.0, 0.0, 0.05, 0.01, 0.04, 2.0, 1.0),       # nadir view
    (6, 0.0, 30.0, 180.0, 0.15, 0.02, 0.09, 2.0, 1.0),     # overhead sun
    (6, 70.0, 60.0, 180.0, 0.30, 0.05, 0.15, 2.5, 1.5),
    (6, 30.0, 20.0, 40.0, 1.00, 0.01, 0.50, 2.0, 1.0),     # white-sky albedo > 1: last kernel's Fourier dropped
    (6, 20.0, 5.0, 0.0, 0.05, 0.30, 0.40, 2.5, 1.5),       # white-sky albedo < 0 and negative direct bounce
    (2, 45.0, 30.0, 90.0, 0.12, 0.02, 0.06, 2.0, 1.0),
    (4, 50.0, 25.0, 120.0, 0.08, 0.015, 0.03, 2.0, 1.0),
    (8, 40.0, 35.0, 10.0, 0.25, 0.04, 0.12, 2.0, 1.0),
    (1, 35.0, 10.0, 79.0, 0.10, 0.03, 0.05, 2.0, 1.0),
]


def supplement(R):
    recs = {k: [] for k in ("F", "F0", "UF", "UF0", "db", "wsa", "status")}
    nmax = max(c[0] for c in SUP_CASES)
    for c in SUP_CASES:
        r = R.vbrdf_modis(*c)
        n = c[0]
        F = np.zeros((2 * nmax, nmax, nmax)); F[:2 * n, :n, :n] = r["F"]
        F0 = np.zeros((2 * nmax, nmax)); F0[:2 * n, :n] = r["F0"]
        UF = np.zeros((2 * nmax, nmax)); UF[:2 * n, :n] = r["UF"]
        UF0 = np.zeros(2 * nmax); UF0[:2 * n] = r["UF0"]
        for k, v in (("F", F), ("F0", F0), ("UF", UF), ("UF0", UF0), ("db", r["db"]), ("wsa", r["wsa"]),
                     ("status", r["status"])):
            recs[k].append(v)
    out = {k: np.asarray(v) for k, v in recs.items()}
    out["cases"] = np.asarray(SUP_CASES, dtype=np.float64)
    return out


RAD_CASES = {
    # name: (ncol, channels, nstreams, nmom, edge)
    "modis_n6_4ch": (6, (354.0, 388.0, 550.0, 870.0), 6, 64, False),
    "modis_n4_edge": (10, (550.0,), 4, 32, True),
    "modis_n2": (5, (550.0,), 2, 32, False),
    "modis_n8": (3, (550.0,), 8, 32, False),
    "modis_n12": (2, (550.0,), 12, 32, False),
    "modis_n16": (2, (550.0,), 16, 64, False),
}


def _same_as_driver(got, want):
    """Harness output equals the reference driver's bit for bit (the driver
    wrapper fills the BR it leaves undefined for a missing geometry with
    MISSING, as the harness does)."""
    for a, b in zip(got, want):
        a, b = np.array(a), np.array(b)
        assert np.array_equal(a, b), np.max(np.abs(a - b))


def radiance_case(R, ncol, channels, nstreams, nmom, edge, seed=11):
    cols = S.make_columns(ncol, 72, channels, seed=seed)
    tabs = S.make_mie_tables(channels, nmom=nmom, npol=1)
    tau, ssa, g, pm, rc = CM.get_aop_vector(tabs, channels, tabs[0].lambda_m, S.TRACERS, cols.qm, cols.rh, nmom, 1)
    assert rc == 0
    kw, par = S.make_modis_kernels(ncol, len(channels), seed=seed)
    sza, vza, raa = cols.sza.copy(), cols.vza.copy(), cols.raa.copy()
    if edge:
        vza[1] = 0.0                       # nadir view
        sza[3] = S.MISSING                 # missing geometry
        kw[4, 0, :] = (0.05, 0.30, 0.40)   # negative direct bounce -> -500
        kw[5, 0, :] = (1.00, 0.01, 0.50)   # white-sky albedo > 1
        kw[6, 0, 1] = S.MISSING            # missing kernel weight (the reference still computes)
        sza[7] = 0.0                       # overhead sun latches nadir FO for later columns
        raa[8] = 0.0
    f = lambda a: np.asarray(a, dtype=np.float64)
    args = (72, cols.channels, nstreams, nmom, cols.rot, cols.depol, f(tau), f(ssa), f(g), f(pm), cols.pe, cols.he,
            cols.te, kw, par, sza, raa, vza)
    rad, refl, br, nf, rcr = R.scalar_landmodis(*args)
    drad, drefl, dbr, drc = oracle.drivers().scalar_landmodis(*args)
    assert drc == rcr
    _same_as_driver((rad, refl, br), (drad, drefl, dbr))
    return dict(km=72, channels=np.asarray(channels, float), nmom=nmom, npol=1, meta_nstreams=nstreams,
                tau=tau, ssa=ssa, g=g, pmom=pm, rot=cols.rot, depol=cols.depol, pe=cols.pe, he=cols.he,
                te=cols.te, kernel_wt=kw, param=par, sza=sza, vza=vza, raa=raa,
                out_radiance=rad, out_reflectance=refl, out_br=br, out_nfourier=nf, out_rc=rcr)


CX_SUP_CASES = [
    # (nstreams, sza, vza, raa, u10m, v10m, mr)
    (6, 30.0, 20.0, 150.0, 3.0, 4.0, 1.34),
    (6, 40.0, 35.0, 178.0, 7.0, 1.0, 1.33),     # near the specular direction
    (6, 10.0, 5.0, 0.0, 0.5, 0.2, 1.34),        # calm sea
    (6, 0.0, 30.0, 180.0, 9.0, -3.0, 1.34),     # overhead sun
    (6, 50.0, 0.0, 90.0, -6.0, 8.0, 1.335),     # nadir view
    (6, 30.0, 30.0, 180.0, 60.0, 0.0, 1.34),    # wind capped at 50 m/s
    (4, 60.0, 50.0, 90.0, 15.0, 12.0, 1.34),
    (8, 35.0, 40.0, 160.0, 2.0, 2.0, 1.34),
    (2, 25.0, 15.0, 120.0, 5.0, 5.0, 1.34),
]


def supplement_cx(R):
    recs = {k: [] for k in ("F", "F0", "UF", "UF0", "db", "wsa", "status")}
    nmax = max(c[0] for c in CX_SUP_CASES)
    for c in CX_SUP_CASES:
        r = R.vbrdf_coxmunk(*c)
        n = c[0]
        F = np.zeros((2 * nmax, nmax, nmax)); F[:2 * n, :n, :n] = r["F"]
        F0 = np.zeros((2 * nmax, nmax)); F0[:2 * n, :n] = r["F0"]
        UF = np.zeros((2 * nmax, nmax)); UF[:2 * n, :n] = r["UF"]
        UF0 = np.zeros(2 * nmax); UF0[:2 * n] = r["UF0"]
        for k, v in (("F", F), ("F0", F0), ("UF", UF), ("UF0", UF0), ("db", r["db"]), ("wsa", r["wsa"]),
                     ("status", r["status"])):
            recs[k].append(v)
    out = {k: np.asarray(v) for k, v in recs.items()}
    out["cases"] = np.asarray(CX_SUP_CASES, dtype=np.float64)
    return out


CX_CASES = {
    # name: (ncol, channels, nstreams, nmom, edge)
    "cx_n6_2ch": (6, (470.0, 660.0), 6, 64, False),
    "cx_n4_edge": (9, (550.0,), 4, 32, True),
    "cx_n8": (3, (550.0,), 8, 32, False),
    "cx_n12": (2, (550.0,), 12, 32, False),
    "cx_n20": (2, (550.0,), 20, 64, False),
}


def cx_case(R, ncol, channels, nstreams, nmom, edge, seed=13):
    cols = S.make_columns(ncol, 72, channels, seed=seed)
    tabs = S.make_mie_tables(channels, nmom=nmom, npol=1)
    tau, ssa, g, pm, rc = CM.get_aop_vector(tabs, channels, tabs[0].lambda_m, S.TRACERS, cols.qm, cols.rh, nmom, 1)
    assert rc == 0
    u, v, mr = S.make_ocean(ncol, channels, seed=seed)
    sza, vza, raa = cols.sza.copy(), cols.vza.copy(), cols.raa.copy()
    if edge:
        vza[1] = 0.0                 # nadir view
        sza[2] = S.MISSING           # missing geometry
        u[3] = S.MISSING             # missing wind: element MISSING
        sza[4], vza[4], raa[4] = 35.0, 35.0, 180.0   # specular glint
        sza[5] = 0.0                 # overhead sun latches nadir FO for later columns
        u[7], v[7] = 60.0, 10.0      # wind capped at 50 m/s
    f = lambda a: np.asarray(a, dtype=np.float64)
    args = (72, cols.channels, nstreams, nmom, cols.rot, cols.depol, f(tau), f(ssa), f(g), f(pm), cols.pe, cols.he,
            cols.te, u, v, mr, sza, raa, vza)
    rad, refl, br, nf, rcr = R.scalar_cx(*args)
    drad, drefl, dbr, drc = oracle.drivers().scalar_cx(*args)
    assert drc == rcr
    _same_as_driver((rad, refl, br), (drad, drefl, dbr))
    return dict(km=72, channels=np.asarray(channels, float), nmom=nmom, npol=1, meta_nstreams=nstreams,
                tau=tau, ssa=ssa, g=g, pmom=pm, rot=cols.rot, depol=cols.depol, pe=cols.pe, he=cols.he,
                te=cols.te, u10m=u, v10m=v, mr=mr, sza=sza, vza=vza, raa=raa,
                out_radiance=rad, out_reflectance=refl, out_br=br, out_nfourier=nf, out_rc=rcr)


VCX_CASES = {
    # name: (ncol, channels, nstreams, nmom, edge)
    "vcx_n4_2ch": (6, (470.0, 660.0), 4, 32, False),
    "vcx_n2_edge": (9, (550.0,), 2, 32, True),
    "vcx_n3": (3, (870.0,), 3, 32, False),
}


def vcx_case(R, ncol, channels, nstreams, nmom, edge, seed=17):
    cols = S.make_columns(ncol, 72, channels, seed=seed)
    tabs = S.make_mie_tables(channels, nmom=nmom, npol=6)
    tau, ssa, g, pm, rc = CM.get_aop_vector(tabs, channels, tabs[0].lambda_m, S.TRACERS, cols.qm, cols.rh, nmom, 6)
    assert rc == 0
    u, v, mr = S.make_ocean(ncol, channels, seed=seed)
    sza, vza, raa = cols.sza.copy(), cols.vza.copy(), cols.raa.copy()
    if edge:
        vza[1] = 0.0                 # nadir view
        sza[2] = S.MISSING           # missing geometry
        u[3] = S.MISSING             # missing wind: element MISSING
        sza[4], vza[4], raa[4] = 35.0, 35.0, 180.0   # specular glint
        sza[5] = 0.0                 # overhead sun latches nadir FO for later columns
        u[7], v[7] = 60.0, 10.0      # wind capped at 50 m/s
    f = lambda a: np.asarray(a, dtype=np.float64)
    args = (72, cols.channels, nstreams, nmom, cols.rot, cols.depol, f(tau), f(ssa), f(pm), cols.pe, cols.he,
            cols.te, u, v, mr, sza, raa, vza)
    rad, refl, q, uu, br, brq, bru, nf, rcr = R.vector_cx(*args)
    d = oracle.drivers().vector_cx(*args)
    assert d[-1] == rcr
    _same_as_driver((rad, refl, q, uu, br, brq, bru), d[:7])
    return dict(km=72, channels=np.asarray(channels, float), nmom=nmom, npol=6, meta_nstreams=nstreams,
                tau=tau, ssa=ssa, pmom=pm, rot=cols.rot, depol=cols.depol, pe=cols.pe, he=cols.he,
                te=cols.te, u10m=u, v10m=v, mr=mr, sza=sza, vza=vza, raa=raa,
                out_radiance=rad, out_reflectance=refl, out_q=q, out_u=uu, out_br=br, out_br_q=brq,
                out_br_u=bru, out_nfourier=nf, out_rc=rcr)


BPDF_SUP_CASES = [
    # (nstreams, sza, vza, raa, modis, fiso|albedo, fgeo, fvol, h/b, b/r, n, NDVI, C)
    (4, 35.0, 20.0, 60.0, 1, 0.10, 0.03, 0.05, 2.0, 1.0, 1.5, 0.4, 6.0),
    (4, 50.0, 40.0, 150.0, 0, 0.08, 0.0, 0.0, 0.0, 0.0, 1.5, 0.7, 5.0),
    (6, 60.0, 10.0, 100.0, 1, 0.20, 0.05, 0.10, 2.0, 1.0, 1.5, -0.1, 7.5),
    (2, 25.0, 0.0, 0.0, 0, 0.12, 0.0, 0.0, 0.0, 0.0, 1.5, 0.2, 4.0),      # nadir view
    (3, 0.0, 30.0, 180.0, 1, 0.15, 0.02, 0.09, 2.0, 1.0, 1.5, 1.4, 6.0),  # overhead sun, NDVI > 1 taken as 0
    (4, 30.0, 20.0, 40.0, 1, 1.00, 0.01, 0.50, 2.0, 1.0, 1.5, 0.3, 6.0),  # white-sky albedo > 1
    (5, 45.0, 30.0, 90.0, 0, 0.05, 0.0, 0.0, 0.0, 0.0, 1.5, 0.5, 8.0),
]


def supplement_bpdf(R):
    recs = {k: [] for k in ("F", "F0", "UF", "DB", "wsa", "status")}
    nmax = max(c[0] for c in BPDF_SUP_CASES)
    for c in BPDF_SUP_CASES:
        r = R.vbrdf_bpdf(*c)
        n = c[0]
        F = np.zeros((2 * nmax, 16, nmax, nmax)); F[:2 * n, :, :n, :n] = r["F"]
        F0 = np.zeros((2 * nmax, 16, nmax)); F0[:2 * n, :, :n] = r["F0"]
        UF = np.zeros((2 * nmax, 16, nmax)); UF[:2 * n, :, :n] = r["UF"]
        for k, v in (("F", F), ("F0", F0), ("UF", UF), ("DB", r["DB"]), ("wsa", r["wsa"]), ("status", r["status"])):
            recs[k].append(v)
    out = {k: np.asarray(v) for k, v in recs.items()}
    out["cases"] = np.asarray(BPDF_SUP_CASES, dtype=np.float64)
    return out


VBPDF_CASES = {
    # name: (ncol, channels, nstreams, nmom, modis, edge)
    "vbpdf_modis_n4_2ch": (5, (470.0, 870.0), 4, 32, True, False),
    "vbpdf_lamb_n4_edge": (10, (550.0,), 4, 32, False, True),
    "vbpdf_modis_n3_edge": (10, (660.0,), 3, 32, True, True),
    "vbpdf_lamb_n6": (3, (550.0, 870.0), 6, 64, False, False),
    "vbpdf_modis_n2": (4, (550.0,), 2, 32, True, False),
}


def vbpdf_case(R, ncol, channels, nstreams, nmom, modis, edge, seed=23):
    cols = S.make_columns(ncol, 72, channels, seed=seed)
    tabs = S.make_mie_tables(channels, nmom=nmom, npol=6)
    tau, ssa, g, pm, rc = CM.get_aop_vector(tabs, channels, tabs[0].lambda_m, S.TRACERS, cols.qm, cols.rh, nmom, 6)
    assert rc == 0
    kw, par = S.make_modis_kernels(ncol, len(channels), seed=seed)
    bp = S.make_bpdf(ncol, len(channels), seed=seed)
    albedo = cols.albedo.copy()
    sza, vza, raa = cols.sza.copy(), cols.vza.copy(), cols.raa.copy()
    if edge:
        vza[1] = 0.0                       # nadir view
        if modis:
            # missing geometry (not in the Lambertian form: its driver then
            # writes Q(j,i), U(j,i) with i left at nch + 1 by the previous
            # column's channel loop, outside the arrays, GEO/VLIDORT_LAMBERT_BPDF.F90:128-134)
            sza[2] = S.MISSING
            kw[3, 0, :] = (0.05, 0.30, 0.40)   # negative direct bounce -> -500
            kw[4, 0, :] = (1.00, 0.01, 0.50)   # white-sky albedo > 1: BPDF Fourier terms left out
            kw[6, 0, 1] = S.MISSING            # missing kernel weight (the reference still computes)
        else:
            albedo[0, 3] = -0.2                # negative direct bounce -> -500
            albedo[0, 4] = 0.97                # white-sky albedo > 1 with the BPDF kernel
            albedo[0, 6] = S.MISSING           # missing albedo: element MISSING
        bp[5, 0, 1] = 1.5                  # NDVI outside [-1, 1] taken as 0
        sza[7] = 0.0                       # overhead sun latches nadir FO for later columns
        raa[8] = 0.0
    f = lambda a: np.asarray(a, dtype=np.float64)
    common = (72, cols.channels, nstreams, nmom, cols.rot, cols.depol, f(tau), f(ssa), f(pm), cols.pe, cols.he,
              cols.te)
    out = R.vector_bpdf(*common, int(modis), albedo, kw, par, bp, sza, raa, vza)
    rad, refl, q, uu, br, brq, bru, nf, rcr = out
    D = oracle.drivers()
    if modis:
        d = D.vector_landmodis_bpdf(*common, kw, par, bp, sza, raa, vza)
    else:
        d = D.vector_lambert_bpdf(*common, albedo, bp, sza, raa, vza)
    assert d[-1] == rcr
    _same_as_driver((rad, refl, q, uu, br, brq, bru), d[:7])
    return dict(km=72, channels=np.asarray(channels, float), nmom=nmom, npol=6, meta_nstreams=nstreams,
                meta_modis=int(modis), tau=tau, ssa=ssa, pmom=pm, rot=cols.rot, depol=cols.depol, pe=cols.pe,
                he=cols.he, te=cols.te, albedo=albedo, kernel_wt=kw, param=par, bpdf=bp, sza=sza, vza=vza, raa=raa,
                out_radiance=rad, out_reflectance=refl, out_q=q, out_u=uu, out_br=br, out_br_q=brq,
                out_br_u=bru, out_nfourier=nf, out_rc=rcr)


def main():
    names = sys.argv[1:]
    if names:   # only the named radiance cases ("vcx": every vector Cox-Munk case)
        R = oracle.drv_harness_modis()
        assert R is not None, "build oracle/_ref first (make -C oracle -f Makefile.ref)"
        if names == ["vcx"]:
            names = list(VCX_CASES)
        if names == ["vbpdf"]:
            np.savez_compressed(os.path.join(OUT, "supplement_bpdf.npz"), **supplement_bpdf(R))
            names = list(VBPDF_CASES)
        for name in names:
            if name in VBPDF_CASES:
                d = vbpdf_case(R, *VBPDF_CASES[name])
            elif name in RAD_CASES:
                d = radiance_case(R, *RAD_CASES[name])
            elif name in CX_CASES:
                d = cx_case(R, *CX_CASES[name])
            else:
                d = vcx_case(R, *VCX_CASES[name])
            np.savez_compressed(os.path.join(OUT, name + ".npz"), **d)
            print(name, d["out_rc"], d["out_radiance"].ravel()[:4], d["out_nfourier"].ravel()[:6])
        return
    os.makedirs(OUT, exist_ok=True)
    with open(os.path.join(OUT, "selftest.json"), "w") as fh:
        json.dump(selftest_vectors(), fh, indent=1)
    R = oracle.drv_harness_modis()
    assert R is not None, "build oracle/_ref first (make -C oracle -f Makefile.ref)"
    np.savez_compressed(os.path.join(OUT, "supplement.npz"), **supplement(R))
    for name, c in RAD_CASES.items():
        d = radiance_case(R, *c)
        np.savez_compressed(os.path.join(OUT, name + ".npz"), **d)
        print(name, d["out_rc"], d["out_radiance"].ravel()[:4], d["out_nfourier"].ravel()[:6])
    np.savez_compressed(os.path.join(OUT, "supplement_cx.npz"), **supplement_cx(R))
    for name, c in CX_CASES.items():
        d = cx_case(R, *c)
        np.savez_compressed(os.path.join(OUT, name + ".npz"), **d)
        print(name, d["out_rc"], d["out_radiance"].ravel()[:4], d["out_nfourier"].ravel()[:6])
    for name, c in VCX_CASES.items():
        d = vcx_case(R, *c)
        np.savez_compressed(os.path.join(OUT, name + ".npz"), **d)
        print(name, d["out_rc"], d["out_radiance"].ravel()[:4], d["out_nfourier"].ravel()[:6])
    np.savez_compressed(os.path.join(OUT, "supplement_bpdf.npz"), **supplement_bpdf(R))
    for name, c in VBPDF_CASES.items():
        d = vbpdf_case(R, *c)
        np.savez_compressed(os.path.join(OUT, name + ".npz"), **d)
        print(name, d["out_rc"], d["out_radiance"].ravel()[:4], d["out_nfourier"].ravel()[:6])


if __name__ == "__main__":
    main()
