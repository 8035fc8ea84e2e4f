"""Python binding of the radiance drivers, with the call contract of the
reference's f2py module ``VLIDORT_POLAR_`` (scalar_lambert / vector_lambert,
src/Components/rtms/leo_vlidort/VLIDORT_POLAR_py.F90:457-671, called from
src/Components/rtms/leo_vlidort/polar_vlidort.py): arrays are given in their
Fortran shapes (e.g. tau[km, nch, nobs]), the dimensions are taken from them,
and the outputs come back as (nobs, nch) arrays plus the return code.

Extra keyword arguments: ``precision`` (0 = FP64, the drop-in) and
``return_nfourier`` (also return the Fourier terms used per element).
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import check, lib, _dp, _ip


def _f64(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def _ptr(a, kind=_dp):
    return a.ctypes.data_as(kind)


def scalar_lambert(channels, nstreams, plane_parallel, ROT, depol, tau, ssa, g, pmom, pe, he, te, albedo,
                   solar_zenith, relat_azymuth, sensor_zenith, MISSING, verbose=0, precision=0,
                   return_nfourier=False):
    """VLIDORT_Scalar_Lambert (src/Components/rtms/geo_vlidort/VLIDORT_LAMBERT.F90:34-168)."""
    channels, ROT, depol, tau, ssa, g, pmom, pe, he, te, albedo, sza, raa, vza = map(
        _f64, (channels, ROT, depol, tau, ssa, g, pmom, pe, he, te, albedo, solar_zenith, relat_azymuth,
               sensor_zenith))
    km, nch, nobs = tau.shape
    if pmom.ndim == 4:
        pmom = _f64(pmom[..., None])
    nmom, npol = pmom.shape[3], pmom.shape[4]
    rad = np.zeros((nobs, nch), order="F")
    refl = np.zeros((nobs, nch), order="F")
    nf = np.zeros((nobs, nch), dtype=np.int32, order="F")
    rc = np.zeros(1, dtype=np.int32)
    e = lib().aa_vlidort_scalar_lambert(km, nch, nobs, _ptr(channels), int(nstreams), int(bool(plane_parallel)),
                                        nmom, npol, _ptr(ROT), _ptr(depol), _ptr(tau), _ptr(ssa), _ptr(g),
                                        _ptr(pmom), _ptr(pe), _ptr(he), _ptr(te), _ptr(albedo), _ptr(sza),
                                        _ptr(raa), _ptr(vza), float(MISSING), int(verbose), int(precision),
                                        _ptr(rad), _ptr(refl), _ptr(nf, _ip), _ptr(rc, _ip))
    check(e, "aa_vlidort_scalar_lambert")
    if return_nfourier:
        return rad, refl, int(rc[0]), nf
    return rad, refl, int(rc[0])


def vector_lambert(channels, nstreams, plane_parallel, ROT, depol, alpha, tau, ssa, pmom, pe, he, te, albedo,
                   solar_zenith, relat_azymuth, sensor_zenith, flux_factor, MISSING, verbose=0, precision=0,
                   return_nfourier=False, DO_2OS_CORRECTION=False, DO_BOA=False):
    """VLIDORT_Vector_Lambert (src/Components/rtms/geo_vlidort/VLIDORT_LAMBERT.F90:172-520).

    Single geometry when the angles are 1-D (nobs); MultiGeom when they are
    2-D (nobs, ngeom) with albedo (nobs, nch, ngeom): outputs are then
    (nobs, nch, ngeom).  DO_2OS_CORRECTION=True runs the vector driver with
    NSTOKES = 1 (Q = U = 0); DO_BOA is not implemented (AA_EUNSUP)."""
    (channels, ROT, depol, alpha, tau, ssa, pmom, pe, he, te, albedo, sza, raa, vza, ff) = map(
        _f64, (channels, ROT, depol, alpha, tau, ssa, pmom, pe, he, te, albedo, solar_zenith, relat_azymuth,
               sensor_zenith, flux_factor))
    km, nch, nobs = tau.shape
    nmom, npol = pmom.shape[3], pmom.shape[4]
    multi = sza.ndim == 2
    ngeom = sza.shape[1] if multi else 1
    shape = (nobs, nch, ngeom) if multi else (nobs, nch)
    out = [np.zeros(shape, order="F") for _ in range(4)]
    nf = np.zeros(shape, dtype=np.int32, order="F")
    rc = np.zeros(1, dtype=np.int32)
    if multi:
        e = lib().aa_vlidort_vector_lambert_multigeom(
            km, nch, nobs, ngeom, _ptr(channels), int(nstreams), int(bool(plane_parallel)), nmom, npol, _ptr(ROT),
            _ptr(depol), _ptr(alpha), _ptr(tau), _ptr(ssa), _ptr(pmom), _ptr(pe), _ptr(he), _ptr(te), _ptr(albedo),
            _ptr(sza), _ptr(raa), _ptr(vza), _ptr(ff), float(MISSING), int(verbose), int(precision),
            int(bool(DO_2OS_CORRECTION)), int(bool(DO_BOA)), *[_ptr(o) for o in out], _ptr(nf, _ip), _ptr(rc, _ip))
        check(e, "aa_vlidort_vector_lambert_multigeom")
    else:
        e = lib().aa_vlidort_vector_lambert_opt(
            km, nch, nobs, _ptr(channels), int(nstreams), int(bool(plane_parallel)), nmom, npol, _ptr(ROT),
            _ptr(depol), _ptr(alpha), _ptr(tau), _ptr(ssa), _ptr(pmom), _ptr(pe), _ptr(he), _ptr(te), _ptr(albedo),
            _ptr(sza), _ptr(raa), _ptr(vza), _ptr(ff), float(MISSING), int(verbose), int(precision),
            int(bool(DO_2OS_CORRECTION)), int(bool(DO_BOA)), *[_ptr(o) for o in out], _ptr(nf, _ip), _ptr(rc, _ip))
        check(e, "aa_vlidort_vector_lambert_opt")
    I, R, Q, U = out
    if return_nfourier:
        return I, R, Q, U, int(rc[0]), nf
    return I, R, Q, U, int(rc[0])


def scalar_landmodis(channels, nstreams, plane_parallel, ROT, depol, tau, ssa, g, pmom, pe, he, te, kernel_wt, param,
                     solar_zenith, relat_azymuth, sensor_zenith, MISSING, verbose=0, precision=0,
                     return_nfourier=False):
    """VLIDORT_Scalar_LandMODIS (src/Components/rtms/geo_vlidort/VLIDORT_BRDF_MODIS.F90:35-202):
    MODIS land-surface BRDF (kernel_wt[3, nch, nobs] = fiso, fgeo, fvol;
    param[2, nch, nobs] = Li-sparse h/b, b/r).  Returns (radiance, reflectance,
    BR, rc) with (nobs, nch) outputs; BR is the exact direct-bounce BRDF."""
    channels, ROT, depol, tau, ssa, g, pmom, pe, he, te, kw, par, sza, raa, vza = map(
        _f64, (channels, ROT, depol, tau, ssa, g, pmom, pe, he, te, kernel_wt, param, solar_zenith, relat_azymuth,
               sensor_zenith))
    km, nch, nobs = tau.shape
    if pmom.ndim == 4:
        pmom = _f64(pmom[..., None])
    nmom, npol = pmom.shape[3], pmom.shape[4]
    rad = np.zeros((nobs, nch), order="F")
    refl = np.zeros((nobs, nch), order="F")
    br = np.zeros((nobs, nch), order="F")
    nf = np.zeros((nobs, nch), dtype=np.int32, order="F")
    rc = np.zeros(1, dtype=np.int32)
    e = lib().aa_vlidort_scalar_landmodis(km, nch, nobs, _ptr(channels), int(nstreams), int(bool(plane_parallel)),
                                          nmom, npol, _ptr(ROT), _ptr(depol), _ptr(tau), _ptr(ssa), _ptr(g),
                                          _ptr(pmom), _ptr(pe), _ptr(he), _ptr(te), _ptr(kw), _ptr(par), _ptr(sza),
                                          _ptr(raa), _ptr(vza), float(MISSING), int(verbose), int(precision),
                                          _ptr(rad), _ptr(refl), _ptr(br), _ptr(nf, _ip), _ptr(rc, _ip))
    check(e, "aa_vlidort_scalar_landmodis")
    if return_nfourier:
        return rad, refl, br, int(rc[0]), nf
    return rad, refl, br, int(rc[0])


def scalar_cx(channels, nstreams, plane_parallel, ROT, depol, tau, ssa, g, pmom, pe, he, te, U10m, V10m, mr,
              solar_zenith, relat_azymuth, sensor_zenith, MISSING, verbose=0, precision=0, return_nfourier=False):
    """VLIDORT_Scalar_CX (src/Components/rtms/geo_vlidort/VLIDORT_BRDF_CX.F90:420-586):
    Cox-Munk ocean surface from the 10 m wind U10m, V10m [nobs] and the
    refractive index mr [nch].  Returns (radiance, reflectance, BRDF[1, nobs,
    nch], rc)."""
    channels, ROT, depol, tau, ssa, g, pmom, pe, he, te, u, v, m, sza, raa, vza = map(
        _f64, (channels, ROT, depol, tau, ssa, g, pmom, pe, he, te, U10m, V10m, mr, solar_zenith, relat_azymuth,
               sensor_zenith))
    km, nch, nobs = tau.shape
    if pmom.ndim == 4:
        pmom = _f64(pmom[..., None])
    nmom, npol = pmom.shape[3], pmom.shape[4]
    rad = np.zeros((nobs, nch), order="F")
    refl = np.zeros((nobs, nch), order="F")
    brdf = np.zeros((1, nobs, nch), order="F")
    nf = np.zeros((nobs, nch), dtype=np.int32, order="F")
    rc = np.zeros(1, dtype=np.int32)
    e = lib().aa_vlidort_scalar_cx(km, nch, nobs, _ptr(channels), int(nstreams), int(bool(plane_parallel)), nmom,
                                   npol, _ptr(ROT), _ptr(depol), _ptr(tau), _ptr(ssa), _ptr(g), _ptr(pmom), _ptr(pe),
                                   _ptr(he), _ptr(te), _ptr(u), _ptr(v), _ptr(m), _ptr(sza), _ptr(raa), _ptr(vza),
                                   float(MISSING), int(verbose), int(precision), _ptr(rad), _ptr(refl), _ptr(brdf),
                                   _ptr(nf, _ip), _ptr(rc, _ip))
    check(e, "aa_vlidort_scalar_cx")
    if return_nfourier:
        return rad, refl, brdf, int(rc[0]), nf
    return rad, refl, brdf, int(rc[0])


def vector_cx(channels, nstreams, plane_parallel, ROT, depol, tau, ssa, pmom, pe, he, te, U10m, V10m, mr,
              solar_zenith, relat_azymuth, sensor_zenith, MISSING, verbose=0, precision=0, return_nfourier=False):
    """VLIDORT_Vector_CX_SingleGeom (src/Components/rtms/geo_vlidort/VLIDORT_BRDF_CX.F90:587-762),
    the contract of the f2py ``VLIDORT_POLAR_.vector_cx``
    (src/Components/rtms/leo_vlidort/VLIDORT_POLAR_py.F90:830-900): the Cox-Munk
    ocean with NSTOKES = 3.  Returns (radiance, reflectance, BR, Q, U, BR_Q,
    BR_U, rc), each array (nobs, nch)."""
    channels, ROT, depol, tau, ssa, pmom, pe, he, te, u, v, m, sza, raa, vza = map(
        _f64, (channels, ROT, depol, tau, ssa, pmom, pe, he, te, U10m, V10m, mr, solar_zenith, relat_azymuth,
               sensor_zenith))
    km, nch, nobs = tau.shape
    nmom, npol = pmom.shape[3], pmom.shape[4]
    rad, refl, br, q, uu, brq, bru = (np.zeros((nobs, nch), order="F") for _ in range(7))
    nf = np.zeros((nobs, nch), dtype=np.int32, order="F")
    rc = np.zeros(1, dtype=np.int32)
    e = lib().aa_vlidort_vector_cx(km, nch, nobs, _ptr(channels), int(nstreams), int(bool(plane_parallel)), nmom,
                                   npol, _ptr(ROT), _ptr(depol), _ptr(tau), _ptr(ssa), _ptr(pmom), _ptr(pe),
                                   _ptr(he), _ptr(te), _ptr(u), _ptr(v), _ptr(m), _ptr(sza), _ptr(raa), _ptr(vza),
                                   float(MISSING), int(verbose), int(precision), _ptr(rad), _ptr(refl), _ptr(q),
                                   _ptr(uu), _ptr(br), _ptr(brq), _ptr(bru), _ptr(nf, _ip), _ptr(rc, _ip))
    check(e, "aa_vlidort_vector_cx")
    if return_nfourier:
        return rad, refl, br, q, uu, brq, bru, int(rc[0]), nf
    return rad, refl, br, q, uu, brq, bru, int(rc[0])


def vector_lambert_bpdf(channels, nstreams, plane_parallel, ROT, depol, tau, ssa, pmom, pe, he, te, albedo,
                        BPDFparam, solar_zenith, relat_azymuth, sensor_zenith, MISSING, verbose=0, precision=0,
                        return_nfourier=False):
    """VLIDORT_Vector_Lambert_BPDF (src/Components/rtms/geo_vlidort/VLIDORT_LAMBERT_BPDF.F90:23-222):
    Lambertian albedo (nobs, nch) plus the polarised BPDF-NDVI kernel,
    BPDFparam (3, nch, nobs) = (refractive index, NDVI, C).  Returns
    (radiance, reflectance, BR, Q, U, BR_Q, BR_U, rc), each array (nobs, nch)."""
    channels, ROT, depol, tau, ssa, pmom, pe, he, te, alb, bp, sza, raa, vza = map(
        _f64, (channels, ROT, depol, tau, ssa, pmom, pe, he, te, albedo, BPDFparam, solar_zenith, relat_azymuth,
               sensor_zenith))
    km, nch, nobs = tau.shape
    nmom, npol = pmom.shape[3], pmom.shape[4]
    rad, refl, br, q, uu, brq, bru = (np.zeros((nobs, nch), order="F") for _ in range(7))
    nf = np.zeros((nobs, nch), dtype=np.int32, order="F")
    rc = np.zeros(1, dtype=np.int32)
    e = lib().aa_vlidort_vector_lambert_bpdf(km, nch, nobs, _ptr(channels), int(nstreams), int(bool(plane_parallel)),
                                             nmom, npol, _ptr(ROT), _ptr(depol), _ptr(tau), _ptr(ssa), _ptr(pmom),
                                             _ptr(pe), _ptr(he), _ptr(te), _ptr(alb), _ptr(bp), _ptr(sza),
                                             _ptr(raa), _ptr(vza), float(MISSING), int(verbose), int(precision),
                                             _ptr(rad), _ptr(refl), _ptr(br), _ptr(q), _ptr(uu), _ptr(brq),
                                             _ptr(bru), _ptr(nf, _ip), _ptr(rc, _ip))
    check(e, "aa_vlidort_vector_lambert_bpdf")
    if return_nfourier:
        return rad, refl, br, q, uu, brq, bru, int(rc[0]), nf
    return rad, refl, br, q, uu, brq, bru, int(rc[0])


def vector_brdf_modis_bpdf(channels, nstreams, plane_parallel, ROT, depol, tau, ssa, pmom, pe, he, te, kernel_wt,
                           RTLSparam, BPDFparam, solar_zenith, relat_azymuth, sensor_zenith, MISSING, verbose=0,
                           precision=0, return_nfourier=False):
    """VLIDORT_Vector_LandMODIS_BPDF (src/Components/rtms/geo_vlidort/VLIDORT_BRDF_MODIS_BPDF.F90:24-222),
    the contract of the f2py ``VLIDORT_POLAR_.vector_brdf_modis_bpdf``
    (src/Components/rtms/leo_vlidort/VLIDORT_POLAR_py.F90:214-293): MODIS
    kernel_wt (3, nch, nobs), RTLSparam (2, nch, nobs) and the BPDF-NDVI
    kernel, BPDFparam (3, nch, nobs).  Returns (radiance, reflectance, BR, Q,
    U, BR_Q, BR_U, rc), each array (nobs, nch)."""
    channels, ROT, depol, tau, ssa, pmom, pe, he, te, kw, kp, bp, sza, raa, vza = map(
        _f64, (channels, ROT, depol, tau, ssa, pmom, pe, he, te, kernel_wt, RTLSparam, BPDFparam, solar_zenith,
               relat_azymuth, sensor_zenith))
    km, nch, nobs = tau.shape
    nmom, npol = pmom.shape[3], pmom.shape[4]
    rad, refl, br, q, uu, brq, bru = (np.zeros((nobs, nch), order="F") for _ in range(7))
    nf = np.zeros((nobs, nch), dtype=np.int32, order="F")
    rc = np.zeros(1, dtype=np.int32)
    e = lib().aa_vlidort_vector_landmodis_bpdf(km, nch, nobs, _ptr(channels), int(nstreams),
                                               int(bool(plane_parallel)), nmom, npol, _ptr(ROT), _ptr(depol),
                                               _ptr(tau), _ptr(ssa), _ptr(pmom), _ptr(pe), _ptr(he), _ptr(te),
                                               _ptr(kw), _ptr(kp), _ptr(bp), _ptr(sza), _ptr(raa), _ptr(vza),
                                               float(MISSING), int(verbose), int(precision), _ptr(rad), _ptr(refl),
                                               _ptr(br), _ptr(q), _ptr(uu), _ptr(brq), _ptr(bru), _ptr(nf, _ip),
                                               _ptr(rc, _ip))
    check(e, "aa_vlidort_vector_landmodis_bpdf")
    if return_nfourier:
        return rad, refl, br, q, uu, brq, bru, int(rc[0]), nf
    return rad, refl, br, q, uu, brq, bru, int(rc[0])


def vbrdf_bpdf(nstreams, sza, vza, raa, bpdf_param, albedo=None, kernel_wt=None, rtls_param=None):
    """The polarised supplement alone for ngeom geometries: VLIDORT_LAMBERTIAN_BPDF
    (albedo [ngeom]) or VLIDORT_LANDMODIS_BPDF (kernel_wt [ngeom, 3],
    rtls_param [ngeom, 2]) with bpdf_param [ngeom, 3].  Returns dict
    F[g, M, q, i, j], F0[g, M, q, i], UF[g, M, q, j], DB[g, q] (Mueller index
    q = 4 o1 + o2) and status[g]."""
    sza, vza, raa = (np.ascontiguousarray(np.atleast_1d(a), dtype=np.float64) for a in (sza, vza, raa))
    bp = np.ascontiguousarray(bpdf_param, dtype=np.float64).reshape(-1, 3)
    ng, n, nm = sza.size, int(nstreams), 2 * int(nstreams)
    lam = None if albedo is None else np.ascontiguousarray(np.atleast_1d(albedo), dtype=np.float64)
    kw = None if kernel_wt is None else np.ascontiguousarray(kernel_wt, dtype=np.float64).reshape(-1, 3)
    kp = None if rtls_param is None else np.ascontiguousarray(rtls_param, dtype=np.float64).reshape(-1, 2)
    F = np.zeros((nm, 16, n, n, ng), order="F")
    F0 = np.zeros((nm, 16, n, ng), order="F")
    UF = np.zeros((nm, 16, n, ng), order="F")
    db = np.zeros((16, ng), order="F")
    st = np.zeros(ng, dtype=np.int32)
    nul = ctypes.POINTER(ctypes.c_double)()
    e = lib().aa_vbrdf_bpdf(n, ng, _ptr(sza), _ptr(vza), _ptr(raa), nul if lam is None else _ptr(lam),
                            nul if kw is None else _ptr(kw), nul if kp is None else _ptr(kp), _ptr(bp), _ptr(F),
                            _ptr(F0), _ptr(UF), _ptr(db), _ptr(st, _ip))
    check(e, "aa_vbrdf_bpdf")
    return dict(F=np.moveaxis(F, 4, 0), F0=np.moveaxis(F0, 3, 0), UF=np.moveaxis(UF, 3, 0), DB=db.T.copy(),
                status=st)


def vbrdf_coxmunk(nstreams, sza, vza, raa, u10m, v10m, mr):
    """The Cox-Munk supplement alone for ngeom geometries (see vbrdf_modis)."""
    sza, vza, raa, u, v, m = (np.ascontiguousarray(np.atleast_1d(a), dtype=np.float64)
                              for a in (sza, vza, raa, u10m, v10m, mr))
    ng, n, nm = sza.size, int(nstreams), 2 * int(nstreams)
    F = np.zeros((nm, n, n, ng), order="F")
    F0 = np.zeros((nm, n, ng), order="F")
    UF = np.zeros((nm, n, ng), order="F")
    UF0 = np.zeros((nm, ng), order="F")
    db = np.zeros(ng)
    st = np.zeros(ng, dtype=np.int32)
    e = lib().aa_vbrdf_coxmunk(n, ng, _ptr(sza), _ptr(vza), _ptr(raa), _ptr(u), _ptr(v), _ptr(m), _ptr(F), _ptr(F0),
                               _ptr(UF), _ptr(UF0), _ptr(db), _ptr(st, _ip))
    check(e, "aa_vbrdf_coxmunk")
    return dict(F=np.moveaxis(F, 3, 0), F0=np.moveaxis(F0, 2, 0), UF=np.moveaxis(UF, 2, 0), UF0=UF0.T.copy(), db=db,
                status=st)


def vbrdf_modis(nstreams, sza, vza, raa, kernel_wt, param):
    """The MODIS BRDF supplement alone for ngeom geometries (degrees), with
    kernel_wt[3, ngeom] and param[2, ngeom]: returns a dict of F[g, M, i, j],
    F0[g, M, i], UF[g, M, j], UF0[g, M], db[g], status[g]."""
    sza, vza, raa = (np.ascontiguousarray(np.atleast_1d(a), dtype=np.float64) for a in (sza, vza, raa))
    kw, par = _f64(kernel_wt), _f64(param)
    ng, n, nm = sza.size, int(nstreams), 2 * int(nstreams)
    F = np.zeros((nm, n, n, ng), order="F")
    F0 = np.zeros((nm, n, ng), order="F")
    UF = np.zeros((nm, n, ng), order="F")
    UF0 = np.zeros((nm, ng), order="F")
    db = np.zeros(ng)
    st = np.zeros(ng, dtype=np.int32)
    e = lib().aa_vbrdf_modis(n, ng, _ptr(sza), _ptr(vza), _ptr(raa), _ptr(kw), _ptr(par), _ptr(F), _ptr(F0),
                             _ptr(UF), _ptr(UF0), _ptr(db), _ptr(st, _ip))
    check(e, "aa_vbrdf_modis")
    return dict(F=np.moveaxis(F, 3, 0), F0=np.moveaxis(F0, 2, 0), UF=np.moveaxis(UF, 2, 0), UF0=UF0.T.copy(), db=db,
                status=st)


def rot_calc(channels, pe, he, te, MISSING, verbose=0):
    """ROT_CALC (src/Components/rtms/leo_vlidort/VLIDORT_POLAR_py.F90:9-41 ->
    VLIDORT_ROT_CALC, src/Components/rtms/geo_vlidort/VLIDORT_ROT.F90:19-85):
    Rayleigh optical thickness ROT[km, nobs, nch] and depolarisation ratio
    depol[nch] from edge pressure/height/temperature pe, he, te [km+1, nobs]
    and channels [nm].  Returns (ROT, depol, rc)."""
    channels, pe, he, te = map(_f64, (channels, pe, he, te))
    if pe.ndim == 1:
        pe, he, te = (_f64(a[:, None]) for a in (pe, he, te))
    km, nobs = pe.shape[0] - 1, pe.shape[1]
    nch = channels.size
    rot = np.zeros((km, nobs, nch), order="F")
    depol = np.zeros(nch)
    rc = np.zeros(1, dtype=np.int32)
    e = lib().aa_rot_calc(km, nch, nobs, _ptr(channels), _ptr(pe), _ptr(he), _ptr(te), float(MISSING),
                          int(verbose), _ptr(rot), _ptr(depol), _ptr(rc, _ip))
    if e != -4:
        check(e, "aa_rot_calc")
    return rot, depol, int(rc[0])
