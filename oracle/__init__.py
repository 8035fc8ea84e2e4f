"""ORACLE TEST INFRASTRUCTURE -- not part of the product.

CPU checkers for the aerosol-optics + discrete-ordinate radiance path:

* ``oracle.chem_mie``   numpy float32 restatement of the Chem_Mie query and the
  per-tracer optics summation;
* ``oracle.ref``        the reference VLIDORT 2.8.3 core compiled from its
  unmodified sources by ``oracle/Makefile.ref`` (into ``oracle/_ref/``) and
  driven through ``oracle/ref_driver.f90`` (a statement-by-statement harness
  of the AeroApps drivers that also returns the Fourier-term counts);
* ``oracle.drivers``    the reference's own AeroApps driver layer
  (VLIDORT_LAMBERT, VLIDORT_ROT, VLIDORT_BRDF_MODIS, VLIDORT_BRDF_CX over
  VLIDORT_ScatMod / VLIDORT_SurfaceMod / VLIDORT_Mod) compiled into
  ``oracle/_ref/libvlidort_drv.so`` and called through ``oracle/ref_calls.f90``;
  the harness in the same library must reproduce it bit for bit;
* ``oracle.port``       the C++ restatement of the same path
  (``oracle/vlidort_oracle.cpp``, built by ``oracle/Makefile``).

Only tests/, ``__graft_entry__.smoke()`` and the bench's ``cpu_baseline`` leg may
import this package, and only as the checker.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_SO = os.path.join(HERE, "_ref", "libvlidort_ref.so")
DRV_SO = os.path.join(HERE, "_ref", "libvlidort_drv.so")
PORT_SO = os.path.join(HERE, "liboracle_vlidort.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _p(a: np.ndarray, dt=np.float64):
    assert a.dtype == dt and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(_dp if dt == np.float64 else _ip)


def _big_stack_call(fn, *args):
    """Run fn(*args) on a thread with a 1 GiB stack: the flang-built reference
    keeps its large work arrays (e.g. 9 MB Greek-matrix inputs) on the stack."""
    import threading
    box = {}

    def run():
        try:
            box["r"] = fn(*args)
        except BaseException as e:  # re-raised in the caller
            box["e"] = e

    old = threading.stack_size()
    threading.stack_size(1 << 30)
    try:
        t = threading.Thread(target=run)
        t.start()
    finally:
        threading.stack_size(old)
    t.join()
    if "e" in box:
        raise box["e"]
    return box.get("r")


def _load(path: str) -> Optional[ctypes.CDLL]:
    if not os.path.exists(path):
        return None
    return ctypes.CDLL(path)


class LambertDriver:
    """ctypes binding of a library exporting the Fortran-layout drivers
    ``*_scalar_lambert`` / ``*_vector_lambert`` (see ref_driver.f90)."""

    def __init__(self, lib: ctypes.CDLL, prefix: str):
        self.lib = lib
        self.fs = getattr(lib, prefix + "_scalar_lambert")
        self.fv = getattr(lib, prefix + "_vector_lambert")
        ci = ctypes.c_int
        self.fs.restype = None
        self.fs.argtypes = [ci, ci, ci, _dp, ci, ci, ci, ci, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                            _dp, _dp, _dp, _dp, ctypes.c_double, _dp, _dp, _ip, _ip]
        self.fv.restype = None
        self.fv.argtypes = [ci, ci, ci, _dp, ci, ci, ci, ci, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                            _dp, _dp, _dp, _dp, _dp, ctypes.c_double, _dp, _dp, _dp, _dp, _ip, _ip]

    def scalar_lambert(self, km, channels, nstreams, nmom, rot, depol, tau, ssa, g, pmom, pe, he, te,
                       albedo, sza, raa, vza, missing=1e15, plane_parallel=False):
        """Arrays in reversed-Fortran C order: rot[nch,nobs,km], tau/ssa/g[nobs,nch,km],
        pmom[npol,nmom,nobs,nch,km], pe/he/te[nobs,km+1], albedo[nch,nobs]."""
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        channels, rot, depol, tau, ssa, g, pmom, pe, he, te, albedo, sza, raa, vza = map(
            f, (channels, rot, depol, tau, ssa, g, pmom, pe, he, te, albedo, sza, raa, vza))
        nch = len(channels)
        nobs = len(sza)
        npol = pmom.shape[0]
        rad = np.zeros((nch, nobs))
        refl = np.zeros((nch, nobs))
        nf = np.zeros((nch, nobs), dtype=np.int32)
        rc = np.zeros(1, dtype=np.int32)
        _big_stack_call(self.fs, km, nch, nobs, _p(channels), nstreams, int(plane_parallel), nmom, npol, _p(rot), _p(depol),
                _p(tau), _p(ssa), _p(g), _p(pmom), _p(pe), _p(he), _p(te), _p(albedo), _p(sza), _p(raa),
                _p(vza), missing, _p(rad), _p(refl), _p(nf, np.int32), _p(rc, np.int32))
        return rad, refl, nf, int(rc[0])

    def vector_lambert(self, km, channels, nstreams, nmom, rot, depol, alpha, tau, ssa, pmom, pe, he, te,
                       albedo, sza, raa, vza, flux_factor, missing=1e15, plane_parallel=False):
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        (channels, rot, depol, alpha, tau, ssa, pmom, pe, he, te, albedo, sza, raa, vza,
         flux_factor) = map(f, (channels, rot, depol, alpha, tau, ssa, pmom, pe, he, te, albedo, sza, raa,
                                vza, flux_factor))
        nch = len(channels)
        nobs = len(sza)
        npol = pmom.shape[0]
        out = [np.zeros((nch, nobs)) for _ in range(4)]
        nf = np.zeros((nch, nobs), dtype=np.int32)
        rc = np.zeros(1, dtype=np.int32)
        _big_stack_call(self.fv, km, nch, nobs, _p(channels), nstreams, int(plane_parallel), nmom, npol, _p(rot), _p(depol),
                _p(alpha), _p(tau), _p(ssa), _p(pmom), _p(pe), _p(he), _p(te), _p(albedo), _p(sza), _p(raa),
                _p(vza), _p(flux_factor), missing, *[_p(o) for o in out], _p(nf, np.int32), _p(rc, np.int32))
        return out[0], out[1], out[2], out[3], nf, int(rc[0])


class ModisRef:
    """ctypes binding of the reference harness's MODIS land-surface routines
    (``ref_vbrdf_modis`` / ``ref_scalar_landmodis`` in ref_driver.f90)."""

    def __init__(self, lib: ctypes.CDLL):
        ci, cd = ctypes.c_int, ctypes.c_double
        self.fb = lib.ref_vbrdf_modis
        self.fb.restype = None
        self.fb.argtypes = [ci] + [cd] * 8 + [_dp] * 6 + [_ip]
        self.fl = lib.ref_scalar_landmodis
        self.fl.restype = None
        self.fl.argtypes = [ci, ci, ci, _dp, ci, ci, ci, ci, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                            _dp, _dp, _dp, _dp, _dp, cd, _dp, _dp, _dp, _ip, _ip]
        self.fbc = lib.ref_vbrdf_coxmunk
        self.fbc.restype = None
        self.fbc.argtypes = [ci] + [cd] * 6 + [_dp] * 6 + [_ip]
        self.fc = lib.ref_scalar_cx
        self.fc.restype = None
        self.fc.argtypes = [ci, ci, ci, _dp, ci, ci, ci, ci, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                            _dp, _dp, _dp, _dp, _dp, _dp, cd, _dp, _dp, _dp, _ip, _ip]
        self.fvc = lib.ref_vector_cx
        self.fvc.restype = None
        self.fvc.argtypes = [ci, ci, ci, _dp, ci, ci, ci, ci] + [_dp] * 14 + [cd] + [_dp] * 7 + [_ip, _ip]
        self.fbp = lib.ref_vbrdf_bpdf
        self.fbp.restype = None
        self.fbp.argtypes = [ci, cd, cd, cd, ci] + [cd] * 8 + [_dp] * 5 + [_ip]
        self.fvp = lib.ref_vector_bpdf
        self.fvp.restype = None
        self.fvp.argtypes = [ci, ci, ci, _dp, ci, ci, ci, ci] + [_dp] * 8 + [ci] + [_dp] * 7 + [cd] + [_dp] * 7 \
            + [_ip, _ip]

    def vector_cx(self, km, channels, nstreams, nmom, rot, depol, tau, ssa, pmom, pe, he, te, u10m, v10m, mr,
                  sza, raa, vza, missing=1e15, plane_parallel=False):
        """VLIDORT_Vector_CX_SingleGeom: arrays as for scalar_cx without g (pmom with 6
        polarisations); returns rad, refl, Q, U, BR, BR_Q, BR_U [nch, nobs], nf, rc."""
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        channels, rot, depol, tau, ssa, pmom, pe, he, te, u10m, v10m, mr, sza, raa, vza = map(
            f, (channels, rot, depol, tau, ssa, pmom, pe, he, te, u10m, v10m, mr, sza, raa, vza))
        nch, nobs, npol = len(channels), len(sza), pmom.shape[0]
        out = [np.zeros((nch, nobs)) for _ in range(7)]
        nf = np.zeros((nch, nobs), dtype=np.int32)
        rc = np.zeros(1, dtype=np.int32)
        _big_stack_call(self.fvc, km, nch, nobs, _p(channels), nstreams, int(plane_parallel), nmom, npol,
                        _p(rot), _p(depol), _p(tau), _p(ssa), _p(pmom), _p(pe), _p(he), _p(te), _p(u10m),
                        _p(v10m), _p(mr), _p(sza), _p(raa), _p(vza), missing, *[_p(o) for o in out],
                        _p(nf, np.int32), _p(rc, np.int32))
        return (*out, nf, int(rc[0]))

    def vbrdf_bpdf(self, nstreams, sza, vza, raa, modis, fiso, fgeo, fvol, p1, p2, b1, b2, b3):
        """Polarised supplement (VLIDORT_LANDMODIS_BPDF when modis, else
        VLIDORT_LAMBERTIAN_BPDF with albedo fiso) for one geometry: dict
        F[M,q,i,j], F0[M,q,i], UF[M,q,j] and DB[q] over the 16 Mueller indices
        q = 4 o1 + o2 (0-based), wsa, status."""
        n, nm = nstreams, 2 * nstreams
        F, F0, UF = np.zeros((n, n, 16, nm)), np.zeros((n, 16, nm)), np.zeros((n, 16, nm))
        db, wsa = np.zeros(16), np.zeros(1)
        st = np.zeros(1, dtype=np.int32)
        _big_stack_call(self.fbp, n, sza, vza, raa, int(modis), fiso, fgeo, fvol, p1, p2, b1, b2, b3, _p(F),
                        _p(F0), _p(UF), _p(db), _p(wsa), _p(st, np.int32))
        return dict(F=F.transpose(3, 2, 1, 0), F0=F0.transpose(2, 1, 0), UF=UF.transpose(2, 1, 0), DB=db,
                    wsa=float(wsa[0]), status=int(st[0]))

    def vector_bpdf(self, km, channels, nstreams, nmom, rot, depol, tau, ssa, pmom, pe, he, te, modis, albedo,
                    kernel_wt, param, bpdf, sza, raa, vza, missing=1e15, plane_parallel=False):
        """VLIDORT_Vector_LandMODIS_BPDF (modis) or VLIDORT_Vector_Lambert_BPDF:
        kernel_wt [nobs, nch, 3], param [nobs, nch, 2], bpdf [nobs, nch, 3]
        (C order of the Fortran (3, nch, nobs) arrays), albedo [nch, nobs];
        returns rad, refl, Q, U, BR, BR_Q, BR_U [nch, nobs], nf, rc."""
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        channels, rot, depol, tau, ssa, pmom, pe, he, te, albedo, kernel_wt, param, bpdf, sza, raa, vza = map(
            f, (channels, rot, depol, tau, ssa, pmom, pe, he, te, albedo, kernel_wt, param, bpdf, sza, raa, vza))
        nch, nobs, npol = len(channels), len(sza), pmom.shape[0]
        out = [np.zeros((nch, nobs)) for _ in range(7)]
        nf = np.zeros((nch, nobs), dtype=np.int32)
        rc = np.zeros(1, dtype=np.int32)
        _big_stack_call(self.fvp, km, nch, nobs, _p(channels), nstreams, int(plane_parallel), nmom, npol,
                        _p(rot), _p(depol), _p(tau), _p(ssa), _p(pmom), _p(pe), _p(he), _p(te), int(modis),
                        _p(albedo), _p(kernel_wt), _p(param), _p(bpdf), _p(sza), _p(raa), _p(vza), missing,
                        *[_p(o) for o in out], _p(nf, np.int32), _p(rc, np.int32))
        return (*out, nf, int(rc[0]))

    def vbrdf_coxmunk(self, nstreams, sza, vza, raa, u10m, v10m, m):
        n, nm = nstreams, 2 * nstreams
        F, F0, UF, UF0 = np.zeros((n, n, nm)), np.zeros((n, nm)), np.zeros((n, nm)), np.zeros(nm)
        db, wsa = np.zeros(1), np.zeros(1)
        st = np.zeros(1, dtype=np.int32)
        _big_stack_call(self.fbc, n, sza, vza, raa, u10m, v10m, m, _p(F), _p(F0), _p(UF), _p(UF0), _p(db),
                        _p(wsa), _p(st, np.int32))
        return dict(F=F.transpose(2, 1, 0), F0=F0.T, UF=UF.T, UF0=UF0, db=float(db[0]), wsa=float(wsa[0]),
                    status=int(st[0]))

    def scalar_cx(self, km, channels, nstreams, nmom, rot, depol, tau, ssa, g, pmom, pe, he, te, u10m, v10m, mr,
                  sza, raa, vza, missing=1e15, plane_parallel=False):
        """VLIDORT_Scalar_CX: as scalar_landmodis with u10m, v10m [nobs], mr [nch]."""
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        channels, rot, depol, tau, ssa, g, pmom, pe, he, te, u10m, v10m, mr, sza, raa, vza = map(
            f, (channels, rot, depol, tau, ssa, g, pmom, pe, he, te, u10m, v10m, mr, sza, raa, vza))
        nch, nobs, npol = len(channels), len(sza), pmom.shape[0]
        rad, refl, br = (np.zeros((nch, nobs)) for _ in range(3))
        nf = np.zeros((nch, nobs), dtype=np.int32)
        rc = np.zeros(1, dtype=np.int32)
        _big_stack_call(self.fc, km, nch, nobs, _p(channels), nstreams, int(plane_parallel), nmom, npol, _p(rot),
                        _p(depol), _p(tau), _p(ssa), _p(g), _p(pmom), _p(pe), _p(he), _p(te), _p(u10m), _p(v10m),
                        _p(mr), _p(sza), _p(raa), _p(vza), missing, _p(rad), _p(refl), _p(br), _p(nf, np.int32),
                        _p(rc, np.int32))
        return rad, refl, br, nf, int(rc[0])

    def vbrdf_modis(self, nstreams, sza, vza, raa, fiso, fgeo, fvol, p1, p2):
        """Dict F[M,i,j], F0[M,i], UF[M,j], UF0[M], db, wsa, status for one geometry."""
        n, nm = nstreams, 2 * nstreams
        F = np.zeros((n, n, nm))        # reversed Fortran F(0:2N-1, N, N)
        F0 = np.zeros((n, nm))
        UF = np.zeros((n, nm))
        UF0 = np.zeros(nm)
        db, wsa = np.zeros(1), np.zeros(1)
        st = np.zeros(1, dtype=np.int32)
        _big_stack_call(self.fb, n, sza, vza, raa, fiso, fgeo, fvol, p1, p2, _p(F), _p(F0), _p(UF), _p(UF0),
                        _p(db), _p(wsa), _p(st, np.int32))
        return dict(F=F.transpose(2, 1, 0), F0=F0.T, UF=UF.T, UF0=UF0, db=float(db[0]), wsa=float(wsa[0]),
                    status=int(st[0]))

    def scalar_landmodis(self, km, channels, nstreams, nmom, rot, depol, tau, ssa, g, pmom, pe, he, te,
                         kernel_wt, param, sza, raa, vza, missing=1e15, plane_parallel=False):
        """Arrays in reversed-Fortran C order as for scalar_lambert, plus
        kernel_wt[nobs,nch,3], param[nobs,nch,2]; returns rad, refl, BR [nch,nobs], nf, rc."""
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        channels, rot, depol, tau, ssa, g, pmom, pe, he, te, kernel_wt, param, sza, raa, vza = map(
            f, (channels, rot, depol, tau, ssa, g, pmom, pe, he, te, kernel_wt, param, sza, raa, vza))
        nch = len(channels)
        nobs = len(sza)
        npol = pmom.shape[0]
        rad, refl, br = (np.zeros((nch, nobs)) for _ in range(3))
        nf = np.zeros((nch, nobs), dtype=np.int32)
        rc = np.zeros(1, dtype=np.int32)
        _big_stack_call(self.fl, km, nch, nobs, _p(channels), nstreams, int(plane_parallel), nmom, npol, _p(rot),
                        _p(depol), _p(tau), _p(ssa), _p(g), _p(pmom), _p(pe), _p(he), _p(te), _p(kernel_wt),
                        _p(param), _p(sza), _p(raa), _p(vza), missing, _p(rad), _p(refl), _p(br),
                        _p(nf, np.int32), _p(rc, np.int32))
        return rad, refl, br, nf, int(rc[0])


class RefDrivers:
    """ctypes binding of ``oracle/ref_calls.f90``: the reference's own driver
    routines.  Arrays in reversed-Fortran C order as for LambertDriver.
    Returns outputs in [nch, nobs] (MultiGeom: [ngeom, nch, nobs]) order."""

    def __init__(self, lib: ctypes.CDLL):
        ci, cd = ctypes.c_int, ctypes.c_double
        self.lib = lib
        self.fs = lib.drv_scalar_lambert
        self.fs.restype = None
        self.fs.argtypes = [ci, ci, ci, _dp, ci, ci, ci, ci] + [_dp] * 13 + [cd, _dp, _dp, _ip]
        self.fv = lib.drv_vector_lambert
        self.fv.restype = None
        self.fv.argtypes = [ci, ci, ci, _dp, ci, ci, ci, ci] + [_dp] * 14 + [cd, ci, ci] + [_dp] * 4 + [_ip]
        self.fm = lib.drv_vector_lambert_mg
        self.fm.restype = None
        self.fm.argtypes = [ci, ci, ci, ci, _dp, ci, ci, ci, ci] + [_dp] * 14 + [cd, ci, ci] + [_dp] * 4 + [_ip]
        self.fr = lib.drv_rot_calc
        self.fr.restype = None
        self.fr.argtypes = [ci, ci, ci, _dp, _dp, _dp, _dp, cd, _dp, _dp, _ip]
        self.fd = lib.drv_coef_depol
        self.fd.restype = cd
        self.fd.argtypes = [cd]
        self.fl = lib.drv_scalar_landmodis
        self.fl.restype = None
        self.fl.argtypes = [ci, ci, ci, _dp, ci, ci, ci, ci] + [_dp] * 14 + [cd, _dp, _dp, _dp, _ip]
        self.fc = lib.drv_scalar_cx
        self.fc.restype = None
        self.fc.argtypes = [ci, ci, ci, _dp, ci, ci, ci, ci] + [_dp] * 15 + [cd, _dp, _dp, _dp, _ip]
        self.fvc = lib.drv_vector_cx
        self.fvc.restype = None
        self.fvc.argtypes = [ci, ci, ci, _dp, ci, ci, ci, ci] + [_dp] * 14 + [cd] + [_dp] * 7 + [_ip]
        self.fvlb = lib.drv_vector_lambert_bpdf
        self.fvlb.restype = None
        self.fvlb.argtypes = [ci, ci, ci, _dp, ci, ci, ci, ci] + [_dp] * 13 + [cd] + [_dp] * 7 + [_ip]
        self.fvmb = lib.drv_vector_landmodis_bpdf
        self.fvmb.restype = None
        self.fvmb.argtypes = [ci, ci, ci, _dp, ci, ci, ci, ci] + [_dp] * 14 + [cd] + [_dp] * 7 + [_ip]

    @staticmethod
    def _f(*arrs):
        return [np.ascontiguousarray(a, dtype=np.float64) for a in arrs]

    def scalar_lambert(self, km, channels, nstreams, nmom, rot, depol, tau, ssa, g, pmom, pe, he, te,
                       albedo, sza, raa, vza, missing=1e15, plane_parallel=False):
        channels, rot, depol, tau, ssa, g, pmom, pe, he, te, albedo, sza, raa, vza = self._f(
            channels, rot, depol, tau, ssa, g, pmom, pe, he, te, albedo, sza, raa, vza)
        nch, nobs, npol = len(channels), len(sza), pmom.shape[0]
        rad, refl = np.zeros((nch, nobs)), np.zeros((nch, nobs))
        rc = np.zeros(1, dtype=np.int32)
        _big_stack_call(self.fs, km, nch, nobs, _p(channels), nstreams, int(plane_parallel), nmom, npol,
                        _p(rot), _p(depol), _p(tau), _p(ssa), _p(g), _p(pmom), _p(pe), _p(he), _p(te),
                        _p(albedo), _p(sza), _p(raa), _p(vza), missing, _p(rad), _p(refl), _p(rc, np.int32))
        return rad, refl, int(rc[0])

    def vector_lambert(self, km, channels, nstreams, nmom, rot, depol, alpha, tau, ssa, pmom, pe, he, te,
                       albedo, sza, raa, vza, flux_factor, missing=1e15, plane_parallel=False,
                       do_2os=False, do_boa=False):
        (channels, rot, depol, alpha, tau, ssa, pmom, pe, he, te, albedo, sza, raa, vza,
         flux_factor) = self._f(channels, rot, depol, alpha, tau, ssa, pmom, pe, he, te, albedo, sza, raa,
                                vza, flux_factor)
        nch, nobs, npol = len(channels), len(sza), pmom.shape[0]
        out = [np.zeros((nch, nobs)) for _ in range(4)]
        rc = np.zeros(1, dtype=np.int32)
        _big_stack_call(self.fv, km, nch, nobs, _p(channels), nstreams, int(plane_parallel), nmom, npol,
                        _p(rot), _p(depol), _p(alpha), _p(tau), _p(ssa), _p(pmom), _p(pe), _p(he), _p(te),
                        _p(albedo), _p(sza), _p(raa), _p(vza), _p(flux_factor), missing, int(do_2os),
                        int(do_boa), *[_p(o) for o in out], _p(rc, np.int32))
        return out[0], out[1], out[2], out[3], int(rc[0])

    def vector_lambert_mg(self, km, channels, nstreams, nmom, rot, depol, alpha, tau, ssa, pmom, pe, he, te,
                          albedo, sza, raa, vza, flux_factor, missing=1e15, plane_parallel=False,
                          do_2os=False, do_boa=False):
        """MultiGeom: albedo[ngeom, nch, nobs], sza/raa/vza[ngeom, nobs]."""
        (channels, rot, depol, alpha, tau, ssa, pmom, pe, he, te, albedo, sza, raa, vza,
         flux_factor) = self._f(channels, rot, depol, alpha, tau, ssa, pmom, pe, he, te, albedo, sza, raa,
                                vza, flux_factor)
        nch, npol = len(channels), pmom.shape[0]
        ngeom, nobs = sza.shape
        out = [np.zeros((ngeom, nch, nobs)) for _ in range(4)]
        rc = np.zeros(1, dtype=np.int32)
        _big_stack_call(self.fm, km, nch, nobs, ngeom, _p(channels), nstreams, int(plane_parallel), nmom, npol,
                        _p(rot), _p(depol), _p(alpha), _p(tau), _p(ssa), _p(pmom), _p(pe), _p(he), _p(te),
                        _p(albedo), _p(sza), _p(raa), _p(vza), _p(flux_factor), missing, int(do_2os),
                        int(do_boa), *[_p(o) for o in out], _p(rc, np.int32))
        return out[0], out[1], out[2], out[3], int(rc[0])

    def rot_calc(self, km, channels, pe, he, te, missing=1e15):
        """VLIDORT_ROT_CALC: returns ROT[nch, nobs, km], depol[nch], rc."""
        channels, pe, he, te = self._f(channels, pe, he, te)
        nch, nobs = len(channels), pe.shape[0]
        rot = np.zeros((nch, nobs, km))
        depol = np.zeros(nch)
        rc = np.zeros(1, dtype=np.int32)
        _big_stack_call(self.fr, km, nch, nobs, _p(channels), _p(pe), _p(he), _p(te), missing, _p(rot),
                        _p(depol), _p(rc, np.int32))
        return rot, depol, int(rc[0])

    def coef_depol(self, wmicron: float) -> float:
        return float(self.fd(float(wmicron)))

    def scalar_landmodis(self, km, channels, nstreams, nmom, rot, depol, tau, ssa, g, pmom, pe, he, te,
                         kernel_wt, param, sza, raa, vza, missing=1e15, plane_parallel=False):
        channels, rot, depol, tau, ssa, g, pmom, pe, he, te, kernel_wt, param, sza, raa, vza = self._f(
            channels, rot, depol, tau, ssa, g, pmom, pe, he, te, kernel_wt, param, sza, raa, vza)
        nch, nobs, npol = len(channels), len(sza), pmom.shape[0]
        rad, refl, br = (np.zeros((nch, nobs)) for _ in range(3))
        rc = np.zeros(1, dtype=np.int32)
        _big_stack_call(self.fl, km, nch, nobs, _p(channels), nstreams, int(plane_parallel), nmom, npol,
                        _p(rot), _p(depol), _p(tau), _p(ssa), _p(g), _p(pmom), _p(pe), _p(he), _p(te),
                        _p(kernel_wt), _p(param), _p(sza), _p(raa), _p(vza), missing, _p(rad), _p(refl),
                        _p(br), _p(rc, np.int32))
        return rad, refl, br, int(rc[0])

    def scalar_cx(self, km, channels, nstreams, nmom, rot, depol, tau, ssa, g, pmom, pe, he, te, u10m, v10m, mr,
                  sza, raa, vza, missing=1e15, plane_parallel=False):
        channels, rot, depol, tau, ssa, g, pmom, pe, he, te, u10m, v10m, mr, sza, raa, vza = self._f(
            channels, rot, depol, tau, ssa, g, pmom, pe, he, te, u10m, v10m, mr, sza, raa, vza)
        nch, nobs, npol = len(channels), len(sza), pmom.shape[0]
        rad, refl, br = (np.zeros((nch, nobs)) for _ in range(3))
        rc = np.zeros(1, dtype=np.int32)
        _big_stack_call(self.fc, km, nch, nobs, _p(channels), nstreams, int(plane_parallel), nmom, npol,
                        _p(rot), _p(depol), _p(tau), _p(ssa), _p(g), _p(pmom), _p(pe), _p(he), _p(te),
                        _p(u10m), _p(v10m), _p(mr), _p(sza), _p(raa), _p(vza), missing, _p(rad), _p(refl),
                        _p(br), _p(rc, np.int32))
        return rad, refl, br, int(rc[0])


    def vector_cx(self, km, channels, nstreams, nmom, rot, depol, tau, ssa, pmom, pe, he, te, u10m, v10m, mr,
                  sza, raa, vza, missing=1e15, plane_parallel=False):
        """Returns rad, refl, Q, U, BR, BR_Q, BR_U [nch, nobs], rc."""
        channels, rot, depol, tau, ssa, pmom, pe, he, te, u10m, v10m, mr, sza, raa, vza = self._f(
            channels, rot, depol, tau, ssa, pmom, pe, he, te, u10m, v10m, mr, sza, raa, vza)
        nch, nobs, npol = len(channels), len(sza), pmom.shape[0]
        out = [np.zeros((nch, nobs)) for _ in range(7)]
        rc = np.zeros(1, dtype=np.int32)
        _big_stack_call(self.fvc, km, nch, nobs, _p(channels), nstreams, int(plane_parallel), nmom, npol,
                        _p(rot), _p(depol), _p(tau), _p(ssa), _p(pmom), _p(pe), _p(he), _p(te), _p(u10m),
                        _p(v10m), _p(mr), _p(sza), _p(raa), _p(vza), missing, *[_p(o) for o in out],
                        _p(rc, np.int32))
        return (*out, int(rc[0]))

    def vector_lambert_bpdf(self, km, channels, nstreams, nmom, rot, depol, tau, ssa, pmom, pe, he, te, albedo,
                            bpdf, sza, raa, vza, missing=1e15, plane_parallel=False):
        """VLIDORT_Vector_Lambert_BPDF; returns rad, refl, Q, U, BR, BR_Q, BR_U [nch, nobs], rc."""
        channels, rot, depol, tau, ssa, pmom, pe, he, te, albedo, bpdf, sza, raa, vza = self._f(
            channels, rot, depol, tau, ssa, pmom, pe, he, te, albedo, bpdf, sza, raa, vza)
        nch, nobs, npol = len(channels), len(sza), pmom.shape[0]
        rad, refl, br, q, u, bq, bu = (np.zeros((nch, nobs)) for _ in range(7))
        rc = np.zeros(1, dtype=np.int32)
        _big_stack_call(self.fvlb, km, nch, nobs, _p(channels), nstreams, int(plane_parallel), nmom, npol,
                        _p(rot), _p(depol), _p(tau), _p(ssa), _p(pmom), _p(pe), _p(he), _p(te), _p(albedo),
                        _p(bpdf), _p(sza), _p(raa), _p(vza), missing, _p(rad), _p(refl), _p(br), _p(q), _p(u),
                        _p(bq), _p(bu), _p(rc, np.int32))
        return rad, refl, q, u, br, bq, bu, int(rc[0])

    def vector_landmodis_bpdf(self, km, channels, nstreams, nmom, rot, depol, tau, ssa, pmom, pe, he, te,
                              kernel_wt, param, bpdf, sza, raa, vza, missing=1e15, plane_parallel=False):
        """VLIDORT_Vector_LandMODIS_BPDF; returns rad, refl, Q, U, BR, BR_Q, BR_U [nch, nobs], rc."""
        channels, rot, depol, tau, ssa, pmom, pe, he, te, kernel_wt, param, bpdf, sza, raa, vza = self._f(
            channels, rot, depol, tau, ssa, pmom, pe, he, te, kernel_wt, param, bpdf, sza, raa, vza)
        nch, nobs, npol = len(channels), len(sza), pmom.shape[0]
        rad, refl, br, q, u, bq, bu = (np.zeros((nch, nobs)) for _ in range(7))
        rc = np.zeros(1, dtype=np.int32)
        _big_stack_call(self.fvmb, km, nch, nobs, _p(channels), nstreams, int(plane_parallel), nmom, npol,
                        _p(rot), _p(depol), _p(tau), _p(ssa), _p(pmom), _p(pe), _p(he), _p(te), _p(kernel_wt),
                        _p(param), _p(bpdf), _p(sza), _p(raa), _p(vza), missing, _p(rad), _p(refl), _p(br),
                        _p(q), _p(u), _p(bq), _p(bu), _p(rc, np.int32))
        return rad, refl, q, u, br, bq, bu, int(rc[0])


_DRV_LIB = None


def _drv_lib() -> Optional[ctypes.CDLL]:
    global _DRV_LIB
    if _DRV_LIB is None:
        _DRV_LIB = _load(DRV_SO)
    return _DRV_LIB


def drivers() -> Optional[RefDrivers]:
    """The reference's own driver layer (built here only; absent on a GPU box)."""
    lib = _drv_lib()
    return None if lib is None else RefDrivers(lib)


def drv_harness() -> Optional[LambertDriver]:
    """The statement-by-statement harness linked into the driver-layer library."""
    lib = _drv_lib()
    return None if lib is None else LambertDriver(lib, "ref")


def drv_harness_modis() -> Optional["ModisRef"]:
    lib = _drv_lib()
    return None if lib is None else ModisRef(lib)


def ref_modis() -> Optional[ModisRef]:
    lib = _load(REF_SO)
    return None if lib is None or not hasattr(lib, "ref_vbrdf_coxmunk") else ModisRef(lib)


def ref() -> Optional[LambertDriver]:
    lib = _load(REF_SO)
    return None if lib is None else LambertDriver(lib, "ref")


def port() -> Optional[LambertDriver]:
    lib = _load(PORT_SO)
    return None if lib is None else LambertDriver(lib, "oracle")
