"""Loader of the in-tree HIP library ``libaeroapps_amd.so`` (ctypes).

The product has no CPU fallback: when the library is missing or cannot be
loaded, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libaeroapps_amd.so")

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_ci = ctypes.c_int
_cd = ctypes.c_double
_cf = ctypes.c_float
_cp = ctypes.c_char_p

# (name, restype, argtypes) of every symbol include/aeroapps_amd.h declares
SIGNATURES = {
    "aa_mie_create": (_ci, [ctypes.POINTER(_vp), _ci, _fp]),
    "aa_mie_add_table": (_ci, [_vp, _ci, _ci, _ci, _ci, _fp, _fp, _fp, _fp, _fp, _fp, _fp, _fp, _cp]),
    "aa_mie_add_table_native": (_ci, [_vp, _ci, _dp, _ci, _ci, _ci, _ci, _ci, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                      _dp, _dp, _dp, _dp, _dp, _cp, _ip]),
    "aa_mie_query_extra": (_ci, [_vp, _ci, _cf, _ci, _fp, _fp, _fp, _fp, _fp, _fp, _fp, _fp]),
    "aa_mie_query_idx": (_ci, [_vp, _cp]),
    "aa_mie_query": (_ci, [_vp, _ci, _cf, _ci, _fp, _fp, _fp, _fp, _fp, _fp, _fp, _fp, _fp, _fp, _ci, _ci, _fp]),
    "aa_mie_dims": (_ci, [_vp, _ip, _ip, _ip, _ip]),
    "aa_mie_channels": (_ci, [_vp, _fp]),
    "aa_mie_tracer_table": (_ci, [_vp, _ci, _ip, _ip, _ip, _ip, _ip, _ip]),
    "aa_mie_registry": (_ci, [_cp, _ci, _ip, _cp, _ip, _ip]),
    "aa_set_tuning": (_ci, [_ci, _ci, _ci, _vp]),
    "aa_get_tuning": (_ci, [_ci, _ci, _ci, _vp]),
    "aa_load_tuning": (_ci, [_cp]),
    "aa_mie_create_rc": (_ci, [ctypes.POINTER(_vp), _cp, _ci, _cp, _ip, _vp, _vp, _ip]),
    "aa_mie_destroy": (_ci, [_vp]),
    "aa_get_aop_vector": (_ci, [_vp, _ci, _ci, _ci, _ci, _fp, _cp, _ci, _fp, _fp, _ci, _ci, _fp, _fp, _fp, _fp, _ip]),
    "aa_get_aop_scalar": (_ci, [_vp, _ci, _ci, _ci, _ci, _fp, _cp, _ci, _fp, _fp, _fp, _fp, _fp, _ip]),
    "aa_get_aop_vector_dev": (_ci, [_vp, _ci, _ci, _ci, _ci, _vp, _vp, _vp, _vp, _ci, _ci, _vp, _vp, _vp, _vp, _vp]),
    "aa_vlidort_scalar_lambert": (_ci, [_ci, _ci, _ci, _dp, _ci, _ci, _ci, _ci, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                        _dp, _dp, _dp, _dp, _dp, _dp, _cd, _ci, _ci, _dp, _dp, _ip, _ip]),
    "aa_vlidort_vector_lambert": (_ci, [_ci, _ci, _ci, _dp, _ci, _ci, _ci, _ci, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                        _dp, _dp, _dp, _dp, _dp, _dp, _dp, _cd, _ci, _ci, _dp, _dp, _dp, _dp, _ip,
                                        _ip]),
    "aa_vlidort_vector_lambert_opt": (_ci, [_ci, _ci, _ci, _dp, _ci, _ci, _ci, _ci, _dp, _dp, _dp, _dp, _dp, _dp,
                                            _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _cd, _ci, _ci, _ci, _ci, _dp,
                                            _dp, _dp, _dp, _ip, _ip]),
    "aa_vlidort_vector_lambert_multigeom": (_ci, [_ci, _ci, _ci, _ci, _dp, _ci, _ci, _ci, _ci, _dp, _dp, _dp, _dp,
                                                  _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _cd, _ci, _ci,
                                                  _ci, _ci, _dp, _dp, _dp, _dp, _ip, _ip]),
    "aa_column_radiance_dev": (_ci, [_vp, _ci, _ci, _ci, _ci, _vp, _vp, _vp, _vp, _ci, _ci, _ci, _ci, _ci, _vp, _vp,
                                     _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _cd, _vp, _vp, _vp, _vp, _vp,
                                     _vp]),
    "aa_column_radiance_dev_noalloc": (_ci, [_vp, _ci, _ci, _ci, _ci, _vp, _vp, _vp, _vp, _ci, _ci, _ci, _ci, _ci,
                                             _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _cd, _vp, _vp,
                                             _vp, _vp, _vp, _vp]),
    "aa_vlidort_scalar_landmodis": (_ci, [_ci, _ci, _ci, _dp, _ci, _ci, _ci, _ci, _dp, _dp, _dp, _dp, _dp, _dp,
                                          _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _cd, _ci, _ci, _dp, _dp, _dp, _ip,
                                          _ip]),
    "aa_vbrdf_modis": (_ci, [_ci, _ci, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _ip]),
    "aa_vlidort_scalar_cx": (_ci, [_ci, _ci, _ci, _dp, _ci, _ci, _ci, _ci, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                   _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _cd, _ci, _ci, _dp, _dp, _dp, _ip, _ip]),
    "aa_vlidort_vector_cx": (_ci, [_ci, _ci, _ci, _dp, _ci, _ci, _ci, _ci, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                   _dp, _dp, _dp, _dp, _dp, _dp, _dp, _cd, _ci, _ci, _dp, _dp, _dp, _dp, _dp,
                                   _dp, _dp, _ip, _ip]),
    "aa_vbrdf_coxmunk": (_ci, [_ci, _ci, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _ip]),
    "aa_vbrdf_bpdf": (_ci, [_ci, _ci, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _ip]),
    "aa_vlidort_vector_lambert_bpdf": (_ci, [_ci, _ci, _ci, _dp, _ci, _ci, _ci, _ci, _dp, _dp, _dp, _dp, _dp, _dp,
                                             _dp, _dp, _dp, _dp, _dp, _dp, _dp, _cd, _ci, _ci, _dp, _dp, _dp, _dp,
                                             _dp, _dp, _dp, _ip, _ip]),
    "aa_vlidort_vector_landmodis_bpdf": (_ci, [_ci, _ci, _ci, _dp, _ci, _ci, _ci, _ci, _dp, _dp, _dp, _dp, _dp, _dp,
                                               _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _cd, _ci, _ci, _dp, _dp, _dp,
                                               _dp, _dp, _dp, _dp, _ip, _ip]),
    "aa_column_radiance_modis_dev": (_ci, [_vp, _ci, _ci, _ci, _ci, _vp, _vp, _vp, _vp, _ci, _ci, _ci, _vp, _vp,
                                           _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _cd, _vp, _vp, _vp,
                                           _vp, _vp]),
    "aa_rot_calc": (_ci, [_ci, _ci, _ci, _dp, _dp, _dp, _dp, _cd, _ci, _dp, _dp, _ip]),
    "aa_rot_calc_dev": (_ci, [_ci, _ci, _ci, _dp, _vp, _vp, _vp, _vp, _dp, _vp]),
    "aa_edge_vars": (_ci, [_ci, _ci, _fp, _fp, _cf, _fp, _fp, _fp]),
    "aa_edge_vars_dev": (_ci, [_ci, _ci, _vp, _vp, _cf, _vp, _vp, _vp, _vp]),
    "aa_calc_qm": (_ci, [_ci, _ci, _ci, _fp, _fp, _fp]),
    "aa_calc_qm_dev": (_ci, [_ci, _ci, _ci, _vp, _vp, _vp, _vp]),
    "aa_set_arena_bytes": (_ci, [ctypes.c_size_t, ctypes.c_size_t]),
    "aa_release_arenas": (_ci, []),
    "aa_version": (ctypes.c_char_p, []),
    "aa_profile_start": (_ci, []),
    "aa_profile_read": (_ci, [_dp, ctypes.POINTER(ctypes.c_longlong), _ci]),
}

ERRORS = {-1: "AA_EARG (bad argument)", -2: "AA_EHIP (HIP runtime error)", -3: "AA_ENOTABLE (no table)",
          -4: "AA_EDIMS (beyond compiled maxima)", -5: "AA_EUNSUP (option not implemented)",
          -7: "AA_ENOFILE (a file cannot be read)", -6: "AA_ENOTRES (tables not resident)"}

_LIB = None


class AAError(RuntimeError):
    pass


def lib() -> ctypes.CDLL:
    """The loaded library; raises when it is not built or does not load."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise AAError(f"{LIB_PATH} is not built (run python -m aeroapps_amd.build_lib)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(code: int, what: str) -> None:
    if code != 0:
        raise AAError(f"{what} failed: {ERRORS.get(code, code)}")
