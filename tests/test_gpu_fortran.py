"""The Fortran drop-in shim (aeroapps_amd/fortran: modules VLIDORT_LAMBERT,
VLIDORT_BRDF_MODIS and Chem_MieMod, external VLIDORT_getAOPvector, all over
ISO_C_BINDING) called from
a Fortran program with the reference's procedure names and dummy arguments."""
import os
import subprocess

import numpy as np
import pytest

from aeroapps_amd import synthetic as S
from aeroapps_amd.chem_mie import _vname
from oracle import chem_mie as CM
from _fixtures import RTOL_FP64, load, max_rel

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "fortran", "_build", "shim_driver")


def _run(tmp_path, payload):
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} not built (make -C aeroapps_amd/fortran)")
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    fin.write_bytes(payload)
    p = subprocess.run([DRIVER, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    return fout.read_bytes()


@pytest.mark.parametrize("name", ["scalar_n4_edge", "vector_n2"])
def test_lambert_drivers_through_fortran_shim(gpu, tmp_path, name):
    print(f"case fortran {name}", flush=True)
    d = load(name)
    vec = bool(d["meta_vector"])
    km, nch, nobs = d["km"], len(d["channels"]), len(d["sza"])
    f8 = lambda a: np.ascontiguousarray(a, dtype=np.float64).tobytes()
    hdr = np.array([2 if vec else 1, km, nch, nobs, d["meta_nstreams"], d["nmom"], d["npol"],
                    d["meta_plane_parallel"]], dtype=np.int32).tobytes() + np.float64(S.MISSING).tobytes()
    body = b"".join(f8(d[k]) for k in ("channels", "rot", "depol", "tau", "ssa", "g", "pmom", "pe", "he", "te",
                                       "albedo", "sza", "raa", "vza", "flux_factor"))
    out = _run(tmp_path, hdr + body)
    rc = int(np.frombuffer(out[:4], dtype=np.int32)[0])
    arr = np.frombuffer(out[4:], dtype=np.float64).reshape(4, nch, nobs)   # Fortran (nobs,nch) each
    assert rc == d["out_rc"]
    assert max_rel(arr[0], d["out_radiance"]) < RTOL_FP64
    assert max_rel(arr[1], d["out_reflectance"]) < RTOL_FP64
    if vec:
        ok = np.abs(d["out_radiance"] / S.MISSING - 1) > 1e-3
        for k, key in ((2, "out_q"), (3, "out_u")):
            err = np.abs(arr[k] - d[key])[ok] / np.abs(d["out_radiance"])[ok]
            assert err.max() < RTOL_FP64


def test_landmodis_through_fortran_shim(gpu, tmp_path):
    print("case fortran modis_n4_edge", flush=True)
    z = np.load(os.path.join(ROOT, "tests", "golden", "brdf", "modis_n4_edge.npz"), allow_pickle=False)
    d = {k: z[k] for k in z.files}
    km, nch, nobs = int(d["km"]), len(d["channels"]), len(d["sza"])
    f8 = lambda a: np.ascontiguousarray(a, dtype=np.float64).tobytes()
    hdr = np.array([4, km, nch, nobs, int(d["meta_nstreams"]), int(d["nmom"]), int(d["npol"]), 0],
                   dtype=np.int32).tobytes() + np.float64(S.MISSING).tobytes()
    body = b"".join(f8(d[k]) for k in ("channels", "rot", "depol", "tau", "ssa", "g", "pmom", "pe", "he", "te",
                                       "kernel_wt", "param", "sza", "raa", "vza"))
    out = _run(tmp_path, hdr + body)
    rc = int(np.frombuffer(out[:4], dtype=np.int32)[0])
    arr = np.frombuffer(out[4:], dtype=np.float64).reshape(3, nch, nobs)
    assert rc == int(d["out_rc"]) == 0
    neg = d["out_radiance"] == -500.0
    assert np.array_equal(arr[0] == -500.0, neg)
    assert max_rel(arr[0][~neg], d["out_radiance"][~neg]) < RTOL_FP64
    assert max_rel(arr[1][~neg], d["out_reflectance"][~neg]) < RTOL_FP64
    assert max_rel(arr[2][~neg], d["out_br"][~neg]) < 1e-14


def test_scalar_cx_through_fortran_shim(gpu, tmp_path):
    print("case fortran cx_n4_edge", flush=True)
    z = np.load(os.path.join(ROOT, "tests", "golden", "brdf", "cx_n4_edge.npz"), allow_pickle=False)
    d = {k: z[k] for k in z.files}
    km, nch, nobs = int(d["km"]), len(d["channels"]), len(d["sza"])
    f8 = lambda a: np.ascontiguousarray(a, dtype=np.float64).tobytes()
    hdr = np.array([5, km, nch, nobs, int(d["meta_nstreams"]), int(d["nmom"]), int(d["npol"]), 0],
                   dtype=np.int32).tobytes() + np.float64(S.MISSING).tobytes()
    body = b"".join(f8(d[k]) for k in ("channels", "rot", "depol", "tau", "ssa", "g", "pmom", "pe", "he", "te",
                                       "u10m", "v10m", "mr", "sza", "raa", "vza"))
    out = _run(tmp_path, hdr + body)
    rc = int(np.frombuffer(out[:4], dtype=np.int32)[0])
    arr = np.frombuffer(out[4:], dtype=np.float64).reshape(3, nch, nobs)
    assert rc == int(d["out_rc"]) == 0
    assert max_rel(arr[0], d["out_radiance"]) < RTOL_FP64
    assert max_rel(arr[1], d["out_reflectance"]) < RTOL_FP64
    assert max_rel(arr[2], d["out_br"]) < 1e-13


def test_get_aop_vector_through_fortran_shim(gpu, tmp_path):
    channels = (550.0, 870.0)
    nmom, npol = 16, 6
    tabs = S.make_mie_tables(channels, nmom=nmom, npol=npol)
    cols = S.make_columns(5, 72, channels, seed=4)
    nq = len(S.TRACERS)
    parts = [np.array([3, len(channels), len(tabs), 72, 5, nq, nmom, npol], dtype=np.int32).tobytes(),
             tabs[0].lambda_m.astype(np.float32).tobytes(), np.asarray(channels, np.float32).tobytes()]
    for t in tabs:
        parts.append(np.array([len(t.rh), t.nbin], dtype=np.int32).tobytes())
        parts.append(t.rh.astype(np.float32).tobytes())
        parts.append(np.stack([t.bext, t.bsca, t.bbck, t.g, t.refr, t.refi]).astype(np.float32).tobytes())
        parts.append(np.ascontiguousarray(t.pmom[:, :nmom]).astype(np.float32).tobytes())
        parts.append(b"".join(n.encode().ljust(16) for n in t.tracers))
    parts += [_vname(S.TRACERS), cols.qm.tobytes(), cols.rh.tobytes()]
    out = _run(tmp_path, b"".join(parts))
    rc, idx = np.frombuffer(out[:8], dtype=np.int32)
    f = np.frombuffer(out[8:], dtype=np.float32)
    n3 = 72 * len(channels) * 5
    tau, ssa, g, pm = f[:n3], f[n3:2 * n3], f[2 * n3:3 * n3], f[3 * n3:]
    t0, s0, g0, p0, rc0 = CM.get_aop_vector(tabs, channels, tabs[0].lambda_m, S.TRACERS, cols.qm, cols.rh, nmom, npol)
    assert rc == rc0 == 0 and idx == 8   # SS::ss003 is the 8th tracer (1-based)
    assert np.array_equal(tau, t0.ravel()) and np.array_equal(ssa, s0.ravel()) and np.array_equal(g, g0.ravel())
    assert np.array_equal(pm, p0.ravel())
