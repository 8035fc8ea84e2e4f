"""The Fortran drop-in shim (aeroapps_amd/fortran: modules VLIDORT_LAMBERT,
VLIDORT_BRDF_MODIS, VLIDORT_BRDF_CX, Chem_RegistryMod and Chem_MieMod,
external VLIDORT_getAOPvector, all over ISO_C_BINDING) called from a Fortran
program with the reference's procedure names and dummy arguments."""
import os
import subprocess
import sys

import numpy as np
import pytest

from aeroapps_amd import chem_mie as AM
from aeroapps_amd import synthetic as S
from aeroapps_amd.chem_mie import _vname
from oracle import chem_mie as CM
from _fixtures import GOLDEN, RTOL_FP64, load, max_rel
from test_mie_rc import FILE_BINS, MIE_RC, ORDER, _files

sys.path.insert(0, GOLDEN)
from golden_inputs import input_checksum, load_sampled, sampled_inputs  # noqa: E402

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "fortran", "_build", "shim_driver")


def _run(tmp_path, payload):
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} not built (make -C aeroapps_amd/fortran)")
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    fin.write_bytes(payload)
    p = subprocess.run([DRIVER, str(fin), str(fout)], capture_output=True, text=True, timeout=120, cwd=tmp_path)
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


def test_vector_cx_through_fortran_shim(gpu, tmp_path):
    print("case fortran vcx_n2_edge", flush=True)
    z = np.load(os.path.join(ROOT, "tests", "golden", "brdf", "vcx_n2_edge.npz"), allow_pickle=False)
    d = {k: z[k] for k in z.files}
    km, nch, nobs = int(d["km"]), len(d["channels"]), len(d["sza"])
    f8 = lambda a: np.ascontiguousarray(a, dtype=np.float64).tobytes()
    hdr = np.array([9, km, nch, nobs, int(d["meta_nstreams"]), int(d["nmom"]), int(d["npol"]), 0],
                   dtype=np.int32).tobytes() + np.float64(S.MISSING).tobytes()
    body = b"".join(f8(d[k]) for k in ("channels", "rot", "depol", "tau", "ssa", "pmom", "pe", "he", "te",
                                       "u10m", "v10m", "mr", "sza", "raa", "vza"))
    out = _run(tmp_path, hdr + body)
    rc = int(np.frombuffer(out[:4], dtype=np.int32)[0])
    arr = np.frombuffer(out[4:], dtype=np.float64).reshape(7, nch, nobs)
    assert rc == int(d["out_rc"]) == 0
    assert max_rel(arr[0], d["out_radiance"]) < RTOL_FP64
    assert max_rel(arr[1], d["out_reflectance"]) < RTOL_FP64
    I = np.abs(d["out_radiance"])
    ok = np.abs(I / S.MISSING - 1) > 1e-3
    for k, key in ((2, "out_q"), (3, "out_u")):
        assert (np.abs(arr[k] - d[key])[ok] / I[ok]).max() < RTOL_FP64, key
    assert max_rel(arr[4], d["out_br"]) < 1e-13
    assert np.array_equal(arr[5], d["out_br_q"]) and np.array_equal(arr[6], d["out_br_u"])


def test_landmodis_bpdf_through_fortran_shim(gpu, tmp_path):
    """VLIDORT_Vector_LandMODIS_BPDF through the Fortran module against the
    reference's own driver (edge columns: missing geometry, negative direct
    bounce, white-sky failure, SZA = 0 latch, RAA 0)."""
    print("case fortran vbpdf_modis_n3_edge", flush=True)
    z = np.load(os.path.join(ROOT, "tests", "golden", "brdf", "vbpdf_modis_n3_edge.npz"), allow_pickle=False)
    d = {k: z[k] for k in z.files}
    km, nch, nobs = int(d["km"]), len(d["channels"]), len(d["sza"])
    f8 = lambda a: np.ascontiguousarray(a, dtype=np.float64).tobytes()
    hdr = np.array([10, km, nch, nobs, int(d["meta_nstreams"]), int(d["nmom"]), int(d["npol"]), 0],
                   dtype=np.int32).tobytes() + np.float64(S.MISSING).tobytes()
    body = b"".join(f8(d[k]) for k in ("channels", "rot", "depol", "tau", "ssa", "pmom", "pe", "he", "te",
                                       "kernel_wt", "param", "bpdf", "sza", "raa", "vza"))
    out = _run(tmp_path, hdr + body)
    rc = int(np.frombuffer(out[:4], dtype=np.int32)[0])
    arr = np.frombuffer(out[4:], dtype=np.float64).reshape(7, nch, nobs)
    assert rc == int(d["out_rc"]) == 0
    assert max_rel(arr[0], d["out_radiance"]) < RTOL_FP64
    assert max_rel(arr[1], d["out_reflectance"]) < RTOL_FP64
    I = np.abs(d["out_radiance"])
    ok = (np.abs(I / S.MISSING - 1) > 1e-3) & (d["out_radiance"] != -500.0) & (d["raa"][None, :] != 0.0)
    for k, key in ((2, "out_q"), (3, "out_u")):
        assert (np.abs(arr[k] - d[key])[ok] / I[ok]).max() < RTOL_FP64, key
    for k, key in ((4, "out_br"), (5, "out_br_q"), (6, "out_br_u")):
        assert np.max(np.abs(arr[k] - d[key]) / np.maximum(np.abs(d[key]), 1e-300)) < 1e-12, key


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
    assert rc == rc0 == 0 and idx == 3   # SS::ss003: bin 3 of the sea-salt table
    assert np.array_equal(tau, t0.ravel()) and np.array_equal(ssa, s0.ravel()) and np.array_equal(g, g0.ravel())
    assert np.array_equal(pm, p0.ravel())


@pytest.mark.parametrize("name,mode", [("vector_multigeom", 6), ("vector_2os", 7)])
def test_vector_multigeom_and_2os_through_fortran_shim(gpu, tmp_path, name, mode):
    """The MultiGeom form of VLIDORT_Vector_Lambert and DO_2OS_CORRECTION =
    .true. against the reference drivers' outputs (sampled fixtures)."""
    print(f"case fortran {name}", flush=True)
    z = load_sampled(name)
    spec = z["spec"]
    d = sampled_inputs(spec)
    assert input_checksum(d) == str(z["checksum"])
    km, nch, nobs = d["km"], len(d["channels"]), spec["ncol"]
    ng = spec.get("ngeom", 1)
    f8 = lambda a: np.ascontiguousarray(a, dtype=np.float64).tobytes()
    hdr = np.array([mode, km, nch, nobs, spec["nstreams"], d["nmom"], d["npol"], 0, ng],
                   dtype=np.int32).tobytes() + np.float64(S.MISSING).tobytes()
    body = b"".join(f8(d[k]) for k in ("channels", "rot", "depol", "tau", "ssa", "pmom", "pe", "he", "te",
                                       "albedo", "sza", "raa", "vza", "flux_factor"))
    out = _run(tmp_path, hdr + body)
    rc = int(np.frombuffer(out[:4], dtype=np.int32)[0])
    arr = np.frombuffer(out[4:], dtype=np.float64).reshape((4, ng, nch, nobs) if ng > 1 else (4, nch, nobs))
    assert rc == int(z["out_rc"]) == 0
    assert max_rel(arr[0], z["out_radiance"]) < RTOL_FP64
    assert max_rel(arr[1], z["out_reflectance"]) < RTOL_FP64
    I = np.abs(z["out_radiance"])
    for k, key in ((2, "out_q"), (3, "out_u")):
        assert (np.abs(arr[k] - z[key]) / I).max() < RTOL_FP64, key


RC_REGISTRY = "\n".join([
    "doing_DU: yes", "nbins_DU: 3", "units_DU: kg/kg",
    "doing_SS: yes", "nbins_SS: 2", "units_SS: kg/kg",
    "doing_SU: yes", "nbins_SU: 4", "units_SU: kg/kg",
    "variable_table_SU::", "DMS kg/kg a", "SO2 kg/kg b", "SO4 kg/kg c", "MSA kg/kg d", "::",
    "doing_BC: yes", "nbins_BC: 2", "units_BC: kg/kg",
    "variable_table_BC::", "BCphobic kg/kg a", "BCphilic kg/kg b", "::",
    "doing_OC: yes", "nbins_OC: 2", "units_OC: kg/kg", ""])


def test_chem_mie_from_resource_files_through_fortran_shim(gpu, tmp_path):
    """Chem_MieCreate(rcfile, rc) with the registry read from the working
    directory and the tables read by a Fortran reader, then Chem_MieQueryIdx,
    Chem_MieQuery (every optional output, by index and by name),
    Chem_MieQueryTauList and Chem_MieQueryAllBand3D: equal to the oracle's
    Chem_MieTableRead + Chem_MieQuery restatement and to the same table set
    built through the C-ABI from Python."""
    print("case fortran chem_mie rc", flush=True)
    files = _files(seed=12)
    (tmp_path / "Chem_MieRegistry.rc").write_text(RC_REGISTRY)
    (tmp_path / "Chem_Mie.rc").write_text(MIE_RC.replace("ExtData/", ""))
    for path, a in files.items():
        nbin, nrh, nch = a["bext"].shape
        npol, nmom = a["pmom"].shape[:2]
        f8 = lambda x: np.ascontiguousarray(x, dtype=np.float64).tobytes()
        (tmp_path / os.path.basename(path)).write_bytes(
            np.array([nch, nrh, nbin, nmom, npol], np.int32).tobytes() +
            b"".join(f8(a[k]) for k in ("lambda_table", "rh", "reff", "bext", "bsca", "bbck", "g", "pback", "pmom",
                                        "gf", "rhop", "refr", "refi")))
    rng = np.random.default_rng(13)
    n = 24
    q = rng.uniform(1e-7, 1e-4, n).astype(np.float32)
    rh = rng.uniform(-0.05, 1.0, n).astype(np.float32)
    out = _run(tmp_path, np.array([8, n], np.int32).tobytes() + q.tobytes() + rh.tobytes())
    hdr = np.frombuffer(out[:4 * 16], dtype=np.int32)
    rcs = hdr[:8]
    ncalls, idx, idx2, idx_dms, nch, nmom, npol, nq = (int(v) for v in hdr[8:])
    assert list(rcs) == [0, 0, 0, 0, 0, 0, 0, 1] and ncalls == 7
    assert (idx, idx2, idx_dms) == (2, 2, -1)
    assert (nch, nmom, npol, nq) == (2, 4, 6, 13)
    f = np.frombuffer(out[4 * 16:], dtype=np.float32)
    o = 0

    def take(m):
        nonlocal o
        o += m
        return f[o - m:o]
    ch = take(2)
    outs = take(n * 16).reshape(16, n)
    pmq = take(nmom * npol * n).reshape(n, npol, nmom)
    tl = take(n)
    h = n // 2
    t3, s3, a3 = (take(2 * h * 2).reshape(2, h, 2) for _ in range(3))
    bc1 = take(2)
    assert o == len(f)
    ch_m = np.array([0.470e-6, 0.550e-6], np.float32)
    assert np.array_equal(ch, ch_m)
    # the oracle's tables
    tabs = {}
    for sp in ORDER:
        a = files[f"ExtData/optics_{sp}.nc"]
        rc0, t = CM.table_read(a["lambda_table"], ch_m, a["rh"], a["reff"], a["bext"], a["bsca"], a["bbck"], a["g"],
                               nmom=4, pback=a["pback"], pmom=a["pmom"], gf=a["gf"], rhop=a["rhop"], refr=a["refr"],
                               refi=a["refi"])
        assert rc0 == 0
        tabs[sp] = t
    ss = tabs["SS"]
    # ss002 = bin 2 of SS, channel 2, every output
    tau, ssa, g, pm = CM.query(ss["bext"], ss["bsca"], ss["g"], ss["pmom"], ss["rhi"], ss["rha"], 1, 1, q, rh)
    assert np.array_equal(outs[0], tau) and np.array_equal(outs[1], ssa) and np.array_equal(outs[2], g)
    assert np.array_equal(np.moveaxis(pmq, 1, 2), pm)
    ex = CM.query_extra(ss, 1, 1, rh)
    for k, key in zip(range(6, 13), ("reff", "p11", "p22", "gf", "rhod", "vol", "area")):
        assert np.array_equal(outs[k], ex[key]), key
    # bext, bsca, bbck, refr, refi: the same table set built from Python through the C-ABI
    reader = lambda path: {k: (v if k in ("lambda_table", "rh") else np.asarray(v).T)
                           for k, v in files["ExtData/" + os.path.basename(path)].items()}
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        mie, rc = AM.MieTables.from_rc("Chem_Mie.rc", reader)
    finally:
        os.chdir(cwd)
    assert rc == 0
    k_ss = mie.query_idx("ss002")
    assert k_ss == FILE_BINS["DU"] + 1
    lib_q = mie.query(k_ss, 2, q, rh, nmom=4, npol=6)
    for k, key in zip((3, 4, 5, 13, 14), ("bext", "bsca", "bbck", "refr", "refi")):
        assert np.array_equal(outs[k], lib_q[key]), key
    # by name: du001 is registry position 1, queried as bin 1 of the selected (SS) table
    t_by_name = CM.query(ss["bext"], ss["bsca"], ss["g"], ss["pmom"], ss["rhi"], ss["rha"], 0, 0, q[:1], rh[:1])[0]
    assert outs[15][0] == t_by_name[0]
    # TauList (channel 1) and AllBand3D (channels 1, 2)
    assert np.array_equal(tl, CM.query(ss["bext"], ss["bsca"], ss["g"], ss["pmom"], ss["rhi"], ss["rha"], 1, 0, q,
                                       rh)[0])
    for c in range(2):
        tq, sq, gq, _ = CM.query(ss["bext"], ss["bsca"], ss["g"], ss["pmom"], ss["rhi"], ss["rha"], 1, c, q[:2 * h],
                                 rh[:2 * h])
        assert np.array_equal(t3[c].ravel(), tq) and np.array_equal(s3[c].ravel(), sq)
        assert np.array_equal(a3[c].ravel(), gq)
    # BCphilic: bin 2 of the black-carbon table, channel 1
    bc = tabs["BC"]
    tb, sb, _, _ = CM.query(bc["bext"], bc["bsca"], bc["g"], bc["pmom"], bc["rhi"], bc["rha"], 1, 0, q[:1], rh[:1])
    assert bc1[0] == tb[0] and bc1[1] == sb[0]
    mie.close()
