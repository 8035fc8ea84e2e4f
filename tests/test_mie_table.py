"""Optics-table ingestion semantics (Chem_MieTableRead after its netCDF reads,
Chem_MieCreateFromRC's resource file and bin assignment): the oracle
restatement against closed forms and the Fortran kind rules it follows."""
import numpy as np
import pytest

from aeroapps_amd import chem_mie as AM
from oracle import chem_mie as CM

f32 = np.float32


def test_polint_follows_fortran_kinds():
    # table wavelengths that float32 cannot hold, target between two nodes
    x = np.array([4.7e-7 + 1e-17, 5.5e-7 + 3e-17, 6.6e-7 + 7e-17])
    y = np.array([0.123456789012345, 0.2345678901234567, 0.3456789012345678])
    xw = f32(5.1e-7)
    i = 1   # 1-based node below the target
    slope = f32((y[i] - y[i - 1]) / (x[i] - x[i - 1]))
    dx = f32(float(xw) - x[i - 1])
    want = f32(y[i - 1] + float(f32(slope * dx)))
    assert CM.polint(x, y, xw) == want
    # a restatement that rounds x(i), y(i) to float32 first (the plausible slip)
    # gives different bits on some of these inputs
    rng = np.random.default_rng(5)
    diffs = 0
    for _ in range(200):
        yy = rng.uniform(0.1, 2.0, 3)
        xt = f32(rng.uniform(4.8e-7, 6.5e-7))
        j = 1 if float(xt) < x[1] else 2
        s = f32((yy[j] - yy[j - 1]) / (x[j] - x[j - 1]))
        slip = f32(f32(yy[j - 1]) + f32(s * f32(xt - f32(x[j - 1]))))
        diffs += int(CM.polint(x, yy, xt) != slip)
    assert diffs > 0


def test_polint_linear_data_and_ranges():
    x = np.array([3.4e-7, 4.4e-7, 5.5e-7, 8.7e-7])
    a, b = 0.3, 2.0e5
    y = a + b * x
    for xw in (3.9e-7, 5.0e-7, 6.6e-7):
        got = CM.polint(x, y, f32(xw))
        exact = a + b * float(f32(xw))
        assert abs(float(got) - exact) <= 2 * np.spacing(f32(exact))
    # below the table: first interval's line from x(1); above: y(n)
    lo = CM.polint(x, y, f32(3.0e-7))
    assert abs(float(lo) - (a + b * float(f32(3.0e-7)))) <= 2 * np.spacing(f32(lo))
    assert CM.polint(x, y, f32(1.0e-6)) == f32(y[-1])
    # a node reproduces its value rounded to float32
    xs = np.array([f32(4.4e-7), f32(5.5e-7)], dtype=np.float64)
    assert CM.polint(xs, np.array([0.1234567890123, 0.2]), f32(4.4e-7)) == f32(0.1234567890123)


def _file_arrays(rng, nl=3, nrh=5, nbin=2, nmom=6, npol=6):
    lam = np.array([4.7e-7, 5.5e-7, 6.6e-7])[:nl] + 1e-17
    rh = np.array([0.0, 0.25, 0.5, 0.75, 0.99, 0.995, 0.999])[:nrh]   # nodes on the 0.001 RH map
    reff = rng.uniform(1e-7, 1e-6, (nbin, nrh))
    arr = lambda: rng.uniform(0.1, 3.0, (nbin, nrh, nl))
    return dict(lambda_table=lam, rh=rh, reff=reff, bext=arr(), bsca=arr(), bbck=arr(), g=arr(),
                pback=rng.uniform(0.1, 1.0, (npol, nbin, nrh, nl)), pmom=rng.uniform(-1, 1, (npol, nmom, nbin, nrh, nl)),
                gf=rng.uniform(1.0, 2.0, (nbin, nrh)), rhop=rng.uniform(1000, 2500, (nbin, nrh)),
                refr=arr(), refi=-arr() * 1e-3)


def test_table_read_fills_and_derived_fields():
    rng = np.random.default_rng(8)
    d = _file_arrays(rng)
    ch = np.array([5.0e-7, 6.0e-7], dtype=f32)
    rc, t = CM.table_read(d["lambda_table"], ch, d["rh"], d["reff"], d["bext"], d["bsca"], d["bbck"], d["g"], nmom=4,
                          pback=d["pback"], pmom=d["pmom"], gf=d["gf"], rhop=d["rhop"], refr=d["refr"],
                          refi=d["refi"])
    assert rc == 0
    assert (t["refi"] > 0).all()                                   # refimag made non-negative
    assert np.array_equal(t["rhod"], np.repeat(t["rhop"][:, :1], 5, axis=1))   # rhod: first RH row of rhop
    vol = d["gf"] ** 3 / np.repeat(d["rhop"][:, :1], 5, axis=1)
    assert np.allclose(t["vol"], vol, rtol=1e-6)
    assert np.allclose(t["area"], vol / (4.0 / 3.0 * d["reff"]), rtol=1e-6)
    assert t["pmom"].shape == (6, 4, 2, 2, 5)
    # absent variables: pback 1, gf/rhop/rhod/refr/refi -999
    rc, t2 = CM.table_read(d["lambda_table"], ch, d["rh"], d["reff"], d["bext"], d["bsca"], d["bbck"], d["g"])
    assert rc == 0 and (t2["pback"] == 1.0).all() and (t2["gf"] == -999.0).all() and (t2["refi"] == -999.0).all()
    assert (t2["rhod"] == -999.0).all() and (t2["refr"] == -999.0).all()
    # more moments asked than the file holds
    rc, _ = CM.table_read(d["lambda_table"], ch, d["rh"], d["reff"], d["bext"], d["bsca"], d["bbck"], d["g"], nmom=7,
                          pmom=d["pmom"])
    assert rc == 99


def test_query_extra_at_table_nodes():
    rng = np.random.default_rng(9)
    d = _file_arrays(rng)
    ch = np.array([5.5e-7], dtype=f32)
    rc, t = CM.table_read(d["lambda_table"], ch, d["rh"], d["reff"], d["bext"], d["bsca"], d["bbck"], d["g"],
                          pback=d["pback"], gf=d["gf"], rhop=d["rhop"])
    # just above an RH node (the snap truncates (rh + 0.001)*1000) the
    # interpolation weight is 0: values of that node
    rh_nodes = (t["rh"][:4] + f32(0.0005)).astype(f32)
    out = CM.query_extra(t, 1, 0, rh_nodes)
    irh, _, arh = CM.snap(rh_nodes, t["rhi"], t["rha"], len(t["rh"]))
    node = irh - 1
    assert (arh == 0).all()
    assert np.array_equal(out["reff"], (f32(1e6) * t["reff"][1][node]).astype(f32))
    assert np.array_equal(out["gf"], t["gf"][1][node])
    assert (out["rhod"] == t["rhod"][1, 0]).all()


RC_TEXT = """#
# Chem_Mie resource file in the reference's format
n_channels: 2
n_moments: 300
r_channels: 0.470e-6 0.550e-6
filename_optical_properties_DU: ExtData/optics_DU.v15_3.nc
filename_optical_properties_SS: ExtData/optics_SS.v3_3.nc
filename_optical_properties_OC: ExtData/optics_OC.v1_3.nc
filename_optical_properties_BC: ExtData/optics_BC.v1_3.nc
filename_optical_properties_SU: ExtData/optics_SU.v1_3.nc
filename_optical_properties_NI: ExtData/optics_NI.v2_5.nc
filename_optical_properties_BRC: ExtData/optics_BRC.v1_5.nc
"""


def test_resource_file_settings():
    s = AM.mie_rc_settings(RC_TEXT)
    assert s["n_channels"] == 2 and s["nmom"] == 300
    assert np.allclose(s["channels_m"], [0.47e-6, 0.55e-6])
    assert s["optics_files"]["NI"] == "ExtData/optics_NI.v2_5.nc"
    assert AM.mie_rc_settings("n_channels: 1\nr_channels: 5.5e-7\n")["nmom"] == 0


def test_sulfate_and_nitrate_bin_filters():
    su = ["DMS", "SO2", "SO4", "MSA", "so4v", "sulfate"]
    assert AM.species_bin_names("SU", su) == ["SO4", "so4v", "sulfate"]
    assert AM.species_bin_names("NI", ["NH3", "NO3an1", "NH4a", "no3an2"]) == ["NO3an1", "no3an2"]
    assert AM.species_bin_names("DU", ["du001", "du002"]) == ["du001", "du002"]
