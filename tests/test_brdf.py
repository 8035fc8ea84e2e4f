"""The BRDF-supplement restatement (oracle/brdf.py) against the reference's own
BRDF self-tester results, the compiled reference's supplement (fixture
tests/golden/brdf/supplement.npz) and closed forms of the MODIS kernels."""
import json
import math
import os

import numpy as np
import pytest

from oracle import brdf as B

HERE = os.path.dirname(os.path.abspath(__file__))
BDIR = os.path.join(HERE, "golden", "brdf")


def _close8(a, b):
    # the self-tester prints 8 significant digits
    return abs(a - b) <= 6e-8 * abs(b) + 1e-15


@pytest.mark.parametrize("name", ["Ross-thick", "Li-sparse"])
def test_reference_selftest_vectors(name):
    d = json.load(open(os.path.join(BDIR, "selftest.json")))
    inp, k = d["inputs"], d["kernels"][name]
    pars = tuple(inp["lisparse_pars"]) if name == "Li-sparse" else (0.0, 0.0, 0.0)
    kern = [(k["index"], inp["factor"], pars)]
    for raa, want in zip(inp["raa"], k["exact"]):
        r = B.vbrdf_master(inp["nstreams"], inp["sza"], inp["vza"], raa, kern, inp["nstreams_brdf"], wsa=False)
        assert _close8(r["db"], want), (raa, r["db"], want)
    r = B.vbrdf_master(inp["nstreams"], inp["sza"], inp["vza"], inp["raa"][0], kern, inp["nstreams_brdf"],
                       wsa=False)
    for m, v in k["fourier"].items():
        m = int(m)
        assert _close8(r["F"][m, 0, 0], v["Stream-Stream"])
        assert _close8(r["F0"][m, 0], v["Solars-Stream"])
        assert _close8(r["UF"][m, 0], v["Stream-Users"])
        assert _close8(r["UF0"][m], v["Solars-Users"])


def test_vs_compiled_reference_supplement():
    z = np.load(os.path.join(BDIR, "supplement.npz"), allow_pickle=False)
    for c, case in enumerate(z["cases"]):
        n = int(case[0])
        sza, vza, raa, fiso, fgeo, fvol, p1, p2 = (float(v) for v in case[1:])
        r = B.vbrdf_master(n, sza, vza, raa, B.landmodis_setup(fiso, fgeo, fvol, (p1, p2)))
        assert r["ok"] == (int(z["status"][c]) == 0), c
        assert abs(r["db"] - z["db"][c]) <= 1e-15 * abs(z["db"][c]) + 1e-18, c
        if r["ok"]:
            assert abs(r["total_wsa"] - z["wsa"][c]) <= 1e-13 * abs(z["wsa"][c]), c
        for key, got in (("F", r["F"]), ("F0", r["F0"]), ("UF", r["UF"]), ("UF0", r["UF0"])):
            want = z[key][c][tuple(slice(0, 2 * n) if i == 0 else slice(0, n) for i in range(got.ndim))]
            scale = np.max(np.abs(want))
            # rounding-level agreement (the high Fourier terms are sums that nearly cancel)
            assert np.max(np.abs(got - want)) <= 1e-14 * scale, (c, key)


def test_white_sky_albedo_failure_drops_last_kernel():
    # total white-sky albedo > 1: the Li-sparse Fourier terms are not added, its direct bounce is
    kern = B.landmodis_setup(1.0, 0.01, 0.5, (2.0, 1.0))
    r = B.vbrdf_master(4, 30.0, 20.0, 40.0, kern)
    assert not r["ok"]
    r2 = B.vbrdf_master(4, 30.0, 20.0, 40.0, kern[:2], wsa=False)
    np.testing.assert_array_equal(r["F"], r2["F"])
    assert r["db"] != r2["db"]


def test_kernel_closed_forms():
    P = (2.0, 1.0, 0.0)
    # nadir illumination and view: both kernels vanish
    assert B.rossthick(P, 1.0, 0.0, 1.0, 0.0, 0.3, math.cos(0.3), math.sin(0.3)) == 0.0
    assert abs(B.lisparse(P, 1.0, 0.0, 1.0, 0.0, 0.3, math.cos(0.3), math.sin(0.3))) < 1e-13
    for th in (10.0, 35.0, 60.0):
        mu, s = math.cos(math.radians(th)), math.sin(math.radians(th))
        # hot spot (view along the illumination; this convention's relative azimuth pi):
        # Ross-thick = pi/(4 mu) - pi/4, Li-sparse (b/r = 1) = sec^2 - sec
        ross = B.rossthick(P, mu, s, mu, s, math.pi, -1.0, math.sin(math.pi))
        li = B.lisparse(P, mu, s, mu, s, math.pi, -1.0, math.sin(math.pi))
        assert abs(ross - (B.PIE / (4 * mu) - B.PIE / 4)) < 1e-13
        assert abs(li - (1 / mu ** 2 - 1 / mu)) < 1e-12
    # reciprocity k(i, v) = k(v, i)
    rng = np.random.default_rng(3)
    for _ in range(50):
        a, b = np.radians(rng.uniform(0, 80, 2))
        ph = rng.uniform(0, math.pi)
        args = (math.cos(ph), math.sin(ph))
        for fn in (B.rossthick, B.lisparse):
            k1 = fn(P, math.cos(a), math.sin(a), math.cos(b), math.sin(b), ph, *args)
            k2 = fn(P, math.cos(b), math.sin(b), math.cos(a), math.sin(a), ph, *args)
            assert abs(k1 - k2) <= 1e-13 * max(1.0, abs(k1))


def test_white_sky_integrals_match_published():
    # Lucht et al. (2000) white-sky integrals of the MODIS kernels: Ross-thick
    # 0.189184, Li-sparse (h/b = 2, b/r = 1) -1.377622; the supplement's
    # white-sky albedo is that integral over pi (Lambertian = 1)
    r = B.vbrdf_master(2, 30.0, 20.0, 40.0, [(B.ROSSTHICK_IDX, 1.0, (0, 0, 0)), (B.LISPARSE_IDX, 1.0, (2.0, 1.0, 0))])
    assert abs(r["wsa"][0] * math.pi / 0.189184 - 1) < 5e-5
    assert abs(r["wsa"][1] * math.pi / -1.377622 - 1) < 5e-5
    r = B.vbrdf_master(2, 30.0, 20.0, 40.0, [(B.LAMBERTIAN_IDX, 1.0, (0, 0, 0))])
    assert r["wsa"] == [1.0]


def test_fourier_series_reconstructs_kernel():
    # sum over M of F_M cos(M phi) approaches the kernel at the quadrature angles
    n = 8
    kern = [(B.ROSSTHICK_IDX, 1.0, (0, 0, 0))]
    sza, vza = 40.0, 25.0
    r = B.vbrdf_master(n, sza, vza, 0.0, kern, 50, wsa=False)
    mu0, muv = math.cos(sza * B.DEG_TO_RAD), math.cos(vza * B.DEG_TO_RAD)
    for phi in (0.3, 1.2, 2.5):
        series = sum(r["UF0"][m] * math.cos(m * phi) for m in range(2 * n))
        exact = B.rossthick(None, mu0, math.sqrt(1 - mu0 ** 2), muv, math.sqrt(1 - muv ** 2), phi, math.cos(phi),
                            math.sin(phi))
        assert abs(series - exact) < 2e-3


# ---------------------------------------------------------------- Cox-Munk
def test_coxmunk_vs_compiled_reference_supplement():
    z = np.load(os.path.join(BDIR, "supplement_cx.npz"), allow_pickle=False)
    for c, case in enumerate(z["cases"]):
        n = int(case[0])
        sza, vza, raa, u, v, m = (float(x) for x in case[1:])
        r = B.vbrdf_master(n, sza, vza, raa, B.coxmunk_setup(u, v, m), 100)
        assert r["ok"] == (int(z["status"][c]) == 0), c
        assert abs(r["db"] - z["db"][c]) <= 1e-15 * abs(z["db"][c]) + 1e-300, c
        for key, got in (("F", r["F"]), ("F0", r["F0"]), ("UF", r["UF"]), ("UF0", r["UF0"])):
            want = z[key][c][tuple(slice(0, 2 * n) if i == 0 else slice(0, n) for i in range(got.ndim))]
            assert np.max(np.abs(got - want)) <= 1e-14 * np.max(np.abs(want)), (c, key)


def test_coxmunk_closed_forms():
    # nadir sun and view: Fresnel reflectance at normal incidence ((m-1)/(m+1))^2
    # over 4 sigma^2, no shadowing (the shadow terms vanish at nadir)
    for m, (u, v) in ((1.34, (3.0, 4.0)), (1.33, (10.0, 0.0))):
        kern = B.coxmunk_setup(u, v, m)[0]
        s2 = kern[2][0]
        k = B.coxmunk((s2, m * m, 1.0), 1.0, 0.0, 1.0, 0.0, 0.0, 1.0, 0.0)
        assert abs(k - ((m - 1) / (m + 1)) ** 2 / (4 * s2)) <= 1e-15
    # sigma^2 = 0.003 + 0.00512 |w| with the reference's float32 literals, |w| capped at 50
    assert B.coxmunk_setup(3.0, 4.0, 1.34)[0][2][0] == float(np.float32(0.003)) + float(np.float32(0.00512)) * 5.0
    assert B.coxmunk_setup(60.0, 0.0, 1.34)[0][2][0] == B.coxmunk_setup(50.0, 0.0, 1.34)[0][2][0]
    # the supplement's erfc fit: fractional error below its published 1.2e-7
    for x in np.linspace(-2.0, 4.0, 61):
        assert abs(B.derfc_e(x) - math.erfc(x)) <= 1.2e-7 * math.erfc(x)
    # reciprocity of the glint kernel with shadowing
    rng = np.random.default_rng(4)
    for _ in range(50):
        a, b = np.radians(rng.uniform(1, 80, 2))
        ph = rng.uniform(0, math.pi)
        p = (0.03, 1.34 ** 2, 1.0)
        k1 = B.coxmunk(p, math.cos(a), math.sin(a), math.cos(b), math.sin(b), ph, math.cos(ph), math.sin(ph))
        k2 = B.coxmunk(p, math.cos(b), math.sin(b), math.cos(a), math.sin(a), ph, math.cos(ph), math.sin(ph))
        assert abs(k1 - k2) <= 1e-13 * max(abs(k1), 1e-300)
