"""The oracle's BPDF-NDVI restatement (oracle/brdf.py: bpdf_ndvi,
fresnel_vector) pinned by the compiled reference's direct-bounce Mueller
elements (tests/golden/brdf/supplement_bpdf.npz: with a Lambertian first
kernel every element but (1,1) is the BPDF kernel's alone) and by closed
forms: at the specular geometry the facet is horizontal and the (1,1), (1,2)
elements are C e^-NDVI e^-tan(theta) / (4 (mu0 + mu)) times (R_par + R_per)/2
and (R_par - R_per)/2 with the textbook Fresnel reflectances; in the
principal plane the sine (U-coupling) elements vanish."""
import math
import os

import numpy as np

from oracle import brdf as OB

HERE = os.path.dirname(os.path.abspath(__file__))
DTR = math.pi / 180.0


def _geom(sza, vza, raa):
    mu0, muv = math.cos(sza * DTR), math.cos(vza * DTR)
    return mu0, math.sqrt(1.0 - mu0 * mu0), muv, math.sqrt(1.0 - muv * muv), math.cos(raa * DTR), math.sin(raa * DTR)


def test_bpdf_kernel_matches_reference_direct_bounce():
    z = np.load(os.path.join(HERE, "golden", "brdf", "supplement_bpdf.npz"), allow_pickle=False)
    for c, case in enumerate(z["cases"]):
        sza, vza, raa, modis, lamb = (float(v) for v in case[1:6])
        k = OB.bpdf_ndvi(case[10:13], *_geom(sza, vza, raa))
        db = z["DB"][c]
        for o1 in range(3):
            for o2 in range(3):
                q = 4 * o1 + o2
                want = db[q] - (lamb if (q == 0 and not modis) else 0.0)
                if q == 0 and modis:
                    continue    # (1,1) also holds the MODIS kernels
                # (elements that vanish analytically, e.g. at nadir view, are
                # rounding-level in both: the bound is relative to the block)
                assert abs(k[o1][o2] - want) <= 1e-12 * np.max(np.abs(db)), (c, o1, o2)


def test_bpdf_specular_closed_form():
    for theta in (15.0, 35.0, 60.0):
        for n, ndvi, cfac in ((1.5, 0.3, 6.0), (1.33, -0.5, 4.0), (1.5, 1.7, 5.0)):
            c, s = math.cos(theta * DTR), math.sin(theta * DTR)
            k = OB.bpdf_ndvi((n, ndvi, cfac), c, s, c, s, 1.0, 0.0)
            ct = math.sqrt(1.0 - s * s / (n * n))
            rper = (c - n * ct) / (c + n * ct)
            rpar = (n * c - ct) / (n * c + ct)
            nd = ndvi if -1.0 <= ndvi <= 1.0 else 0.0
            amp = cfac * math.exp(-nd) * math.exp(-math.tan(theta * DTR)) / (4.0 * (c + c))
            assert math.isclose(k[0][0], amp * (rpar ** 2 + rper ** 2) / 2.0, rel_tol=1e-12)
            assert math.isclose(k[0][1], amp * (rpar ** 2 - rper ** 2) / 2.0, rel_tol=1e-10)
            for o1, o2 in ((0, 2), (1, 2), (2, 0), (2, 1)):
                assert k[o1][o2] == 0.0


def test_bpdf_principal_plane_sine_elements_vanish():
    for sza, vza, raa in ((30.0, 50.0, 0.0), (45.0, 20.0, 180.0), (60.0, 10.0, 0.0)):
        k = OB.bpdf_ndvi((1.5, 0.4, 6.0), *_geom(sza, vza, raa))
        assert k[0][0] > 0.0
        for o1, o2 in ((0, 2), (1, 2), (2, 0), (2, 1)):
            assert abs(k[o1][o2]) <= 1e-14 * k[0][0]
