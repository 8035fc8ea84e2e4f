"""GPU input preparation through the C-ABI against the oracle restatement:
Rayleigh optical thickness and depolarisation bit-exact (float64 with the
reference's float32 literals, same operation order, no contraction), edge
pressure/height and tracer column masses bit-exact (float32), edge
temperature within 4 float32 ulp (logf/powf of the device math library vs the
host's)."""
import numpy as np
import pytest

from aeroapps_amd import chem_mie as AM
from aeroapps_amd import vlidort as AV
from oracle import prep as P
from tests.test_prep import _log_linear_profile

pytestmark = pytest.mark.gpu


def test_rot_calc_bit_exact(gpu):
    print("case rot_calc km=72 nobs=40 nch=4", flush=True)
    km, nobs = 72, 40
    rng = np.random.default_rng(11)
    z = np.sort(rng.uniform(0.0, 70.0, (km + 1, nobs)), axis=0)[::-1]
    z[-1] = rng.uniform(0.0, 2.0, nobs)
    pe = 101325.0 * np.exp(-z / 7.4) * rng.uniform(0.95, 1.05, nobs)
    te = rng.uniform(190.0, 300.0, (km + 1, nobs))
    he = z * 1000.0
    ch = [354.0, 388.0, 550.0, 870.0]
    rot, depol, rc = AV.rot_calc(ch, pe, he, te, -1.0e20, 0)
    rot0, depol0 = P.rot_calc(ch, pe, he, te)
    assert rc == 0
    assert np.array_equal(rot, rot0)
    assert np.array_equal(depol, depol0)


def test_rot_calc_layer_limit(gpu):
    pe = np.ones((80, 2)); he = np.ones((80, 2)); te = np.ones((80, 2))
    _, _, rc = AV.rot_calc([550.0], pe, he, te, -1.0e20, 0)
    assert rc == 3


def test_edge_vars(gpu):
    print("case edge_vars km=72 nobs=300", flush=True)
    km, nobs = 72, 300
    ad, dp, ptop, *_ = _log_linear_profile(km, nobs, seed=7)
    pe, ze, te = AM.edge_vars(ad, dp, ptop)
    pe0, ze0, te0 = P.edge_vars(ad, dp, ptop)
    assert np.array_equal(pe, pe0)
    assert np.array_equal(ze, ze0)
    ulp = np.abs(te.view(np.int32).astype(np.int64) - te0.view(np.int32).astype(np.int64))
    assert ulp.max() <= 4


def test_get_edge_vars_transposed_input(gpu):
    class Aer:
        pass
    km, nobs = 20, 6
    ad, dp, ptop, *_ = _log_linear_profile(km, nobs, seed=9)
    aer = Aer()
    aer.AIRDENS, aer.DELP, aer.PS = ad.T.copy(), dp.T.copy(), np.ones(nobs)
    pe, ze, te = AM.getEdgeVars(aer, ptop)
    pe1, ze1, te1 = AM.edge_vars(ad, dp, ptop)
    assert np.array_equal(pe, pe1) and np.array_equal(ze, ze1) and np.array_equal(te, te1)


def test_calc_qm_bit_exact(gpu):
    rng = np.random.default_rng(13)
    q = rng.lognormal(np.log(1e-9), 1.5, (72, 15, 50)).astype(np.float32)
    dp = rng.uniform(200, 2000, (72, 50)).astype(np.float32)
    assert np.array_equal(AM.calc_qm(q, dp), P.calc_qm(q, dp))
