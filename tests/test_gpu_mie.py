"""GPU Chem_Mie optics (kernel a) through the C-ABI: bit-exact float32 against
the oracle restatement, including the reference's return codes."""
import numpy as np
import pytest

from aeroapps_amd import chem_mie as AM
from aeroapps_amd import synthetic as S
from oracle import chem_mie as CM

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("channels,nmom,npol,ncol", [
    ((354.0, 388.0, 550.0, 870.0), 64, 1, 64),
    ((550.0, 870.0), 32, 6, 32),
    ((550.0,), 32, 1, 257),
])
def test_get_aop_vector_bit_exact(gpu, channels, nmom, npol, ncol):
    print(f"case mie nch={len(channels)} nmom={nmom} npol={npol} ncol={ncol}", flush=True)
    tabs = S.make_mie_tables(channels, nmom=nmom, npol=npol)
    cols = S.make_columns(ncol, 72, channels, seed=99)
    mie = AM.MieTables.from_synthetic(tabs)
    t0, s0, g0, p0, rc0 = CM.get_aop_vector(tabs, channels, tabs[0].lambda_m, S.TRACERS, cols.qm, cols.rh, nmom, npol)
    t1, s1, g1, p1, rc1 = AM.get_aop_vector(mie, np.asarray(channels), S.TRACERS, 0, cols.qm.T, cols.rh.T, nmom, npol)
    assert rc1 == rc0 == 0
    assert np.array_equal(t1, t0.T)
    assert np.array_equal(s1, s0.T)
    assert np.array_equal(g1, g0.T)
    assert np.array_equal(p1, p0.T)
    mie.close()


def test_return_codes(gpu):
    channels = (550.0,)
    tabs = S.make_mie_tables(channels, nmom=16, npol=1)
    cols = S.make_columns(4, 8, channels, seed=1)
    mie = AM.MieTables.from_synthetic(tabs)
    # channel not in the tables
    *_, rc = AM.get_aop_vector(mie, np.asarray([600.0]), S.TRACERS, 0, cols.qm.T, cols.rh.T, 8, 1)
    assert rc == 99
    # more moments asked than the tables hold
    *_, rc = AM.get_aop_vector(mie, np.asarray(channels), S.TRACERS, 0, cols.qm.T, cols.rh.T, 32, 1)
    assert rc == 666
    # unknown last tracer name: the reference reports rc=1 and still sums the others
    names = list(S.TRACERS[:-1]) + ["nosuch"]
    t1, s1, g1, p1, rc = AM.get_aop_vector(mie, np.asarray(channels), names, 0, cols.qm.T, cols.rh.T, 8, 1)
    t0, s0, g0, p0, rc0 = CM.get_aop_vector(tabs, channels, tabs[0].lambda_m, names, cols.qm, cols.rh, 8, 1)
    assert rc == rc0 == 1
    assert np.array_equal(t1, t0.T) and np.array_equal(p1, p0.T)
    mie.close()


def test_query_matches_oracle(gpu):
    channels = (550.0, 870.0)
    tabs = S.make_mie_tables(channels, nmom=8, npol=1)
    mie = AM.MieTables.from_synthetic(tabs)
    rng = np.random.default_rng(0)
    q = rng.uniform(0, 1e-6, 500).astype(np.float32)
    rh = rng.uniform(-0.1, 1.2, 500).astype(np.float32)
    idx = mie.query_idx("DU::du003")
    assert idx == 2
    out = mie.query(idx, 2.0, q, rh, nmom=8, npol=1)
    t = tabs[0]
    rhi, rha = CM.rh_map(t.rh)
    tau, ssa, g, pm = CM.query(t.bext, t.bsca, t.g, t.pmom, rhi, rha, 2, 1, q, rh)
    assert np.array_equal(out["tau"], tau)
    assert np.array_equal(out["ssa"], ssa)
    assert np.array_equal(out["gasym"], g)
    mie.close()
