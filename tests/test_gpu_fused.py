"""The fused device path (aa_column_radiance_dev: Chem_Mie optics on the GPU
feeding the radiance kernels) equals the two-step host-pointer path
(aa_get_aop_vector, then aa_vlidort_scalar_lambert) bit for bit, and the
radiances match the oracle."""
import numpy as np
import pytest

import oracle
from aeroapps_amd import _lib
from aeroapps_amd import chem_mie as AM
from aeroapps_amd import synthetic as S
from aeroapps_amd import vlidort as V
from oracle import chem_mie as CM
from _fixtures import RTOL_FP64, max_rel

pytestmark = pytest.mark.gpu


def fused(torch, mie, cols, channels, N, nmom, npol, env=None):
    dev = torch.device("cuda:0")
    T = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    f32, f64 = torch.float32, torch.float64
    nch, ncol, km = len(channels), cols.nobs, cols.km
    qm, rh = T(cols.qm, f32), T(cols.rh, f32)
    rot, depol = T(cols.rot, f64), T(cols.depol, f64)
    pe, he, te = T(cols.pe, f64), T(cols.he, f64), T(cols.te, f64)
    alb, sza, raa, vza = T(cols.albedo, f64), T(cols.sza, f64), T(cols.raa, f64), T(cols.vza, f64)
    tr = T(np.array([mie.query_idx(n) for n in S.TRACERS]), torch.int32)
    ci = T(np.arange(nch), torch.int32)
    rad = torch.zeros((nch, ncol), dtype=f64, device=dev)
    refl = torch.zeros_like(rad)
    nf = torch.zeros((nch, ncol), dtype=torch.int32, device=dev)
    p = lambda t: t.data_ptr()
    st = torch.cuda.current_stream(dev)
    e = _lib.lib().aa_column_radiance_dev(mie.handle, km, ncol, nch, len(S.TRACERS), p(tr), p(ci), p(qm), p(rh),
                                          nmom, npol, N, 0, 0, None, None, p(rot), p(depol), p(pe), p(he), p(te),
                                          p(alb), p(sza), p(raa), p(vza), None, S.MISSING, p(rad), p(refl), None,
                                          None, p(nf), st.cuda_stream)
    _lib.check(e, "aa_column_radiance_dev")
    torch.cuda.synchronize()
    return rad.cpu().numpy(), refl.cpu().numpy(), nf.cpu().numpy()


@pytest.mark.parametrize("ncol,channels,N,nmom", [(24, (354.0, 388.0, 550.0, 870.0), 6, 64),
                                                  (33, (550.0,), 4, 32)])
def test_fused_equals_host_path(gpu, monkeypatch, ncol, channels, N, nmom):
    import torch
    print(f"case fused ncol={ncol} N={N}", flush=True)
    cols = S.make_columns(ncol, 72, channels, seed=31, col0=777)
    tabs = S.make_mie_tables(channels, nmom=nmom, npol=1)
    mie = AM.MieTables.from_synthetic(tabs)
    rad, refl, nf = fused(torch, mie, cols, channels, N, nmom, 1)
    # host two-step path
    tau, ssa, g, pm, rc = AM.get_aop_vector(mie, np.asarray(channels), S.TRACERS, 0, cols.qm.T, cols.rh.T, nmom, 1)
    I, R, rc1, nf1 = V.scalar_lambert(cols.channels, N, False, cols.rot.T, cols.depol, tau, ssa, g, pm, cols.pe.T,
                                      cols.he.T, cols.te.T, cols.albedo.T, cols.sza, cols.raa, cols.vza, S.MISSING,
                                      return_nfourier=True)
    assert np.array_equal(rad, I.T) and np.array_equal(refl, R.T) and np.array_equal(nf, nf1.T)
    # and the oracle
    t0, s0, g0, p0, _ = CM.get_aop_vector(tabs, channels, tabs[0].lambda_m, S.TRACERS, cols.qm, cols.rh, nmom, 1)
    f = lambda a: np.asarray(a, dtype=np.float64)
    I0, R0, nf0, _ = oracle.port().scalar_lambert(72, cols.channels, N, nmom, cols.rot, cols.depol, f(t0), f(s0),
                                                  f(g0), f(p0), cols.pe, cols.he, cols.te, cols.albedo, cols.sza,
                                                  cols.raa, cols.vza)
    assert max_rel(rad, I0) < RTOL_FP64
    assert (nf == nf0).all()
    # several optics blocks and work chunks give the same bits
    monkeypatch.setenv("AA_OPTICS_BYTES", str(72 * len(channels) * (3 + nmom) * 4 * 5))
    monkeypatch.setenv("AA_CHUNK_ITEMS", "7")
    rad2, refl2, nf2 = fused(torch, mie, cols, channels, N, nmom, 1)
    assert np.array_equal(rad, rad2) and np.array_equal(nf, nf2)
    # the blocks on two library streams instead of one after the other on the
    # caller stream
    monkeypatch.setenv("AA_BLOCK_STREAMS", "2")
    rad4, refl4, nf4 = fused(torch, mie, cols, channels, N, nmom, 1)
    monkeypatch.delenv("AA_BLOCK_STREAMS")
    assert np.array_equal(rad, rad4) and np.array_equal(refl, refl4) and np.array_equal(nf, nf4)
    # the same through budgets set by the API (they override the environment),
    # after the arenas were released
    monkeypatch.delenv("AA_OPTICS_BYTES")
    monkeypatch.delenv("AA_CHUNK_ITEMS")
    L = _lib.lib()
    _lib.check(L.aa_release_arenas(), "aa_release_arenas")
    _lib.check(L.aa_set_arena_bytes(1 << 20, 72 * len(channels) * (3 + nmom) * 4 * 3), "aa_set_arena_bytes")
    try:
        rad3, refl3, nf3 = fused(torch, mie, cols, channels, N, nmom, 1)
    finally:
        _lib.check(L.aa_set_arena_bytes(0, 0), "aa_set_arena_bytes")
    assert np.array_equal(rad, rad3) and np.array_equal(refl, refl3) and np.array_equal(nf, nf3)
    mie.close()
