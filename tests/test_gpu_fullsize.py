"""The 500,000-column, 4-channel, 12-stream workload at its full size, through
the fused device path (inputs generated on the device, several optics blocks
and work chunks): size-independent properties of the whole run.

* Batch invariance: about 500 columns drawn across the whole range (with the first
  and last column and neighbours of the block boundaries) run alone give the
  same bits as inside the full run, in FP64 and in mixed precision.
* Parity: 32 of those columns (128 items) against the oracle (the C++
  restatement, itself pinned to the reference's drivers) fed the oracle's
  float32 Chem_Mie optics of the same inputs: FP64 radiance within 1e-9 and
  identical Fourier-term counts; mixed within 1e-5.
* Reflectance = pi I / (cos(SZA) F) element by element (the drivers' own
  relation), counts within [1, 2N], no MISSING.
"""
import numpy as np
import pytest

import oracle
from aeroapps_amd import _lib
from aeroapps_amd import chem_mie as AM
from aeroapps_amd import synthetic as S
from oracle import chem_mie as CM
from _fixtures import RTOL_FP64, max_rel

pytestmark = pytest.mark.gpu
RTOL_MIXED = 1e-5


def _run(torch, mie, d, cols, N, nmom, precision):
    """aa_column_radiance_dev over the columns `cols` (None: all) of the device inputs d."""
    dev = torch.device("cuda:0")
    sel = (lambda t, dim: t) if cols is None else (lambda t, dim: t.index_select(dim, cols).contiguous())
    qm, rh = sel(d["qm"], 0), sel(d["rh"], 0)
    rot, alb = sel(d["rot"], 1), sel(d["albedo"], 1)
    pe, he, te = sel(d["pe"], 0), sel(d["he"], 0), sel(d["te"], 0)
    sza, vza, raa, ff = sel(d["sza"], 0), sel(d["vza"], 0), sel(d["raa"], 0), sel(d["flux_factor"], 0)
    nch, ncol, km = rot.shape[0], rot.shape[1], rot.shape[2]
    tr = torch.tensor([mie.query_idx(n) for n in S.TRACERS], dtype=torch.int32, device=dev)
    ci = torch.arange(nch, dtype=torch.int32, device=dev)
    rad = torch.zeros((nch, ncol), dtype=torch.float64, device=dev)
    refl = torch.zeros_like(rad)
    nf = torch.zeros((nch, ncol), dtype=torch.int32, device=dev)
    p = lambda t: t.data_ptr()
    st = torch.cuda.current_stream(dev)
    e = _lib.lib().aa_column_radiance_dev(mie.handle, km, ncol, nch, len(S.TRACERS), p(tr), p(ci), p(qm), p(rh),
                                          nmom, 1, N, 0, precision, None, None, p(rot), p(d["depol"]), p(pe), p(he),
                                          p(te), p(alb), p(sza), p(raa), p(vza), p(ff), S.MISSING, p(rad), p(refl),
                                          None, None, p(nf), st.cuda_stream)
    _lib.check(e, "aa_column_radiance_dev")
    torch.cuda.synchronize()
    return rad.cpu().numpy(), refl.cpu().numpy(), nf.cpu().numpy()


def test_full_workload_properties(gpu):
    import torch
    wl = S.WORKLOADS["scalar_500k_4ch"]
    ncol, km, ch, N, nmom = wl["ncol"], wl["km"], wl["channels"], wl["nstreams"], wl["nmom"]
    dev = torch.device("cuda:0")
    print(f"case full size {ncol} columns x {len(ch)} channels N={N}", flush=True)
    parts = [S.make_columns_torch(min(100_000, ncol - c0), km, ch, seed=7, col0=c0, device=dev)
             for c0 in range(0, ncol, 100_000)]
    d = {k: torch.cat([p[k] for p in parts], dim=(1 if k in ("rot", "albedo") else 0)).contiguous()
         for k in ("qm", "rh", "rot", "albedo", "pe", "he", "te", "sza", "vza", "raa", "flux_factor")}
    d["depol"] = parts[0]["depol"]
    del parts
    tabs = S.make_mie_tables(ch, nmom=nmom, npol=1)
    mie = AM.MieTables.from_synthetic(tabs)
    rng = np.random.default_rng(2024)
    pick = set(rng.choice(ncol, 480, replace=False).tolist()) | {0, ncol - 1}
    for m in range(1, 10):   # around the optics-block edges (blocks of 55,556 columns at the bench arena)
        pick |= {55_556 * m - 1, 55_556 * m}
    pick = np.array(sorted(c for c in pick if 0 <= c < ncol))
    idx = torch.tensor(pick, dtype=torch.int64, device=dev)
    try:
        full = {}
        for precision in (0, 1):
            print(f"full run precision={precision}", flush=True)
            full[precision] = _run(torch, mie, d, None, N, nmom, precision)
            alone = _run(torch, mie, d, idx, N, nmom, precision)
            rad, refl, nf = full[precision]
            assert np.array_equal(alone[0], rad[:, pick]) and np.array_equal(alone[2], nf[:, pick])
        rad, refl, nf = full[0]
        # the drivers' reflectance relation and the count bounds, every element
        assert (nf >= 1).all() and (nf <= 2 * N).all() and (rad != S.MISSING).all()
        pif = float(np.float32(4.0) * np.arctan(np.float32(1.0)))
        mu0 = np.cos(d["sza"].cpu().numpy() * pif / 180.0)
        ffh = d["flux_factor"].cpu().numpy().T                      # [nch, ncol]
        assert max_rel(refl, pif * rad / (mu0[None, :] * ffh)) < 1e-14
        # parity of 32 columns with the oracle on the same (device-generated) inputs
        o = pick[::16][:32]
        f = lambda t: t.cpu().numpy()
        qm, rhh = f(d["qm"])[o], f(d["rh"])[o]
        tau, ssa, g, pm, rc = CM.get_aop_vector(tabs, ch, tabs[0].lambda_m, S.TRACERS, qm, rhh, nmom, 1)
        assert rc == 0
        P = oracle.port()
        f64 = lambda a: np.asarray(a, dtype=np.float64)
        I0, R0, nf0, rc0 = P.scalar_lambert(km, np.asarray(ch, float), N, nmom, f(d["rot"])[:, o], f(d["depol"]),
                                            f64(tau), f64(ssa), f64(g), f64(pm), f(d["pe"])[o], f(d["he"])[o],
                                            f(d["te"])[o], f(d["albedo"])[:, o], f(d["sza"])[o], f(d["raa"])[o],
                                            f(d["vza"])[o])
        assert rc0 == 0
        assert max_rel(rad[:, o], I0) < RTOL_FP64 and (nf[:, o] == nf0).all()
        radm, _, nfm = full[1]
        same = nfm[:, o] == nf0
        assert same.mean() > 0.95
        assert (np.abs(radm[:, o] - I0)[same] / np.abs(I0)[same]).max() < RTOL_MIXED
    finally:
        mie.close()
