"""Uneven batches on the fused device path: per-column layer counts (terrain
truncation, 30..72 layers) and per-column stream counts (4, 8, 12, 16 streams)
in one call.  Each column must match the oracle run on that column alone
(its own layers and streams) within 1e-9 with the same Fourier terms, and a
column computed alone must give the same bits as in the mixed batch."""
import numpy as np
import pytest

import oracle
from aeroapps_amd import chem_mie as AM
from aeroapps_amd import synthetic as S
from oracle import chem_mie as CM
from _fixtures import RTOL_FP64, max_rel
from _gpu_helpers import fused

pytestmark = pytest.mark.gpu


def sub_columns(cols, idx):
    """Columns `idx` of a batch, as a new batch (same arrays sliced)."""
    import dataclasses
    return dataclasses.replace(cols, qm=cols.qm[idx], rh=cols.rh[idx], rot=cols.rot[:, idx], pe=cols.pe[idx],
                               he=cols.he[idx], te=cols.te[idx], albedo=cols.albedo[:, idx], sza=cols.sza[idx],
                               vza=cols.vza[idx], raa=cols.raa[idx], flux_factor=cols.flux_factor[idx],
                               nlayers=cols.nlayers[idx])


def test_uneven_layers_and_streams(gpu):
    import torch
    channels = (550.0,)
    nmom = 32
    ncol = 40
    print(f"case uneven ncol={ncol}", flush=True)
    cols = S.make_columns(ncol, 72, channels, seed=7, col0=300, uneven_layers=(30, 72))
    nstr = S.stream_choice(ncol, seed=7, col0=300)
    tabs = S.make_mie_tables(channels, nmom=nmom, npol=1)
    mie = AM.MieTables.from_synthetic(tabs)
    rad, refl, nf = fused(torch, mie, cols, channels, 0, nmom, 1, nlayers=cols.nlayers, nstreams_col=nstr)
    tau, ssa, g, pm, _ = CM.get_aop_vector(tabs, channels, tabs[0].lambda_m, S.TRACERS, cols.qm, cols.rh, nmom, 1)
    P = oracle.port()
    f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    for o in range(ncol):
        L, N = int(cols.nlayers[o]), int(nstr[o])
        sl = slice(o, o + 1)
        I0, R0, nf0, rc0 = P.scalar_lambert(L, cols.channels, N, nmom, f(cols.rot[:, sl, :L]), cols.depol,
                                            f(tau[sl, :, :L]), f(ssa[sl, :, :L]), f(g[sl, :, :L]),
                                            f(pm[:, :, sl, :, :L]), f(cols.pe[sl, :L + 1]), f(cols.he[sl, :L + 1]),
                                            f(cols.te[sl, :L + 1]), f(cols.albedo[:, sl]), cols.sza[sl],
                                            cols.raa[sl], cols.vza[sl])
        assert max_rel(rad[:, sl], I0) < RTOL_FP64, (o, L, N)
        assert (nf[:, sl] == nf0).all(), (o, L, N)
    # alone == in the mix, bit for bit
    for o in (0, 7, 23):
        one = sub_columns(cols, slice(o, o + 1))
        r1, _, n1 = fused(torch, mie, one, channels, 0, nmom, 1, nlayers=one.nlayers, nstreams_col=nstr[o:o + 1])
        assert np.array_equal(r1, rad[:, o:o + 1]) and np.array_equal(n1, nf[:, o:o + 1])
    mie.close()
