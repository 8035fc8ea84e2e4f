"""Serving loop: granules admitted while others run on the GPU (continuous
batching over several streams) give the same bits as each granule alone."""
import numpy as np
import pytest

from aeroapps_amd import chem_mie as AM
from aeroapps_amd import serving as SV
from aeroapps_amd import synthetic as S

pytestmark = pytest.mark.gpu


def test_granule_alone_equals_in_mix(gpu):
    import torch
    channels = (354.0, 388.0, 550.0, 870.0)
    nmom, N = 64, 6
    print("case serving mix", flush=True)
    tabs = S.make_mie_tables(channels, nmom=nmom, npol=1)
    mie = AM.MieTables.from_synthetic(tabs)
    pool = SV.ColumnPool(torch, 3000, 72, channels, torch.device("cuda:0"))
    grans = SV.make_granules(8, 3000, channels, seed=5, lo=50, hi=1500)
    # alone: one slot, one granule at a time
    alone = SV.Server(torch, mie, pool, N, nmom, slots=1, hbm_budget=60e9, arena_bytes=20e9, keep_results=True)
    ref = [SV.Granule(g.gid, g.ncol, g.channels, g.col0) for g in grans]
    for g in ref:
        alone.serve([g], [0.0])
    # in the mix: all arrive together, four slots
    mix = SV.Server(torch, mie, pool, N, nmom, slots=4, hbm_budget=200e9, arena_bytes=20e9, keep_results=True)
    stats = mix.serve(grans, [0.0] * len(grans))
    assert stats["granules"] == 8 and stats["p99_s"] >= stats["p50_s"] > 0
    for a, b in zip(ref, grans):
        assert np.array_equal(a.radiance, b.radiance), a.gid
        assert np.array_equal(a.nfourier, b.nfourier), a.gid
    mie.close()
