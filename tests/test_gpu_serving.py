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


def test_mixed_serving_fresh_tables_against_fp64(gpu):
    """The serving configuration as benchmarked: mixed precision, four slots,
    tables handed to the server before anything uploaded them (the first
    kernels of the four slots race to upload them; the table object
    serialises that).  Every granule is within the mixed-precision tolerance
    (1e-5 relative) of the same granule run alone in FP64, where the
    Fourier-term counts agree; fewer than 1% of the counts may differ."""
    import torch
    channels = (354.0, 388.0, 550.0, 870.0)
    nmom, N = 64, 6
    print("case serving mixed fresh tables", flush=True)
    pool = SV.ColumnPool(torch, 4000, 72, channels, torch.device("cuda:0"), seed=9)
    grans = SV.make_granules(8, 4000, channels, seed=21, lo=200, hi=2000)
    mie = AM.MieTables.from_synthetic(S.make_mie_tables(channels, nmom=nmom, npol=1))   # not yet uploaded
    mix = SV.Server(torch, mie, pool, N, nmom, slots=4, hbm_budget=200e9, arena_bytes=20e9, keep_results=True,
                    precision=1)
    stats = mix.serve(grans, [0.0] * len(grans))
    assert stats["granules"] == 8 and stats["failed"] == 0
    ref = [SV.Granule(g.gid, g.ncol, g.channels, g.col0) for g in grans]
    alone = SV.Server(torch, mie, pool, N, nmom, slots=1, hbm_budget=60e9, arena_bytes=20e9, keep_results=True,
                      precision=0)
    for g in ref:
        alone.serve([g], [0.0])
    differ = total = 0
    for a, b in zip(ref, grans):
        same = a.nfourier == b.nfourier
        differ += int((~same).sum())
        total += same.size
        rel = np.abs(b.radiance - a.radiance)[same] / np.abs(a.radiance)[same]
        assert rel.max() < 1e-5, (a.gid, rel.max())
    print(f"fourier-count mismatches {differ}/{total}", flush=True)
    assert differ < 0.01 * total
    mie.close()
