"""Every launch shape the tuner may pick gives bit-identical radiances and
Fourier-term counts (fused device path, FP64 and mixed)."""
import pytest

from aeroapps_amd import tune as T

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nstreams,nstokes", [(2, 1), (6, 1), (2, 3)])
def test_all_shapes_bit_identical(gpu, nstreams, nstokes):
    import torch
    print(f"case tuning shapes N={nstreams} S={nstokes}", flush=True)
    res = T.tune_class(torch, nstreams, nstokes, (0, 1), ncol=300, reps=1, log=lambda m: None)
    assert set(res) == {0, 1}
    for prec, (best, times) in res.items():
        assert best in times and len(times) == len(T.shape_candidates(nstreams, nstokes))
