"""Host logic of the serving loop (no GPU): admission, slot reuse and
failure handling with a stand-in for the per-granule GPU run."""
import threading
import time

import pytest

from aeroapps_amd import serving as SV


class _Pool:
    km = 72


def _server(slots, fail_gids=(), budget=1e12):
    """A Server whose GPU run is replaced by a short sleep (or an error for
    the granules in fail_gids)."""
    sv = SV.Server.__new__(SV.Server)
    sv.slots = slots
    sv.budget = budget
    sv.in_use = 0.0
    sv.cv = threading.Condition()
    sv.done = []
    sv.free_slots = list(range(slots))
    sv.pool = _Pool()
    sv.ran = []

    def run(g, slot):
        if g.gid in fail_gids:
            raise RuntimeError("aa_column_radiance_dev failed: AA_EHIP (HIP runtime error)")
        time.sleep(0.01)
        sv.ran.append((g.gid, slot))
        g.finish = time.perf_counter()

    sv._run = run
    return sv


def _granules(n):
    return [SV.Granule(gid=k, ncol=1000 + k, channels=[550.0], col0=0) for k in range(n)]


def test_failed_granule_does_not_stop_the_loop():
    sv = _server(2, fail_gids={1, 4})
    grans = _granules(8)
    with pytest.raises(SV.ServingError) as ei:
        sv.serve(grans, [0.0] * 8)
    assert sorted(g.gid for g in ei.value.failed) == [1, 4]
    # every other granule ran, on both slots
    assert sorted(gid for gid, _ in sv.ran) == [0, 2, 3, 5, 6, 7]
    assert {slot for _, slot in sv.ran} == {0, 1}
    assert sv.in_use == 0.0 and sorted(sv.free_slots) == [0, 1]


def test_failures_reported_without_raising():
    sv = _server(3, fail_gids={0})
    stats = sv.serve(_granules(5), [0.0, 0.0, 0.01, 0.02, 0.02], raise_on_error=False)
    assert stats["failed"] == 1 and stats["granules"] == 5


def test_repeated_serve_calls_wait_for_their_own_granules():
    sv = _server(1)
    for k in range(3):
        g = SV.Granule(gid=k, ncol=10, channels=[550.0], col0=0)
        sv.serve([g], [0.0])
        assert g.finish > 0 and g.error is None


def test_admission_respects_budget():
    # budget for about one granule at a time: granules still all complete
    nb = SV.granule_bytes(1000, 72, 1)
    sv = _server(4, budget=1.5 * nb)
    stats = sv.serve(_granules(6), [0.0] * 6)
    assert stats["failed"] == 0 and len(sv.ran) == 6
