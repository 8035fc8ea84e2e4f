"""The launch-shape table of the radiance kernels (host side, no GPU): the
shapes each class accepts, set/get/clear, and the table file format."""
import os

import pytest

from aeroapps_amd import tune as T
from aeroapps_amd._lib import LIB_PATH, AAError

pytestmark = pytest.mark.skipif(not os.path.exists(LIB_PATH), reason="library not built")


def test_shape_candidates_per_class():
    # n = N*S <= 4: band solver 8 or 4 lanes; 5 <= n <= 8: 8 or 16; above: one shape
    assert T.shape_candidates(2, 1) == [(4, 16), (4, 64), (8, 16), (8, 64)]
    assert T.shape_candidates(6, 1) == [(8, 16), (8, 64), (16, 16), (16, 64)]
    assert T.shape_candidates(4, 3) == [(32, 16)]             # n = 12: 32 lanes, 16 pairs only
    assert T.shape_candidates(12, 1) == [(32, 16)]


def test_set_get_clear():
    d = T.get_tuning(6, 1, 1)
    assert d == {"bvp_lanes": 8, "sab_pairs": 64, "chunk_items": 0} or os.path.exists(T.DEFAULT_TABLE)
    T.set_tuning(6, 1, 1, bvp_lanes=16, sab_pairs=16, chunk_items=5000)
    assert T.get_tuning(6, 1, 1) == {"bvp_lanes": 16, "sab_pairs": 16, "chunk_items": 5000}
    T.clear_tuning(6, 1, 1)
    assert T.get_tuning(6, 1, 1) == {"bvp_lanes": 8, "sab_pairs": 64, "chunk_items": 0}
    for bad in (dict(bvp_lanes=4), dict(bvp_lanes=32), dict(sab_pairs=32), dict(chunk_items=-1)):
        with pytest.raises(AAError):
            T.set_tuning(6, 1, 1, **bad)
    with pytest.raises(AAError):
        T.set_tuning(6, 2, 1)


def test_table_file(tmp_path):
    good = tmp_path / "t.txt"
    good.write_text("# comment\n\n2 1 0 4 16 0   # times\n6 1 1 16 64 100000\n")
    T.load_tuning(str(good))
    assert T.get_tuning(2, 1, 0) == {"bvp_lanes": 4, "sab_pairs": 16, "chunk_items": 0}
    assert T.get_tuning(6, 1, 1) == {"bvp_lanes": 16, "sab_pairs": 64, "chunk_items": 100000}
    assert T.get_tuning(2, 1, 1)["bvp_lanes"] == 8        # classes not in the file: defaults
    bad = tmp_path / "b.txt"
    bad.write_text("6 1 1 12 64 0\n")                      # 12 lanes: no such shape
    with pytest.raises(AAError):
        T.load_tuning(str(bad))
    with pytest.raises(AAError):
        T.load_tuning(str(tmp_path / "none.txt"))
    # back to the library's own table (or the defaults when it has none)
    if os.path.exists(T.DEFAULT_TABLE):
        T.load_tuning(None)
    else:
        for c in ((2, 1, 0), (6, 1, 1)):
            T.clear_tuning(*c)
