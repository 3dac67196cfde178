"""Pick table (aggregat_list.cpp:54-66, 109-141): the product's chunked table
with exact offsets against a full re-sort and plain sequential sums after
every merge (bit-exact cumulative sums, reference std::sort picks)."""
import pytest

import mcac_amd


@pytest.mark.parametrize("n,events,chunk,seed", [
    (50, 300, 16, 1), (1000, 990, 16, 2), (1000, 990, 64, 3), (5000, 1500, 128, 4),
    (3000, 2990, 32, 6), (200, 199, 16, 7), (2000, 1990, 2048, 8), (100000, 40, 2048, 9),
])
def test_pick_table_matches_resort(n, events, chunk, seed):
    assert mcac_amd.load_sim().mcac_pick_table_selftest(n, events, chunk, seed, 0) == 0


@pytest.mark.parametrize("n,events,chunk,seed,mode", [
    (17, 15, 16, 11, 1), (40, 38, 16, 12, 1), (3000, 600, 64, 13, 1), (20000, 60, 512, 14, 1),
    (200000, 8, 2048, 15, 1), (5000, 400, 128, 16, 2), (100000, 20, 2048, 17, 2), (600000, 3, 2048, 18, 2),
])
def test_pick_table_ties_follow_std_sort(n, events, chunk, seed, mode):
    """Picks inside runs of equal keys follow the reference's std::sort permutation."""
    L = mcac_amd.load_sim()
    t0 = L.mcac_pick_table_selftest_ties()
    assert L.mcac_pick_table_selftest(n, events, chunk, seed, mode) == 0
    if mode == 1:
        assert L.mcac_pick_table_selftest_ties() - t0 > events
