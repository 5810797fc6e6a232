"""Lower bounds through the contraction hierarchy (nigiri_amd_tt_lower_bounds,
the host half of the device PHAST sweep) against the CPU oracle's restatement
of nigiri's dial-bucket Dijkstra (dijkstra.cc:18-67), which is itself pinned
by Bellman-Ford; the values must be equal everywhere, cut-off included."""
import numpy as np
import pytest

import nigiri_amd as na
import oracle
from test_oracle_gtfs import transfer_time_tt

_TT = {}


def tt_for(key):
    if key not in _TT:
        cfg, scale = key
        _TT[key] = na.Timetable.generate(cfg, scale=scale)
    return _TT[key]


@pytest.mark.parametrize("key", [(1, 1.0), (2, 0.1), (3, 0.03)])
@pytest.mark.parametrize("direction", [na.DIR_FWD, na.DIR_BWD])
def test_ch_lower_bounds_equal_dijkstra(key, direction):
    tt = tt_for(key)
    rng = np.random.default_rng(17 + key[0])
    L = tt.n_locations
    for trial in range(12):
        n = 1 if trial < 6 else int(rng.integers(2, 12))
        seeds = [(int(rng.integers(9, L)), int(rng.integers(0, 40)) if n > 1 else 0) for _ in range(n)]
        qs = na.Queries.build([dict(start=[(9, 0)], dest=seeds, time=tt.view.date_range_from + 600,
                                    dir=direction, dest_match_mode=na.MATCH_INTERMODAL if n > 1 else 0)])
        want = oracle.lower_bounds(tt, qs[0])
        got = tt.lower_bounds(seeds, direction)
        assert np.array_equal(got, want), (trial, np.argwhere(got != want)[:5].tolist())


def test_ch_lower_bounds_cutoff_and_unreachable():
    # a far seed offset pushes every bound past the 7201-minute cut-off except
    # the ones below it; locations without a path stay unreachable
    tt = tt_for((1, 1.0))
    qs = na.Queries.build([dict(start=[(9, 0)], dest=[(100, 7150)], time=tt.view.date_range_from + 600,
                                dest_match_mode=na.MATCH_INTERMODAL)])
    want = oracle.lower_bounds(tt, qs[0])
    got = tt.lower_bounds([(100, 7150)], na.DIR_FWD)
    assert np.array_equal(got, want)
    assert (got == 0xFFFF).any() and (got[got != 0xFFFF] <= 7200).all()
    # the gtfs fixture: stops far apart, no footpaths; special stations unreachable
    g = transfer_time_tt()
    lb = g.lower_bounds([(g.location("C"), 0)], na.DIR_FWD)
    assert lb[g.location("C")] == 0 and (lb[:9] == 0xFFFF).all()
    assert lb[g.location("A")] == 19  # A -> B 10 min, B -> C 9 min (T2)
