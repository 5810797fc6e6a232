"""Device lower bounds through nigiri_amd_lower_bounds (upward search of the
contraction hierarchy on the host, PHAST sweep on the device) against the CPU
oracle's restatement of nigiri's dial-bucket Dijkstra (dijkstra.cc:18-67)."""
import numpy as np
import pytest

import nigiri_amd as na
import oracle
from hipmem import DeviceArray

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cfg,scale", [(1, 1.0), (2, 0.1), (3, 0.03)])
def test_gpu_lower_bounds_equal_dijkstra(cfg, scale):
    tt = na.Timetable.generate(cfg, scale=scale)
    img = na.Image(tt)
    L = tt.n_locations
    rng = np.random.default_rng(5 + cfg)
    for trial in range(8):
        direction = na.DIR_FWD if trial % 2 == 0 else na.DIR_BWD
        n = 1 if trial < 4 else int(rng.integers(2, 10))
        seeds = [(int(rng.integers(9, L)), 0 if n == 1 else int(rng.integers(0, 40))) for _ in range(n)]
        print(f"lower bounds cfg={cfg} trial={trial} seeds={seeds}", flush=True)
        out = DeviceArray(L, np.uint16)
        out.upload(np.full(L, 0x1234, np.uint16))
        img.lower_bounds(seeds, direction, out.ptr.value)
        got = out.download()
        out.free()
        qs = na.Queries.build([dict(start=[(9, 0)], dest=seeds, time=tt.view.date_range_from + 600,
                                    dir=direction, dest_match_mode=na.MATCH_INTERMODAL)])
        assert np.array_equal(got, oracle.lower_bounds(tt, qs[0]))


def footpath_neighbourhood(tt, l, hops, rng):
    """Locations within `hops` footpaths of l (forward footpaths)."""
    off, fp = tt.array("fp_out_offsets"), tt.array("fp_out")
    seen, frontier = {l}, [l]
    for _ in range(hops):
        nxt = []
        for u in frontier:
            for e in range(off[u], off[u + 1]):
                t = int(fp[e] & ((1 << 23) - 1))
                if t not in seen:
                    seen.add(t)
                    nxt.append(t)
        frontier = nxt
    return sorted(seen)


@pytest.mark.parametrize("cfg,scale", [(1, 1.0), (3, 0.03)])
def test_gpu_fastest_direct_equals_oracle(cfg, scale):
    # nigiri_amd_fastest_direct (get_fastest_direct.cc:12-100 on the device)
    # against the oracle: destinations among the starts' footpath
    # neighbours (1-3 footpaths away) and random far ones, both directions,
    # several start offsets, default and non-default transfer-time settings,
    # fastest_direct_factor 1 and 1.5, a query with its own value
    tt = na.Timetable.generate(cfg, scale=scale)
    img = na.Image(tt)
    L = tt.n_locations
    rng = np.random.default_rng(11 + cfg)
    specs = []
    for i in range(96):
        s0 = int(rng.integers(9, L))
        near = footpath_neighbourhood(tt, s0, 1 + i % 3, rng)
        starts = [(s0, int(rng.integers(0, 6)))]
        if i % 4 == 1:
            starts += [(int(x), int(rng.integers(0, 10))) for x in rng.choice(near, size=min(3, len(near)))]
        dests = [(int(x), int(rng.integers(0, 20))) for x in rng.choice(near, size=min(1 + i % 5, len(near)))]
        dests += [(int(rng.integers(9, L)), int(rng.integers(0, 20))) for _ in range(i % 3)]
        spec = dict(start=starts, dest=dests, time=tt.view.date_range_from + 600,
                    dir=na.DIR_FWD if i % 2 == 0 else na.DIR_BWD,
                    fastest_direct_factor=1.5 if i % 7 == 3 else 1.0)
        if i % 3 == 2:
            spec.update(tts_default=0, tts_factor=1.5 + (i % 2), tts_min_transfer_time=3 * (i % 4),
                        tts_additional_time=i % 5)
        if i == 40:
            spec["fastest_direct"] = 77
        specs.append(spec)
    qs = na.Queries.build(specs)
    print(f"fastest direct cfg={cfg} n={len(qs)}", flush=True)
    got = img.fastest_direct(qs)
    want = np.array([oracle.fastest_direct(tt, qs[i]) for i in range(len(qs))], np.int16)
    assert np.array_equal(got, want), np.argwhere(got != want)[:5]
    assert got[40] == 77
    assert (want != 32767).sum() > len(qs) // 4  # many of the queries have a direct walk
