"""The allocation-free lower-bound entry point (nigiri_amd_lower_bounds_noalloc,
static device scratch) against the allocating one (nigiri_amd_lower_bounds)
and the CPU oracle's Dijkstra: two calls in a row, the second with a smaller
timetable and fewer seeds, so the shared scratch must not leak the first
call's labels into the second."""
import numpy as np
import pytest

import nigiri_amd as na
import oracle
from hipmem import DeviceArray

pytestmark = pytest.mark.gpu


def device_lbs(img, seeds, direction, L, noalloc):
    out = DeviceArray(L, np.uint16)
    out.upload(np.full(L, 0x1234, np.uint16))
    img.lower_bounds(seeds, direction, out.ptr.value, noalloc=noalloc)
    got = out.download()
    out.free()
    return got


def oracle_lbs(tt, seeds, direction):
    qs = na.Queries.build([dict(start=[(9, 0)], dest=seeds, time=tt.view.date_range_from + 600,
                                dir=direction, dest_match_mode=na.MATCH_INTERMODAL)])
    return oracle.lower_bounds(tt, qs[0])


def test_no_alloc_lower_bounds_equal_the_allocating_entry_point():
    big = na.Timetable.generate(2, scale=0.1)
    small = na.Timetable.generate(1)
    cases = [
        (big, [(100, 0), (2000, 12), (4000, 30)], na.DIR_FWD),  # first call: larger input
        (small, [(500, 5)], na.DIR_BWD),                        # second call: smaller input
    ]
    images = {}
    for tt, seeds, direction in cases:
        img = images.setdefault(id(tt), na.Image(tt))
        L = tt.n_locations
        print(f"no-alloc lower bounds: L={L} seeds={seeds} dir={direction}", flush=True)
        got = device_lbs(img, seeds, direction, L, noalloc=True)
        ref = device_lbs(img, seeds, direction, L, noalloc=False)
        assert np.array_equal(got, ref)
        assert np.array_equal(got, oracle_lbs(tt, seeds, direction))
        assert (got != 0xFFFF).any() and got[seeds[0][0]] <= seeds[0][1]
