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
