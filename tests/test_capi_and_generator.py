"""C ABI exports and synthetic timetable invariants (CPU only)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import nigiri_amd as na

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "nigiri_amd.h")).read()
    return sorted(set(re.findall(r"\b(nigiri_amd_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(na.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(na.SIGNATURES), set(syms) ^ set(na.SIGNATURES)


def test_oracle_exports():
    import oracle
    h = oracle.lib()
    for s in ("oracle_search", "oracle_lb", "oracle_fastest_direct", "oracle_starts",
              "oracle_search_many"):
        assert hasattr(h, s)


def test_bad_arguments_return_status_codes():
    lib = na.lib()
    h = C.c_void_p()
    assert lib.nigiri_amd_tt_generate(9, 1, 1.0, C.byref(h)) == na.EINVAL
    assert b"workload" in lib.nigiri_amd_last_error()
    assert lib.nigiri_amd_tt_generate(1, 1, 0.0, C.byref(h)) == na.EINVAL
    assert lib.nigiri_amd_engine_create(None, 64, 1, 16, C.byref(h)) == na.EINVAL


@pytest.mark.parametrize("cfg,scale", [(1, 1.0), (2, 0.05), (3, 0.01), (5, 0.004)])
def test_generated_timetable_invariants(cfg, scale):
    tt = na.Timetable.generate(cfg, scale=scale)
    v = tt.view
    L, R = v.n_locations, v.n_routes
    rso = tt.array("route_stop_offsets")
    rtf = tt.array("route_transport_from")
    rstf = tt.array("route_stop_time_from")
    ev = tt.array("event_times")
    stops = tt.array("route_stops")
    bf = tt.array("bitfields").reshape(-1, 8)
    tdays = tt.array("transport_traffic_days")
    rdays = tt.array("route_traffic_days")
    assert rso[-1] == len(stops) and rtf[-1] == v.n_transports and rstf[-1] == len(ev)
    for r in range(R):
        n = rso[r + 1] - rso[r]
        nt = rtf[r + 1] - rtf[r]
        assert n >= 2 and nt >= 1
        assert rstf[r + 1] - rstf[r] == 2 * (n - 1) * nt
        block = ev[rstf[r]:rstf[r + 1]].reshape(2 * (n - 1), nt)
        mam = block >> 5
        days = block & 31
        count = days.astype(np.int64) * 1440 + mam
        # times increase along the trip, FIFO order modulo 1440 across trips
        assert np.all(np.diff(count, axis=0) >= 0)
        assert np.all(np.diff(mam, axis=1) >= 0)
        # route traffic days cover every transport's traffic days
        for t in range(rtf[r], rtf[r + 1]):
            assert np.all((bf[tdays[t]] & ~bf[rdays[r]]) == 0)
        locs = stops[rso[r]:rso[r + 1]] & ((1 << 28) - 1)
        assert len(set(locs.tolist())) == len(locs) and locs.min() >= 9 and locs.max() < L
    # footpaths: out lists sorted by duration; in lists are the exact transpose
    fo, fp = tt.array("fp_out_offsets"), tt.array("fp_out")
    io, ip = tt.array("fp_in_offsets"), tt.array("fp_in")
    out_edges, in_edges = set(), set()
    for l in range(L):
        d = fp[fo[l]:fo[l + 1]] >> 23
        assert np.all(np.diff(d) >= 0)
        for x in fp[fo[l]:fo[l + 1]]:
            out_edges.add((l, int(x) & ((1 << 23) - 1), int(x) >> 23))
        for x in ip[io[l]:io[l + 1]]:
            in_edges.add((int(x) & ((1 << 23) - 1), l, int(x) >> 23))
    assert out_edges == in_edges
    # location_routes is the inverse of route_location_seq
    lro, lr = tt.array("location_route_offsets"), tt.array("location_routes")
    pairs = {(int(l), r) for r in range(R) for l in stops[rso[r]:rso[r + 1]] & ((1 << 28) - 1)}
    assert pairs == {(l, int(r)) for l in range(L) for r in lr[lro[l]:lro[l + 1]]}


def test_generator_is_deterministic():
    a = na.Timetable.generate(2, scale=0.02)
    b = na.Timetable.generate(2, scale=0.02)
    for name in ("event_times", "route_stops", "fp_out", "bitfields", "lb_fwd"):
        assert np.array_equal(a.array(name), b.array(name))
    qa = na.Queries.generate(a, 5, 50)
    qb = na.Queries.generate(b, 5, 50)
    for i in range(50):
        assert bytes(qa[i])[:32] == bytes(qb[i])[:32]


def test_workload_sizes_at_full_scale_match_the_spec():
    tt = na.Timetable.generate(1)
    assert tt.view.n_locations == 2009 and tt.view.n_transports == 4000
    qs = na.Queries.generate(tt, 5, 2000) if False else None
    tt3 = na.Timetable.generate(3, scale=0.05)
    qs = na.Queries.generate(tt3, 3, 1000)
    ranges = sum(qs[i].is_interval for i in range(1000))
    assert 240 < ranges < 360  # 30 % range queries
    windows = [qs[i].interval_to - qs[i].interval_from for i in range(1000) if qs[i].is_interval]
    assert min(windows) >= 60 and max(windows) <= 240
    qi = na.Queries.generate(tt3, 4, 200)
    assert all(qi[i].start_match_mode == na.MATCH_INTERMODAL and 1 <= qi[i].n_start <= 50
               for i in range(200))
