"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle.

Bar: bit-exact journeys (start, dest_time, dest, transfers) and bit-exact
final state (round_times rows, best, tmp, time_at_dest, lower bounds) per query.
"""
import sys

import numpy as np
import pytest

import nigiri_amd as na
import oracle
from golden import raptor_abc as G
from test_oracle_golden import abc_timetable, hm

pytestmark = pytest.mark.gpu

STATE_KEYS = ("round_times", "best", "tmp", "time_at_dest", "lb")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


_images = {}


def image_for(key, make_tt):
    if key not in _images:
        tt = make_tt()
        _images[key] = (tt, na.Image(tt))
    return _images[key]


def compare(tt, eng, qs, idx=None, check_state=True):
    res = eng.search(qs)
    n = len(qs)
    for i in range(n):
        ref = oracle.search(tt, qs[i], state=check_state)
        got = res[i]
        assert got["status"] == 0
        assert got["journeys"] == ref["journeys"], (i, got["journeys"], ref["journeys"])
        assert got["legs"] == ref["legs"], (i, got["legs"], ref["legs"])
        assert got["base_day"] == ref["base_day"]
        assert got["n_rounds"] == ref["stats"]["n_rounds"], (i, got["n_rounds"], ref["stats"]["n_rounds"])
        if check_state and len(qs) <= eng.slots:
            st = eng.download(got["slot"])
            K = eng.max_k
            for key in STATE_KEYS:
                a = st[key]
                b = ref[key][:K] if key == "round_times" else ref[key]
                if not np.array_equal(a, b):
                    diff = np.argwhere(a != b)
                    raise AssertionError(f"query {i} {key} differs at {diff[:5].tolist()}: "
                                         f"gpu {a[tuple(diff[0])]} oracle {b[tuple(diff[0])]}")
    return res


def test_gpu_golden_abc_range_fwd_bwd():
    tt, loc, day0 = abc_timetable()
    img = na.Image(tt)
    eng = na.Engine(img, lanes=64, groups=1, max_k=16)
    iv = (day0 + hm(G.INTERVAL[0]), day0 + hm(G.INTERVAL[1]))
    for d, (s, t) in ((na.DIR_FWD, ("A", "C")), (na.DIR_BWD, ("C", "A"))):
        say(f"golden abc dir={d}")
        qs = na.Queries.build([dict(start=[(loc[s], 0)], dest=[(loc[t], 0)], interval=iv, dir=d)])
        compare(tt, eng, qs)


def test_gpu_golden_abc_intermodal_fwd():
    tt, loc, day0 = abc_timetable()
    img = na.Image(tt)
    eng = na.Engine(img, lanes=64, groups=1, max_k=16)
    iv = (day0 + hm(G.INTERVAL[0]), day0 + hm(G.INTERVAL[1]))
    (s, sd), (d, dd) = G.INTERMODAL_START, G.INTERMODAL_DEST
    say("golden abc intermodal fwd")
    qs = na.Queries.build([dict(start=[(loc[s], sd)], dest=[(loc[d], dd)], interval=iv,
                                start_match_mode=na.MATCH_INTERMODAL,
                                dest_match_mode=na.MATCH_INTERMODAL)])
    res = compare(tt, eng, qs)
    want = [(day0 + hm(a), day0 + hm(b), 1, k) for a, b, k in G.INTERMODAL_FWD_EXPECTED]
    assert sorted(res[0]["journeys"]) == sorted(want)


@pytest.mark.parametrize("lanes", [64, 1])
def test_gpu_point_queries_workload1(lanes):
    tt, img = image_for("w1", lambda: na.Timetable.generate(1))
    groups = 2 if lanes == 64 else 8
    eng = na.Engine(img, lanes=lanes, groups=groups, max_k=16)
    qs = na.Queries.generate(tt, 1, 96 if lanes == 64 else 8)
    say(f"workload1 point lanes={lanes} n={len(qs)}")
    compare(tt, eng, qs)


def test_gpu_large_batch_workload1():
    # a batch of four groups: the lower-bound sweeps of all four run together
    tt, img = image_for("w1", lambda: na.Timetable.generate(1))
    eng = na.Engine(img, lanes=64, groups=4, max_k=16)
    qs = na.Queries.generate(tt, 1, 256, seed=0x7172 + 11)
    say("workload1 large batch n=256")
    compare(tt, eng, qs)


def test_gpu_backward_point_queries_workload1():
    tt, img = image_for("w1", lambda: na.Timetable.generate(1))
    eng = na.Engine(img, lanes=64, groups=1, max_k=16)
    qs = na.Queries.generate(tt, 1, 64)
    for i in range(64):
        qs[i].dir = na.DIR_BWD
    say("workload1 backward n=64")
    compare(tt, eng, qs)


def test_gpu_range_and_point_mix_workload3():
    tt, img = image_for("w3s", lambda: na.Timetable.generate(3, scale=0.03))
    eng = na.Engine(img, lanes=64, groups=2, max_k=10)
    qs = na.Queries.generate(tt, 3, 128)
    say("workload3 mix n=128")
    compare(tt, eng, qs)


def test_gpu_intermodal_workload4():
    tt, img = image_for("w3s", lambda: na.Timetable.generate(3, scale=0.03))
    eng = na.Engine(img, lanes=64, groups=1, max_k=10)
    qs = na.Queries.generate(tt, 4, 64)
    say("workload4 intermodal n=64")
    compare(tt, eng, qs)


def test_gpu_workload2_plain_and_batched_agree():
    tt, img = image_for("w2s", lambda: na.Timetable.generate(2, scale=0.1))
    qs = na.Queries.generate(tt, 2, 128)
    say("workload2 batched n=128")
    eng = na.Engine(img, lanes=64, groups=2, max_k=16)
    compare(tt, eng, qs, check_state=False)
    say("workload2 plain n=16")
    plain = na.Engine(img, lanes=1, groups=1, max_k=16)
    compare(tt, plain, qs.subset(list(range(16))), check_state=False)


def test_gpu_more_queries_than_slots_reuses_slots():
    tt, img = image_for("w1", lambda: na.Timetable.generate(1))
    eng = na.Engine(img, lanes=64, groups=1, max_k=9)
    qs = na.Queries.generate(tt, 1, 300)
    say("workload1 300 queries through 64 slots")
    compare(tt, eng, qs, check_state=False)


def test_gpu_served_results_equal_alone_and_in_mix():
    tt, img = image_for("w3s", lambda: na.Timetable.generate(3, scale=0.03))
    qs = na.Queries.generate(tt, 5, 96, seed=0x7172 + 3)
    say("serve mix n=96")
    eng = na.Engine(img, lanes=64, groups=1, max_k=10)
    arrivals = np.cumsum(np.random.default_rng(1).exponential(2.0, len(qs)))
    served, wall = eng.serve(qs, arrivals)
    alone = na.Engine(img, lanes=64, groups=1, max_k=10)
    for i in range(len(qs)):
        one = alone.search(qs.subset([i]))[0]
        assert served[i]["journeys"] == one["journeys"], i
        assert served[i]["legs"] == one["legs"], i
        ref = oracle.search(tt, qs[i])
        assert served[i]["journeys"] == ref["journeys"] and served[i]["legs"] == ref["legs"], i
        assert served[i]["latency_ms"] >= 0.0


def with_destinations(qs, n_locations, rng, n_exact, n_egress):
    """Even queries get n_exact exact destinations, odd ones n_egress
    intermodal egress offsets (0-30 min), at random locations."""
    L = n_locations
    for i in range(len(qs)):
        q = qs[i]
        n = n_exact if i % 2 == 0 else n_egress
        loc = rng.choice(np.arange(9, L), size=n, replace=False).astype(np.uint32)
        dur = (rng.integers(0, 31, size=n) if i % 2 else np.zeros(n)).astype(np.int16)
        q.dest_match_mode = na.MATCH_EXACT if i % 2 == 0 else na.MATCH_INTERMODAL
        q.n_dest = n
        q.dest_loc = loc.ctypes.data_as(q.dest_loc.__class__)
        q.dest_dur = dur.ctypes.data_as(q.dest_dur.__class__)
        qs._keep += [loc, dur]
    return qs


def test_gpu_many_destinations_and_egress_offsets():
    # 40 exact destinations and 200 egress offsets per query (nigiri's
    # collect_destinations has no limit; the engines take up to 64 and 256)
    tt, img = image_for("w3s", lambda: na.Timetable.generate(3, scale=0.03))
    qs = na.Queries.generate(tt, 3, 32, seed=0x7172 + 83)
    with_destinations(qs, tt.n_locations, np.random.default_rng(83), 40, 200)
    say("many destinations batched n=32")
    compare(tt, na.Engine(img, lanes=64, groups=1, max_k=10), qs)
    say("many destinations plain n=4")
    compare(tt, na.Engine(img, lanes=1, groups=1, max_k=10), qs.subset([0, 1, 2, 3]))


def test_gpu_work_counters_equal_oracle_counters():
    # the algorithmic byte model counts the same units on both sides
    tt, img = image_for("w3s", lambda: na.Timetable.generate(3, scale=0.03))
    eng = na.Engine(img, lanes=64, groups=1, max_k=10)
    qs = na.Queries.generate(tt, 3, 64)
    say("work counters n=64")
    eng.stats(reset=True)
    eng.search(qs)
    w = eng.stats()["work"]
    ref = {"route_stops": 0, "events": 0, "route_day_tests": 0, "transport_day_tests": 0,
           "marked_stations": 0, "location_routes": 0, "footpaths": 0, "transfer_stations": 0}
    for i in range(len(qs)):
        s = oracle.search(tt, qs[i])["stats"]
        ref["route_stops"] += s["n_route_stops_visited"]
        ref["events"] += s["n_events_scanned"]
        ref["route_day_tests"] += s["n_route_day_tests"]
        ref["transport_day_tests"] += s["n_transport_day_tests"]
        ref["marked_stations"] += s["n_marked_stations"]
        ref["location_routes"] += s["n_location_routes_read"]
        ref["footpaths"] += s["n_footpaths_visited"]
        ref["transfer_stations"] += s["n_transfer_stations"]
    for k, v in ref.items():
        assert w[k] == v, (k, w[k], v)


def test_gpu_batched_engine_routes_beyond_its_lds_staging():
    # routes of 300 and 257 transports (beyond the LDS event staging) next to
    # a small one: the batched engine scans them lane by lane from global
    # memory; both engines equal the oracle, state included
    from ttbuilder import Builder
    b = Builder(20000, 6, 6)
    day = 20001
    b.add_line([9, 10, 11], [([i * 4, i * 4 + 5, i * 4 + 6, i * 4 + 12], [day]) for i in range(300)])
    b.add_line([11, 12, 13], [([i * 5 + 3, i * 5 + 9, i * 5 + 10, i * 5 + 15], [day, day + 1])
                              for i in range(257)])
    b.add_line([9, 13], [([i * 30 + 1, i * 30 + 50], [day]) for i in range(20)])
    tt = b.build()
    img = na.Image(tt)
    # max_transfers 6: end_k 8 rows, what the engines below hold
    specs = [dict(start=[(9, 0)], dest=[(13, 0)], time=day * 1440 + 100 + 37 * i, max_transfers=6)
             for i in range(40)]
    specs += [dict(start=[(13, 0)], dest=[(9, 0)], time=day * 1440 + 900 + 11 * i, dir=na.DIR_BWD,
                   max_transfers=6) for i in range(8)]
    specs += [dict(start=[(10, 0)], dest=[(12, 0)], interval=(day * 1440 + 60, day * 1440 + 240),
                   max_transfers=6)]
    for lanes in (64, 1):
        eng = na.Engine(img, lanes=lanes, groups=1 if lanes == 64 else 4, max_k=8)
        for d in (na.DIR_FWD, na.DIR_BWD):
            qs = na.Queries.build([q for q in specs if q.get("dir", na.DIR_FWD) == d])
            say(f"routes of 300/257 transports, lanes={lanes} dir={d} n={len(qs)}")
            compare(tt, eng, qs, check_state=len(qs) <= eng.slots)


def test_gpu_algo_reconstruct_equals_search_legs():
    # raptor::reconstruct through the Algo entry point, against the state the
    # (single) epoch of each point query left in its slot
    tt, img = image_for("w1", lambda: na.Timetable.generate(1))
    eng = na.Engine(img, lanes=64, groups=1, max_k=16)
    qs = na.Queries.generate(tt, 1, 64, seed=0x7172 + 5)
    say("algo reconstruct workload1 n=64")
    res = compare(tt, eng, qs, check_state=False)
    n = 0
    for i in range(len(qs)):
        for j, legs in zip(res[i]["journeys"], res[i]["legs"]):
            assert eng.reconstruct(res[i]["slot"], qs[i], j) == legs
            n += 1
    assert n > 0


def test_gpu_golden_intermodal_bwd_interval_extension():
    # raptor_intermodal_test.cc:85-251: min_connection_count 3 widens the
    # interval; the epochs of the new parts run on the same slot
    from test_oracle_golden import intermodal_bwd_query
    tt, loc, day0 = abc_timetable()
    img = na.Image(tt)
    eng = na.Engine(img, lanes=64, groups=1, max_k=16)
    say("golden abc intermodal bwd with interval extension")
    res = compare(tt, eng, intermodal_bwd_query(loc, day0))
    assert len(res[0]["journeys"]) == len(G.INTERMODAL_BWD_EXPECTED)


@pytest.mark.parametrize("lanes", [64, 1])
def test_gpu_interval_extension_workload3(lanes):
    # min_connection_count with each extension direction, forward and backward
    # (early termination for one-sided extensions, resets when extending the
    # wrong way)
    tt, img = image_for("w3s", lambda: na.Timetable.generate(3, scale=0.03))
    eng = na.Engine(img, lanes=lanes, groups=1 if lanes == 64 else 8, max_k=10)
    n = 48 if lanes == 64 else 8
    qs = na.Queries.generate(tt, 3, n, seed=0x7172 + 31)
    for i in range(n):
        qs[i].min_connection_count = (2, 3, 5)[i % 3]
        qs[i].extend_interval_earlier, qs[i].extend_interval_later = ((1, 0), (0, 1), (1, 1))[(i // 3) % 3]
    # one direction per search call: a search runs its backward queries after
    # the forward ones, on the same slots, so only one direction's states remain
    say(f"workload3 interval extension lanes={lanes} n={n} fwd")
    compare(tt, eng, qs)
    for i in range(n):
        qs[i].dir = na.DIR_BWD
    say(f"workload3 interval extension lanes={lanes} n={n} bwd")
    compare(tt, eng, qs)


def test_gpu_graph_rounds_equal_eager_rounds():
    # the round sequence replayed from captured hipGraphs gives the eager
    # launches' journeys, legs and state; each graph is replayed many times
    # with the runner's own copies, epoch starts and syncs between replays
    tt, img = image_for("w3s", lambda: na.Timetable.generate(3, scale=0.03))
    qs = na.Queries.generate(tt, 3, 128, seed=0x7172 + 41)
    eager = na.Engine(img, lanes=64, groups=2, max_k=10)
    eager.configure(graphs=False)
    graph = na.Engine(img, lanes=64, groups=2, max_k=10)
    graph.stats(reset=True)
    say("graph rounds vs eager rounds, workload3 mix n=128")
    a = eager.search(qs)
    b = compare(tt, graph, qs)
    st = graph.stats()
    assert st["graph_rounds"] == st["rounds"] > 0
    # one graph per (direction, sweep on/off): captured once, replayed
    assert 1 <= st["graph_captures"] <= 2, st["graph_captures"]
    assert eager.stats()["graph_rounds"] == 0
    for i in range(len(qs)):
        assert a[i]["journeys"] == b[i]["journeys"], i
        assert a[i]["legs"] == b[i]["legs"], i
        sa, sb = eager.download(a[i]["slot"]), graph.download(b[i]["slot"])
        for key in STATE_KEYS:
            assert np.array_equal(sa[key], sb[key]), (i, key)
    # profiling splits the round into per-segment graphs with events between
    graph.profile(True)
    c = graph.search(qs)
    assert [r["journeys"] for r in c] == [r["journeys"] for r in a]
    st = graph.stats()
    assert st["route_scan_launches"] > 0 and st["route_scan_ms"] > 0.0


def test_gpu_launch_knobs_do_not_change_results():
    # every launch-shape knob away from its default: results stay bit-exact
    tt, img = image_for("w3s", lambda: na.Timetable.generate(3, scale=0.03))
    qs = na.Queries.generate(tt, 3, 128, seed=0x7172 + 43)
    eng = na.Engine(img, lanes=64, groups=2, max_k=10)
    eng.configure(grid_waves=512, route_claim=4, transfer_claim=3, pack_max_entries=1,
                  transfer_pack_max_entries=2, loose_lanes=17)
    cfg = eng.get_config()
    assert (cfg["grid_waves"], cfg["route_claim"], cfg["transfer_claim"], cfg["pack_max_entries"],
            cfg["transfer_pack_max_entries"], cfg["loose_lanes"]) == (512, 4, 3, 1, 2, 17)
    say("knobs away from defaults, workload3 mix n=128")
    compare(tt, eng, qs)
    eng.configure(route_claim=2, transfer_claim=64, pack_max_entries=3,
                  transfer_pack_max_entries=1, loose_lanes=1)
    say("knobs, second setting")
    compare(tt, eng, qs)


def test_gpu_autotune_picks_candidates_and_keeps_results():
    tt, img = image_for("w3s", lambda: na.Timetable.generate(3, scale=0.03))
    qs = na.Queries.generate(tt, 3, 64, seed=0x7172 + 47)
    eng = na.Engine(img, lanes=64, groups=1, max_k=10)
    say("autotune workload3 n=64")
    chosen, trials = eng.autotune(qs)
    assert len(trials) == 4 + 3 + 3 + 5 + 5 + 4
    for knob, value in chosen.items():
        tried = [(v, ms) for k, v, ms in trials if k == knob]
        assert value in [v for v, _ in tried], (knob, value)
        assert all(ms > 0.0 for _, ms in tried)
    assert eng.get_config()["pack_max_entries"] == chosen["pack_max_entries"]
    compare(tt, eng, qs)


def test_gpu_whole_batch_admission_serving():
    # admit_when_idle: queries queued behind a running batch start together
    # once it has finished; results equal the oracle's
    tt, img = image_for("w1", lambda: na.Timetable.generate(1))
    qs = na.Queries.generate(tt, 1, 128, seed=0x7172 + 53)
    eng = na.Engine(img, lanes=64, groups=1, max_k=9)
    eng.configure(admission_wait_ms=500.0, admit_when_idle=True)
    assert eng.get_config()["admit_when_idle"] == 1
    say("whole-batch admission, 128 queries through 64 slots")
    res, wall = eng.serve(qs, np.zeros(len(qs)))
    # queries without start labels complete at admission without a slot,
    # so the batches are told apart by admission time, not by index: a
    # batch starts together (one admission step), and the second only once
    # every slot of the first is idle (its last queries complete, legs
    # written, right after that admission)
    queue = np.array([r["queue_ms"] for r in res])
    lat = np.array([r["latency_ms"] for r in res])
    t_second = queue.max()
    second = queue > queue.min() + 0.5 * (t_second - queue.min())
    first = ~second
    assert 64 - 8 <= second.sum() <= 64, second.sum()
    assert queue[second].max() - queue[second].min() < 1.0, queue[second]
    assert queue[first].max() - queue[first].min() < 1.0, queue[first]
    assert t_second >= lat[first].max() - 5.0, (t_second, lat[first].max())
    assert t_second <= lat[first].max() + 1e-3, (t_second, lat[first].max())
    for i in range(0, 128, 8):
        ref = oracle.search(tt, qs[i])
        assert res[i]["journeys"] == ref["journeys"], i
        assert res[i]["legs"] == ref["legs"], i


# ---- full configured sizes (samples small enough for the oracle) -----------

def test_gpu_full_size_workload2_sample():
    # 50k stops, 365-day bitfields: 64 batched queries with state, 8 plain
    tt, img = image_for("w2", lambda: na.Timetable.generate(2))
    qs = na.Queries.generate(tt, 2, 64, seed=0x7172 + 61)
    say("full-size workload2 batched n=64")
    compare(tt, na.Engine(img, lanes=64, groups=1, max_k=16), qs)
    say("full-size workload2 plain n=8")
    compare(tt, na.Engine(img, lanes=1, groups=1, max_k=16), qs.subset(list(range(8))), check_state=False)


def test_gpu_full_size_workload3_mix_sample():
    # 200k stops: point and range queries (1-4 h windows) and intermodal ones
    tt, img = image_for("w3", lambda: na.Timetable.generate(3))
    qs = na.Queries.generate(tt, 3, 24, seed=0x7172 + 67)
    say("full-size workload3 point/range n=24")
    compare(tt, na.Engine(img, lanes=64, groups=1, max_k=10), qs)
    qi = na.Queries.generate(tt, 4, 16, seed=0x7172 + 71)
    say("full-size workload4 intermodal n=16")
    compare(tt, na.Engine(img, lanes=64, groups=1, max_k=10), qi)


def test_gpu_full_size_workload5_mix_sample():
    # 500k stops, the serving mix (point / range / intermodal)
    tt, img = image_for("w5", lambda: na.Timetable.generate(5))
    qs = na.Queries.generate(tt, 5, 12, seed=0x7172 + 73)
    say("full-size workload5 mix n=12")
    compare(tt, na.Engine(img, lanes=64, groups=1, max_k=10), qs)


def test_gpu_full_size_workload5_served_mix():
    # the serving loop on the 500k-stop timetable: 48 queries of the mix arrive
    # (Poisson, 5 ms apart on average) while others run on 128 slots; every
    # result equals the same query searched alone on an idle engine, and the
    # first 12 (the queries of the sample above) equal the oracle's
    tt, img = image_for("w5", lambda: na.Timetable.generate(5))
    qs = na.Queries.generate(tt, 5, 48, seed=0x7172 + 73)
    say("full-size workload5 served n=48")
    eng = na.Engine(img, lanes=64, groups=2, max_k=10)
    arrivals = np.cumsum(np.random.default_rng(5).exponential(5.0, len(qs)))
    served, _ = eng.serve(qs, arrivals)
    alone = na.Engine(img, lanes=64, groups=1, max_k=10)
    for i in range(len(qs)):
        one = alone.search(qs.subset([i]))[0]
        assert served[i]["status"] == 0 and one["status"] == 0, i
        assert served[i]["journeys"] == one["journeys"], i
        assert served[i]["legs"] == one["legs"], i
    say("full-size workload5 served: oracle for the first 12")
    for i in range(12):
        ref = oracle.search(tt, qs[i])
        assert served[i]["journeys"] == ref["journeys"] and served[i]["legs"] == ref["legs"], i


def compare_pong(tt, eng, q, check_work=False):
    """The plain engine's pong against the oracle's: journeys, legs, interval and
    the number of executes; with check_work also the route-scan work counters,
    which the ping bounds prune."""
    if check_work:
        eng.stats(reset=True)
    got = eng.pong(q)
    ref = oracle.pong(tt, q)
    assert got["status"] == 0
    assert got["journeys"] == ref["journeys"], (got["journeys"], ref["journeys"])
    assert got["legs"] == ref["legs"]
    assert got["interval"] == ref["interval"]
    assert got["n_epochs"] == ref["stats"]["n_epochs"]
    if check_work:
        w = eng.stats()["work"]
        s = ref["stats"]
        assert (w["route_stops"], w["events"], w["route_day_tests"], w["transport_day_tests"]) == (
            s["n_route_stops_visited"], s["n_events_scanned"], s["n_route_day_tests"],
            s["n_transport_day_tests"])
    return got


def test_gpu_pong_golden():
    # pong_minlookahead_test.cc and pong_gap_truncation_test.cc on the plain
    # engine, plus the golden range queries of the A-B-C network
    from golden import pong as P
    from test_oracle_pong import line_tt
    say("pong golden: min lookahead")
    tt, loc, day0 = line_tt(P.MINLOOKAHEAD)
    eng = na.Engine(na.Image(tt), lanes=1, groups=2, max_k=16)
    for when, max_tt, want in P.MINLOOKAHEAD["CASES"]:
        spec = dict(start=[(loc["S0"], 0)], dest=[(loc["S2"], 0)], max_travel_time=max_tt,
                    min_connection_count=1, extend_interval_later=1)
        if isinstance(when, tuple):
            spec["interval"] = (day0 + hm(when[0]), day0 + hm(when[1]))
        else:
            spec["time"] = day0 + hm(when)
        got = compare_pong(tt, eng, na.Queries.build([spec])[0])
        assert [j[:2] for j in got["journeys"]] == [(day0 + hm(d), day0 + hm(a)) for d, a, _ in want]
    say("pong golden: gap truncation")
    tt, loc, day0 = line_tt(P.GAP)
    eng = na.Engine(na.Image(tt), lanes=1, groups=2, max_k=16)
    iv = (day0 + hm(P.GAP["INTERVAL"][0]), day0 + hm(P.GAP["INTERVAL"][1]))
    q = na.Queries.build([dict(start=[(loc["A"], 0)], dest=[(loc["B"], 0)], interval=iv,
                               min_connection_count=P.GAP["MIN_CONNECTIONS"],
                               extend_interval_later=1)])[0]
    got = compare_pong(tt, eng, q)
    assert sorted(j[0] for j in got["journeys"]) == [day0 + hm(t) for t in P.GAP["EXPECTED_DEPARTURES"]]
    say("pong golden: abc range and intermodal, both directions")
    tt, loc, day0 = abc_timetable()
    eng = na.Engine(na.Image(tt), lanes=1, groups=2, max_k=16)
    iv = (day0 + hm(G.INTERVAL[0]), day0 + hm(G.INTERVAL[1]))
    (s, sd), (d, dd) = G.INTERMODAL_START, G.INTERMODAL_DEST
    specs = [dict(start=[(loc["A"], 0)], dest=[(loc["C"], 0)], interval=iv),
             dict(start=[(loc["C"], 0)], dest=[(loc["A"], 0)], interval=iv, dir=na.DIR_BWD),
             dict(start=[(loc[s], sd)], dest=[(loc[d], dd)], interval=iv,
                  start_match_mode=na.MATCH_INTERMODAL, dest_match_mode=na.MATCH_INTERMODAL)]
    qs = na.Queries.build(specs)
    for i in range(len(qs)):
        assert len(compare_pong(tt, eng, qs[i])["journeys"]) > 0


@pytest.mark.parametrize("direction", [na.DIR_FWD, na.DIR_BWD])
def test_gpu_pong_generated(direction):
    # range queries of the workload-3 generator (2k stops): pong on the plain
    # engine equals the oracle's pong, work counters included
    tt, img = image_for("w3p", lambda: na.Timetable.generate(3, scale=0.01))
    eng = na.Engine(img, lanes=1, groups=2, max_k=10)
    qs = na.Queries.generate(tt, 3, 48, seed=0x7172 + 61)
    idx = [i for i in range(len(qs)) if qs[i].is_interval][:10]
    n = 0
    for i in idx:
        qs[i].dir = direction
        say(f"pong workload3 small dir={direction} query {i}")
        n += len(compare_pong(tt, eng, qs[i], check_work=True)["journeys"])
    assert n > 0


def test_gpu_pong_needs_the_plain_engine():
    tt, loc, day0 = abc_timetable()
    eng = na.Engine(na.Image(tt), lanes=64, groups=1, max_k=16)
    q = na.Queries.build([dict(start=[(loc["A"], 0)], dest=[(loc["C"], 0)],
                               interval=(day0 + hm(G.INTERVAL[0]), day0 + hm(G.INTERVAL[1])))])[0]
    with pytest.raises(na.NigiriError) as e:
        eng.pong(q)
    assert e.value.code == na.EUNSUPPORTED


def test_gpu_pong_on_concurrent_engines():
    # plain engines driven from several host threads at once (each with its
    # own state and stream) give each query the oracle's pong result
    import threading
    tt, img = image_for("w3p", lambda: na.Timetable.generate(3, scale=0.01))
    qs = na.Queries.generate(tt, 3, 96, seed=0x7172 + 71)
    idx = [i for i in range(len(qs)) if qs[i].is_interval][:12]
    engines = [na.Engine(img, lanes=1, groups=2, max_k=10) for _ in range(4)]
    got = {}
    errors = []

    def work(e, part):
        try:
            for i in part:
                got[i] = e.pong(qs[i])
        except Exception as ex:  # reported below
            errors.append(ex)

    say(f"pong on 4 engines from 4 threads, {len(idx)} queries")
    ths = [threading.Thread(target=work, args=(e, idx[k::4])) for k, e in enumerate(engines)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    assert not errors, errors
    for i in idx:
        ref = oracle.pong(tt, qs[i])
        assert got[i]["journeys"] == ref["journeys"]
        assert got[i]["legs"] == ref["legs"]
