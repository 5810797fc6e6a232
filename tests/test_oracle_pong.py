"""Pins the oracle's pong range search (raptor/pong.cc:62-428): golden
journeys of nigiri's pong tests, and the cross-check nigiri runs on every
range query of its routing tests (test/raptor_search.cc:204-232): the
journeys pong finds in the query interval equal those of the range search."""
import pytest

import nigiri_amd as na
import oracle
from golden import pong as P
from golden import raptor_abc as G
from test_oracle_golden import abc_timetable, egress_timetable, hm, intermodal_bwd_query
from ttbuilder import Builder, day_number


def line_tt(spec):
    b = Builder(day_number(*spec["RANGE_FROM"]), spec["RANGE_DAYS"], len(spec["STOPS"]))
    loc = {name: 9 + i for i, name in enumerate(spec["STOPS"])}
    day = day_number(*spec["DATE"])
    for stops, times in spec["TRIPS"].values():
        b.add_line([loc[s] for s in stops], [(times, [day])])
    return b.build(), loc, day * 1440


def printed(leg):
    """What nigiri prints of a leg: a footpath shows its duration, not the
    footpath object's target (the source of an incoming footpath when a
    backward search reconstructs it)."""
    if leg[0] == na.LEG_FOOTPATH:
        return leg[:5] + (leg[5][1],)
    return leg


def in_window(q, out):
    """(departure, arrival, transfers) of the journeys whose anchor (departure
    forward, arrival backward) lies in the query interval, with their legs.
    The reference compares printed journeys, which show legs, not dest_
    (pong journeys keep the origin as dest_)."""
    lo, hi = q.interval_from, q.interval_to
    res = []
    for j, legs in zip(out["journeys"], out["legs"]):
        dep, arr = min(j[:2]), max(j[:2])
        anchor = dep if q.dir == na.DIR_FWD else arr
        if lo <= anchor < hi:
            res.append(((dep, arr, j[3]), [printed(l) for l in legs]))
    return sorted(res)


def assert_pong_equals_range(tt, q, legs=True):
    """legs=False on generated networks: there, equally good itineraries
    exist, and the backward reconstruction of a pong journey may pick
    another one than the forward reconstruction (the times agree)."""
    a = in_window(q, oracle.search(tt, q))
    b = in_window(q, oracle.pong(tt, q))
    assert [j for j, _ in a] == [j for j, _ in b]
    for (dep, arr, _), ls in b:
        assert ls[0][3] == dep and ls[-1][4] == arr
        assert all(x[4] <= y[3] for x, y in zip(ls, ls[1:]))  # legs in travel order
    if legs:
        assert [l for _, l in a] == [l for _, l in b]
    return b


@pytest.mark.parametrize("case", range(len(P.MINLOOKAHEAD["CASES"])))
def test_pong_min_lookahead_golden(case):
    tt, loc, day0 = line_tt(P.MINLOOKAHEAD)
    when, max_tt, want = P.MINLOOKAHEAD["CASES"][case]
    spec = dict(start=[(loc["S0"], 0)], dest=[(loc["S2"], 0)], max_travel_time=max_tt,
                min_connection_count=1, extend_interval_later=1)
    if isinstance(when, tuple):
        spec["interval"] = (day0 + hm(when[0]), day0 + hm(when[1]))
    else:
        spec["time"] = day0 + hm(when)
    out = oracle.pong(tt, na.Queries.build([spec])[0])
    # a pong journey keeps the backward search's destination, the origin
    # (pong.cc:403-405 swaps only the times)
    assert out["journeys"] == [(day0 + hm(d), day0 + hm(a), loc["S0"], k) for d, a, k in want]
    if want:
        kinds = {"run": na.LEG_RUN, "footpath": na.LEG_FOOTPATH}
        got = [(l[0], l[1], l[2], l[3], l[4]) for l in out["legs"][0]]
        assert got == [(kinds[k], loc[f], loc[t], day0 + hm(d), day0 + hm(a))
                       for k, f, t, d, a in P.MINLOOKAHEAD["LEGS"]]


def test_pong_gap_truncation_golden():
    tt, loc, day0 = line_tt(P.GAP)
    iv = (day0 + hm(P.GAP["INTERVAL"][0]), day0 + hm(P.GAP["INTERVAL"][1]))
    q = na.Queries.build([dict(start=[(loc["A"], 0)], dest=[(loc["B"], 0)], interval=iv,
                               min_connection_count=P.GAP["MIN_CONNECTIONS"],
                               extend_interval_later=1)])[0]
    deps = sorted(j[0] for j in oracle.pong(tt, q)["journeys"])
    assert deps == [day0 + hm(t) for t in P.GAP["EXPECTED_DEPARTURES"]]
    # the range search gives the same departures (the test asserts them on it)
    assert sorted(j[0] for j in oracle.search(tt, q)["journeys"]) == deps


@pytest.mark.parametrize("direction", ["fwd", "bwd"])
def test_pong_equals_range_abc(direction):
    tt, loc, day0 = abc_timetable()
    iv = (day0 + hm(G.INTERVAL[0]), day0 + hm(G.INTERVAL[1]))
    if direction == "fwd":
        spec = dict(start=[(loc["A"], 0)], dest=[(loc["C"], 0)], interval=iv, dir=na.DIR_FWD)
    else:
        spec = dict(start=[(loc["C"], 0)], dest=[(loc["A"], 0)], interval=iv, dir=na.DIR_BWD)
    got = assert_pong_equals_range(tt, na.Queries.build([spec])[0])
    assert len(got) == len(G.FWD_EXPECTED if direction == "fwd" else G.BWD_EXPECTED)


def test_pong_equals_range_intermodal():
    tt, loc, day0 = abc_timetable()
    iv = (day0 + hm(G.INTERVAL[0]), day0 + hm(G.INTERVAL[1]))
    (s, sd), (d, dd) = G.INTERMODAL_START, G.INTERMODAL_DEST
    q = na.Queries.build([dict(start=[(loc[s], sd)], dest=[(loc[d], dd)], interval=iv,
                               start_match_mode=na.MATCH_INTERMODAL,
                               dest_match_mode=na.MATCH_INTERMODAL)])[0]
    assert len(assert_pong_equals_range(tt, q)) == len(G.INTERMODAL_FWD_EXPECTED)
    assert_pong_equals_range(tt, intermodal_bwd_query(loc, day0)[0])
    tt, loc, day0, E = egress_timetable()
    q = na.Queries.build([dict(start=[(loc[s], d) for s, d in E.START],
                               dest=[(loc[s], d) for s, d in E.DEST],
                               interval=(day0 + hm(E.INTERVAL[0]), day0 + hm(E.INTERVAL[1])),
                               start_match_mode=na.MATCH_INTERMODAL,
                               dest_match_mode=na.MATCH_INTERMODAL,
                               min_connection_count=E.MIN_CONNECTIONS)])[0]
    got = assert_pong_equals_range(tt, q)
    dep, arr, k = E.MUST_FIND
    assert (day0 + hm(dep), day0 + hm(arr), k) in [j for j, _ in got]


def range_queries(tt, cfg, n, seed):
    qs = na.Queries.generate(tt, cfg, 4 * n, seed=seed)
    return [i for i in range(len(qs)) if qs[i].is_interval][:n], qs


@pytest.mark.parametrize("direction", [na.DIR_FWD, na.DIR_BWD])
def test_pong_equals_range_generated(direction):
    # range queries of the 200k-stop workload's generator, scaled down
    tt = na.Timetable.generate(3, scale=0.01)
    idx, qs = range_queries(tt, 3, 12, 0x7172 + 61)
    assert len(idx) >= 8
    n_journeys = 0
    for i in idx:
        q = qs[i]
        q.dir = direction
        n_journeys += len(assert_pong_equals_range(tt, q, legs=False))
    assert n_journeys > 0
