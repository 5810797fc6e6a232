"""Pins the CPU oracle with golden journeys of nigiri's own tests."""
import pytest

import nigiri_amd as na
import oracle
from golden import raptor_abc as G
from golden import start_times_simple as ST
from ttbuilder import Builder, day_number


def hm(s):
    h, m = s.split(":")
    return int(h) * 60 + int(m)


def abc_timetable():
    b = Builder(day_number(*G.RANGE_FROM), G.RANGE_DAYS, 3)
    loc = {"A": 9, "B": 10, "C": 11}
    day = day_number(*G.DATE)
    for name, (fr, to, first, travel, n, hw) in G.TRAINS.items():
        trips = []
        for i in range(n):
            dep = first + i * hw
            trips.append(([dep, dep + travel], [day]))
        b.add_line([loc[fr], loc[to]], trips)
    return b.build(), loc, day * 1440


def run(tt, q):
    return oracle.search(tt, q)["journeys"]


@pytest.mark.parametrize("direction", ["fwd", "bwd"])
def test_raptor_abc_range(direction):
    tt, loc, day0 = abc_timetable()
    iv = (day0 + hm(G.INTERVAL[0]), day0 + hm(G.INTERVAL[1]))
    if direction == "fwd":
        spec = dict(start=[(loc["A"], 0)], dest=[(loc["C"], 0)], interval=iv, dir=na.DIR_FWD)
        expected = G.FWD_EXPECTED
    else:
        spec = dict(start=[(loc["C"], 0)], dest=[(loc["A"], 0)], interval=iv, dir=na.DIR_BWD)
        expected = G.BWD_EXPECTED
    qs = na.Queries.build([spec])
    js = run(tt, qs[0])
    got = []
    for start, dest, _, transfers in js:
        dep, arr = min(start, dest), max(start, dest)
        got.append(((dep - day0) // 60, (dep - day0) % 60, (arr - day0) // 60, (arr - day0) % 60,
                    transfers))
    want = [(hm(d) // 60, hm(d) % 60, hm(a) // 60, hm(a) % 60, t) for d, a, t in expected]
    assert got == want


def test_raptor_abc_point_query_closed_form():
    # ontrip at 05:10: next 1337 at 05:30 -> B 06:30 (+2 transfer) -> 7331 06:45 -> C 07:45
    tt, loc, day0 = abc_timetable()
    qs = na.Queries.build([dict(start=[(loc["A"], 0)], dest=[(loc["C"], 0)], time=day0 + hm("05:10"))])
    js = run(tt, qs[0])
    assert js == [(day0 + hm("05:10"), day0 + hm("07:45"), loc["C"], 1)]


def test_transfer_time_is_respected():
    # arrival at B 06:00; with a 16 minute transfer time the 06:15 connection is missed
    b = Builder(day_number(*G.RANGE_FROM), G.RANGE_DAYS, 3)
    day = day_number(*G.DATE)
    b.transfer[10] = 16
    b.add_line([9, 10], [([300, 360], [day])])
    b.add_line([10, 11], [([375, 435], [day]), ([390, 450], [day])])
    tt = b.build()
    qs = na.Queries.build([dict(start=[(9, 0)], dest=[(11, 0)], time=day * 1440 + 290)])
    assert run(tt, qs[0]) == [(day * 1440 + 290, day * 1440 + 450, 11, 1)]


def test_start_dest_overlap_is_empty():
    tt, loc, day0 = abc_timetable()
    qs = na.Queries.build([dict(start=[(loc["A"], 0)], dest=[(loc["A"], 0)], time=day0 + 300)])
    assert run(tt, qs[0]) == []


def line_timetable(spec):
    b = Builder(day_number(*spec.RANGE_FROM), spec.RANGE_DAYS, 3)
    names = sorted({n for fr, to, *_ in spec.TRAINS.values() for n in (fr, to)})
    loc = {n: 9 + i for i, n in enumerate(names)}
    day = day_number(*spec.DATE)
    for name, (fr, to, first, travel, n, hw) in spec.TRAINS.items():
        b.add_line([loc[fr], loc[to]], [([first + i * hw, first + i * hw + travel], [day])
                                        for i in range(n)])
    return b.build(), loc, day * 1440


def test_start_times_golden_labels():
    # get_starts (start_times.cc) against start_times_test.cc's printed list
    tt, loc, day0 = line_timetable(ST)
    mar = day0 - 30 * 1440  # minutes of 2020-03-00 00:00 UTC

    def t(s):
        d, hhmm = s.split(" ")
        return mar + int(d) * 1440 + hm(hhmm)

    qs = na.Queries.build([dict(start=[(loc[s], d) for s, d in ST.OFFSETS], dest=[(loc["A"], 0)],
                                interval=(t(ST.INTERVAL[0]), t(ST.INTERVAL[1])))])
    got = oracle.starts(tt, qs[0])
    uniq = [x for i, x in enumerate(got) if i == 0 or got[i - 1] != x]
    name = {v: k for k, v in loc.items()}
    assert [(a, b, name[s]) for a, b, s in uniq] == [(t(a), t(b), s) for a, b, s in ST.EXPECTED]


def test_raptor_intermodal_forward_golden():
    tt, loc, day0 = abc_timetable()
    iv = (day0 + hm(G.INTERVAL[0]), day0 + hm(G.INTERVAL[1]))
    s, sd = G.INTERMODAL_START
    d, dd = G.INTERMODAL_DEST
    qs = na.Queries.build([dict(start=[(loc[s], sd)], dest=[(loc[d], dd)], interval=iv,
                                start_match_mode=na.MATCH_INTERMODAL,
                                dest_match_mode=na.MATCH_INTERMODAL)])
    js = run(tt, qs[0])
    want = [(day0 + hm(a), day0 + hm(b), 1, k) for a, b, k in G.INTERMODAL_FWD_EXPECTED]
    assert sorted(js) == sorted(want)


def expected_legs(spec_legs, loc, day0, traffic_day):
    name = {"START": 0, "END": 1, **loc}
    out = []
    for leg in spec_legs:
        kind, fr, to, dep, arr = leg[:5]
        row = [name[fr], name[to], day0 + hm(dep), day0 + hm(arr)]
        if kind == "run":
            # the line is fixed by the stops (A -> B only on 1337, B -> C only on 7331)
            out.append((na.LEG_RUN, *row, (None, None, traffic_day, 0, 2)))
        elif kind == "footpath":
            out.append((na.LEG_FOOTPATH, *row, (name[fr], leg[5])))
        else:
            out.append((na.LEG_OFFSET, *row, (name[leg[5]], leg[6])))
    return out


def comparable(legs):
    # runs: compare traffic day and stop range; with the stops and times they fix the trip
    return [l if l[0] != na.LEG_RUN else (*l[:5], l[5][2:]) for l in legs]


@pytest.mark.parametrize("direction", ["fwd", "bwd", "intermodal"])
def test_reconstructed_legs_golden(direction):
    # legs printed by raptor_test.cc:18-112 and raptor_intermodal_test.cc:16-50
    tt, loc, day0 = abc_timetable()
    traffic_day = (day0 - tt.view.internal_from) // 1440
    iv = (day0 + hm(G.INTERVAL[0]), day0 + hm(G.INTERVAL[1]))
    if direction == "fwd":
        spec = dict(start=[(loc["A"], 0)], dest=[(loc["C"], 0)], interval=iv, dir=na.DIR_FWD)
        want = G.FWD_LEGS
    elif direction == "bwd":
        spec = dict(start=[(loc["C"], 0)], dest=[(loc["A"], 0)], interval=iv, dir=na.DIR_BWD)
        want = G.BWD_LEGS
    else:
        s, sd = G.INTERMODAL_START
        d, dd = G.INTERMODAL_DEST
        spec = dict(start=[(loc[s], sd)], dest=[(loc[d], dd)], interval=iv,
                    start_match_mode=na.MATCH_INTERMODAL, dest_match_mode=na.MATCH_INTERMODAL)
        want = G.INTERMODAL_FWD_LEGS
    out = oracle.search(tt, na.Queries.build([spec])[0])
    order = sorted(range(len(out["journeys"])), key=lambda i: min(out["journeys"][i][:2]))
    got = [comparable(out["legs"][i]) for i in order]
    assert got == [comparable(expected_legs(w, loc, day0, traffic_day)) for w in want]


def intermodal_bwd_query(loc, day0):
    iv = (day0 + hm(G.INTERVAL[0]), day0 + hm(G.INTERVAL[1]))
    (s, sd), (d, dd) = G.INTERMODAL_BWD_START, G.INTERMODAL_BWD_DEST
    return na.Queries.build([dict(start=[(loc[s], sd)], dest=[(loc[d], dd)], interval=iv,
                                  dir=na.DIR_BWD, start_match_mode=na.MATCH_INTERMODAL,
                                  dest_match_mode=na.MATCH_INTERMODAL,
                                  min_connection_count=G.INTERMODAL_BWD_MIN_CONNECTIONS,
                                  extend_interval_earlier=1, extend_interval_later=1)])


def test_raptor_intermodal_backward_golden():
    # raptor_intermodal_test.cc:85-251: the interval extension (search.h:237-317,
    # interval_estimate.h) widens [05:00, 06:00) until 3 journeys start in it
    tt, loc, day0 = abc_timetable()
    traffic_day = (day0 - tt.view.internal_from) // 1440
    out = oracle.search(tt, intermodal_bwd_query(loc, day0)[0])
    got = sorted(zip(out["journeys"], out["legs"]), key=lambda x: x[0][1])
    want = [(day0 + hm(a), day0 + hm(d), 1, k) for d, a, k in G.INTERMODAL_BWD_EXPECTED]
    assert [j for j, _ in got] == want
    assert [comparable(l) for _, l in got] == [
        comparable(expected_legs(w, loc, day0, traffic_day)) for w in G.INTERMODAL_BWD_LEGS]


def egress_timetable():
    from golden import intermodal_egress as E
    b = Builder(day_number(*E.RANGE_FROM), E.RANGE_DAYS, 4)
    loc = {"A": 9, "A2": 10, "B": 11, "C": 12}
    day = day_number(*E.DATE)
    for stops, times in E.TRIPS.values():
        b.add_line([loc[x] for x in stops], [(times, [day])])
    return b.build(), loc, day * 1440, E


def test_intermodal_egress_at_bound_golden():
    # gpu_intermodal_egress_test.cc:89-119: the one-transfer journey via A and B
    # leaves the start at 09:55 and reaches END at 11:00
    tt, loc, day0, E = egress_timetable()
    q = na.Queries.build([dict(start=[(loc[s], d) for s, d in E.START],
                               dest=[(loc[s], d) for s, d in E.DEST],
                               interval=(day0 + hm(E.INTERVAL[0]), day0 + hm(E.INTERVAL[1])),
                               start_match_mode=na.MATCH_INTERMODAL,
                               dest_match_mode=na.MATCH_INTERMODAL,
                               min_connection_count=E.MIN_CONNECTIONS)])
    js = oracle.search(tt, q[0])["journeys"]
    dep, arr, k = E.MUST_FIND
    assert (day0 + hm(dep), day0 + hm(arr), 1, k) in js
