"""Pins the oracle's journey reconstruction (src/routing/raptor/reconstruct.cc and
optimize_footpaths.cc) beyond the golden legs of test_oracle_golden.py:

* every reconstructed journey is a feasible itinerary of the timetable: its runs
  are active trips whose times are the timetable's events, consecutive legs
  meet in place and time, walks are footpaths or transfers of their own
  duration, and it has transfers + 1 runs;
* optimize_footpaths on the journeys of test/routing/optmize_footpaths_test.cc
  with the via stops removed; the expected legs follow by hand from the rules of
  optimize_initial_departure, optimize_last_arrival and optimize_transfers
  (the reference test keeps the vias, whose only effect is to block these moves).
"""
import random

import numpy as np
import pytest

import nigiri_amd as na
import oracle
from test_oracle_bruteforce import FIRST, random_network
from ttbuilder import Builder, day_number


class TtArrays:
    def __init__(self, tt):
        for n in ("route_stop_offsets", "route_stops", "route_transport_from",
                  "route_stop_time_from", "event_times", "transport_traffic_days", "bitfields",
                  "transfer_time", "fp_out_offsets", "fp_out"):
            setattr(self, n, tt.array(n).astype(np.int64))
        self.internal_from = tt.view.internal_from

    def stop_loc(self, r, i):
        return int(self.route_stops[self.route_stop_offsets[r] + i]) & ((1 << 28) - 1)

    def event(self, r, t, day, i, arr):
        n_t = self.route_transport_from[r + 1] - self.route_transport_from[r]
        d = int(self.event_times[self.route_stop_time_from[r] + n_t * (2 * i - arr)
                                 + t - self.route_transport_from[r]])
        return self.internal_from + day * 1440 + (d & 31) * 1440 + (d >> 5)

    def active(self, t, day):
        bf = int(self.transport_traffic_days[t])
        return day < 512 and (int(self.bitfields[bf * 8 + day // 64]) >> (day % 64)) & 1 == 1

    def footpath(self, a, b):
        for e in range(self.fp_out_offsets[a], self.fp_out_offsets[a + 1]):
            if int(self.fp_out[e]) & ((1 << 23) - 1) == b:
                return int(self.fp_out[e]) >> 23
        return None


def check_journey(ta, j, legs, starts, fwd):
    start, dest_time, dest, transfers = j
    assert legs, j
    runs = [l for l in legs if l[0] == na.LEG_RUN]
    assert len(runs) == transfers + 1
    for a, b in zip(legs, legs[1:]):
        assert a[2] == b[1] and a[4] <= b[3], (j, legs)
    for kind, fr, to, dep, arr, what in legs:
        assert dep <= arr
        if kind == na.LEG_RUN:
            t, r, day, s0, s1 = what
            assert ta.route_transport_from[r] <= t < ta.route_transport_from[r + 1]
            assert s0 < s1 - 1 and ta.active(t, day)
            assert ta.stop_loc(r, s0) == fr and ta.stop_loc(r, s1 - 1) == to
            assert ta.event(r, t, day, s0, 0) == dep and ta.event(r, t, day, s1 - 1, 1) == arr
        elif kind == na.LEG_FOOTPATH:
            if fr == to:
                assert arr - dep == int(ta.transfer_time[fr])
            else:
                assert ta.footpath(fr, to) == arr - dep
    first, last = legs[0], legs[-1]
    if fwd:
        assert first[1] in starts and last[2] == dest
        assert first[3] >= start and last[4] <= dest_time
    else:
        assert first[1] == dest and last[2] in starts
        assert first[3] >= dest_time and last[4] <= start


def check_network(seed):
    b, lines, fps = random_network(seed, n_stops=7, n_lines=5)
    tt = b.build()
    ta = TtArrays(tt)
    rnd = random.Random(77 + seed)
    n_checked = 0
    for _ in range(8):
        src, dst = rnd.sample(range(9, 16), 2)
        fwd = rnd.random() < 0.6
        spec = dict(start=[(src, 0)], dest=[(dst, 0)], dir=na.DIR_FWD if fwd else na.DIR_BWD,
                    use_start_footpaths=rnd.random() < 0.5, max_transfers=rnd.randint(0, 5))
        t = (FIRST + 1) * 1440 + rnd.randint(0, 1439)
        if rnd.random() < 0.5:
            spec["time"] = t
        else:
            spec["interval"] = (t, t + rnd.randint(30, 600))
        out = oracle.search(tt, na.Queries.build([spec])[0])
        for j, legs in zip(out["journeys"], out["legs"]):
            check_journey(ta, j, legs, {src}, fwd)
            n_checked += 1
    return n_checked


@pytest.mark.parametrize("seed", range(40))
def test_reconstructed_journeys_are_feasible_itineraries(seed):
    check_network(seed)


def test_feasibility_check_sees_journeys():
    assert sum(check_network(seed) for seed in range(10)) >= 20


def optimize_timetable(extra_fp=False):
    # optmize_footpaths_test.cc:49-87: T0 A 10:10 -> B 10:12 -> C 10:15,
    # T1 C 10:17 -> D 10:33 -> E 10:40 on 2019-05-01 (times taken as UTC here)
    first = day_number(2019, 5, 1)
    bld = Builder(first, 1, 5)
    A, B, C, D, E = range(9, 14)
    bld.add_line([A, B, C], [([610, 612, 612, 615], [first])])
    bld.add_line([C, D, E], [([617, 633, 633, 640], [first])])
    if extra_fp:
        bld.add_footpath(B, D, 1, both=False)
    return bld.build(), (A, B, C, D, E), first * 1440, bld.internal_day(first)


def leg(kind, fr, to, dep, arr, transport=0, route=0, day=0, stop_from=0, stop_to=0,
        duration=0, target=0):
    return na.Leg(kind, fr, to, dep, arr, transport, route, day, stop_from, stop_to, duration,
                  target)


def test_optimize_initial_departure_without_via():
    # start offset {B, 8}: boarding T0 at B (10:12) leaves the start at 10:04,
    # later than the 10:00 of the {A, 10} offset, so the offset moves to B
    tt, (A, B, C, D, E), d0, day = optimize_timetable()
    q = na.Queries.build([dict(start=[(B, 8)], dest=[(E, 0)], time=d0 + 600,
                               start_match_mode=na.MATCH_INTERMODAL)])[0]
    legs = [leg(na.LEG_OFFSET, 0, A, d0 + 600, d0 + 610, duration=10, target=A),
            leg(na.LEG_RUN, A, C, d0 + 610, d0 + 615, 0, 0, day, 0, 3)]
    got = oracle.optimize_footpaths(tt, q, legs)
    assert got == [(na.LEG_OFFSET, 0, B, d0 + 604, d0 + 612, (B, 8)),
                   (na.LEG_RUN, B, C, d0 + 612, d0 + 615, (0, 0, day, 1, 3))]


def test_optimize_last_arrival_without_via():
    # destination offset {B, 8}: leaving T0 at B (10:12) reaches END at 10:20,
    # before the 10:25 of the {C, 10} egress, so the egress moves to B
    tt, (A, B, C, D, E), d0, day = optimize_timetable()
    q = na.Queries.build([dict(start=[(A, 0)], dest=[(B, 8)], time=d0 + 610,
                               dest_match_mode=na.MATCH_INTERMODAL)])[0]
    legs = [leg(na.LEG_RUN, A, C, d0 + 610, d0 + 615, 0, 0, day, 0, 3),
            leg(na.LEG_OFFSET, C, 1, d0 + 615, d0 + 625, duration=10, target=1)]
    got = oracle.optimize_footpaths(tt, q, legs)
    assert got == [(na.LEG_RUN, A, B, d0 + 610, d0 + 612, (0, 0, day, 0, 2)),
                   (na.LEG_OFFSET, B, 1, d0 + 612, d0 + 620, (B, 8))]


def test_optimize_transfers_without_via():
    # transfer C -> C (2 min, no buffer): penalty 2.  The footpath B -> D (1 min)
    # reaches T1 at D with 20 min to spare: penalty 1 - 0.2 * 15 = -2, so the
    # transfer moves to B -> D; no other stop pair has a footpath
    tt, (A, B, C, D, E), d0, day = optimize_timetable(extra_fp=True)
    q = na.Queries.build([dict(start=[(A, 0)], dest=[(E, 0)], time=d0 + 610)])[0]
    legs = [leg(na.LEG_RUN, A, C, d0 + 610, d0 + 615, 0, 0, day, 0, 3),
            leg(na.LEG_FOOTPATH, C, C, d0 + 615, d0 + 617, duration=2, target=C),
            leg(na.LEG_RUN, C, E, d0 + 617, d0 + 640, 1, 1, day, 0, 3)]
    got = oracle.optimize_footpaths(tt, q, legs)
    assert got == [(na.LEG_RUN, A, B, d0 + 610, d0 + 612, (0, 0, day, 0, 2)),
                   (na.LEG_FOOTPATH, B, D, d0 + 612, d0 + 613, (D, 1)),
                   (na.LEG_RUN, D, E, d0 + 633, d0 + 640, (1, 1, day, 1, 3))]


def test_optimize_keeps_a_transfer_without_better_footpath():
    # without the B -> D footpath nothing improves on the transfer at C
    tt, (A, B, C, D, E), d0, day = optimize_timetable()
    q = na.Queries.build([dict(start=[(A, 0)], dest=[(E, 0)], time=d0 + 610)])[0]
    legs = [leg(na.LEG_RUN, A, C, d0 + 610, d0 + 615, 0, 0, day, 0, 3),
            leg(na.LEG_FOOTPATH, C, C, d0 + 615, d0 + 617, duration=2, target=C),
            leg(na.LEG_RUN, C, E, d0 + 617, d0 + 640, 1, 1, day, 0, 3)]
    got = oracle.optimize_footpaths(tt, q, legs)
    assert got == [l.as_tuple() for l in legs]
