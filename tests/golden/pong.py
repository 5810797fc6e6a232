"""Golden vectors of nigiri's pong tests, as data.

MINLOOKAHEAD: test/routing/pong_minlookahead_test.cc:21-178.  GTFS
timetable in Europe/London (UTC+1 in June), service MON 2026-06-01 only,
date range [2026-06-01, 2026-06-07).  In UTC:
  R0: S0 04:00 -> S1 05:00
  R1: S1 07:00 -> S2 09:00
Transfer time 2 min at every stop (the GTFS default).  Queries: forward
S0 -> S2, no start footpaths, min_connection_count 1, extend later.
  * interval [00:00, 04:01) and [00:00, 04:00), max travel 5 h: one journey
    04:00 -> 09:00 with 1 transfer (pong_ignores_interval);
  * ontrip at 00:00, max travel 5 h: the same journey; max travel 4 h 59:
    none (pong_min_lookahead).
Expected legs (test file lines 63-77): R0 S0 04:00 -> S1 05:00, footpath
S1 -> S1 05:00 -> 05:02 (2 min), R1 S1 07:00 -> S2 09:00.

GAP: test/routing/pong_gap_truncation_test.cc:21-123.  Etc/UTC, service
2024-06-19 only, date range [2024-06-18, 2024-06-20):
  PH: A 09:59 -> B 11:00
  EA: A 10:05 -> C 10:20,  EB: C 10:30 -> B 11:00
  L:  A 15:00 -> B 15:30
Query: forward A -> B, interval [10:00, 10:03), no start footpaths,
min_connection_count 2, extend later.  Expected departures: 10:05, 15:00
(no journey departs before 10:00).
"""

MINLOOKAHEAD = dict(
    DATE=(2026, 6, 1),
    RANGE_FROM=(2026, 6, 1),
    RANGE_DAYS=6,
    STOPS=["S0", "S1", "S2"],
    TRIPS={"R0": (["S0", "S1"], [240, 300]), "R1": (["S1", "S2"], [420, 540])},
    # (interval or ontrip time, max_travel_time, expected journeys (dep, arr, transfers))
    CASES=[
        (("00:00", "04:01"), 300, [("04:00", "09:00", 1)]),
        (("00:00", "04:00"), 300, [("04:00", "09:00", 1)]),
        ("00:00", 300, [("04:00", "09:00", 1)]),
        ("00:00", 299, []),
    ],
    # (kind, from, to, dep, arr)
    LEGS=[("run", "S0", "S1", "04:00", "05:00"), ("footpath", "S1", "S1", "05:00", "05:02"),
          ("run", "S1", "S2", "07:00", "09:00")],
)

GAP = dict(
    DATE=(2024, 6, 19),
    RANGE_FROM=(2024, 6, 18),
    RANGE_DAYS=2,
    STOPS=["A", "C", "B"],
    TRIPS={"PH": (["A", "B"], [599, 660]), "EA": (["A", "C"], [605, 620]),
           "EB": (["C", "B"], [630, 660]), "L": (["A", "B"], [900, 930])},
    INTERVAL=("10:00", "10:03"),
    MIN_CONNECTIONS=2,
    EXPECTED_DEPARTURES=["10:05", "15:00"],
)
