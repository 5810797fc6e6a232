"""Golden journeys of nigiri's test/routing/raptor_test.cc (forward :18-66,
backward :68-112) on the HRD timetable files_abc()
(test/loader/hrd/hrd_timetable.h:164-178), transcribed as data.

Timetable (HRD times are local, +02:00 from 2020-03-29 02:00, so UTC is
local - 2 h on 2020-03-30; bitfield 000005 = 30.03 only; the HRD count is
the number of repetitions after the first trip):
  RE 1337  A -> B : 1 + 48 trips every 30 min, first A dep 02:30 local = 00:30 UTC,
                    B arr 03:30 local = 01:30 UTC
  RE 7331  B -> C : 1 + 92 trips every 15 min, first B dep 02:30 local = 00:30 UTC,
                    C arr 03:30 local = 01:30 UTC
  transfer time 2 min (printed as "FOOTPATH (duration=2)" at B); no footpaths
  between A, B and C.
  date range [2020-03-28, 2020-03-31) (hrd_timetable.h:18-23).

Queries: A -> C forward and C -> A backward, interval
[2020-03-30 05:00, 2020-03-30 06:00) UTC, query defaults otherwise
(test/raptor_search.cc:87-99).
Expected journeys as (departure, arrival, transfers) in UTC "HH:MM" of
2020-03-30, in the printed result order.
"""

DATE = (2020, 3, 30)
RANGE_FROM = (2020, 3, 28)
RANGE_DAYS = 3
TRAINS = {
    # name: (from, to, first dep UTC minutes, travel minutes, n trips, headway)
    "1337": ("A", "B", 30, 60, 49, 30),
    "7331": ("B", "C", 30, 60, 93, 15),
}
INTERVAL = ("05:00", "06:00")
FWD_EXPECTED = [("05:00", "07:15", 1), ("05:30", "07:45", 1)]
BWD_EXPECTED = [("03:00", "05:15", 1), ("03:30", "05:45", 1)]

# test/routing/raptor_intermodal_test.cc:16-83 (raptor_intermodal_forward) on
# the same timetable: start offset {A, 10 min}, destination offset {C, 15 min},
# interval [05:00, 06:00) UTC, forward.  Journeys START -> END as
# (departure, arrival, transfers).
INTERMODAL_START = ("A", 10)
INTERMODAL_DEST = ("C", 15)
INTERMODAL_FWD_EXPECTED = [("05:20", "08:00", 1), ("05:50", "08:30", 1)]

# The legs of the journeys above as nigiri's tests print them (the
# reconstruction, src/routing/raptor/reconstruct.cc, then optimize_footpaths):
# ("run", from, to, dep, arr, train) with the stop range [0, 2) of the
# two-stop line, ("footpath", from, to, dep, arr, duration),
# ("offset", from, to, dep, arr, target, duration); times "HH:MM" UTC.
FWD_LEGS = [
    [("run", "A", "B", "05:00", "06:00", "1337"), ("footpath", "B", "B", "06:00", "06:02", 2),
     ("run", "B", "C", "06:15", "07:15", "7331")],
    [("run", "A", "B", "05:30", "06:30", "1337"), ("footpath", "B", "B", "06:30", "06:32", 2),
     ("run", "B", "C", "06:45", "07:45", "7331")],
]
BWD_LEGS = [
    [("run", "A", "B", "03:00", "04:00", "1337"), ("footpath", "B", "B", "04:00", "04:02", 2),
     ("run", "B", "C", "04:15", "05:15", "7331")],
    [("run", "A", "B", "03:30", "04:30", "1337"), ("footpath", "B", "B", "04:30", "04:32", 2),
     ("run", "B", "C", "04:45", "05:45", "7331")],
]
INTERMODAL_FWD_LEGS = [
    [("offset", "START", "A", "05:20", "05:30", "A", 10),
     ("run", "A", "B", "05:30", "06:30", "1337"), ("footpath", "B", "B", "06:30", "06:32", 2),
     ("run", "B", "C", "06:45", "07:45", "7331"), ("offset", "C", "END", "07:45", "08:00", "C", 15)],
    [("offset", "START", "A", "05:50", "06:00", "A", 10),
     ("run", "A", "B", "06:00", "07:00", "1337"), ("footpath", "B", "B", "07:00", "07:02", 2),
     ("run", "B", "C", "07:15", "08:15", "7331"), ("offset", "C", "END", "08:15", "08:30", "C", 15)],
]

# test/routing/raptor_intermodal_test.cc:85-251 (raptor_intermodal_backward)
# on the same timetable: start offset {C, 15 min}, destination offset
# {A, 10 min}, interval [05:00, 06:00) UTC, backward, min_connection_count 3,
# extending the interval earlier and later.  The printed journeys run from
# END (departure) to START (arrival), every 30 minutes from 00:20 to 03:50:
# (departure at END, arrival at START, transfers).
INTERMODAL_BWD_START = ("C", 15)
INTERMODAL_BWD_DEST = ("A", 10)
INTERMODAL_BWD_MIN_CONNECTIONS = 3
INTERMODAL_BWD_EXPECTED = [
    ("00:20", "03:00", 1), ("00:50", "03:30", 1), ("01:20", "04:00", 1), ("01:50", "04:30", 1),
    ("02:20", "05:00", 1), ("02:50", "05:30", 1), ("03:20", "06:00", 1), ("03:50", "06:30", 1),
]


def _hm(m):
    return f"{m // 60:02d}:{m % 60:02d}"


# their legs: END -> A offset (10), A -> B (1337), transfer at B (2),
# B -> C (7331), C -> START offset (15)
INTERMODAL_BWD_LEGS = [
    [("offset", "END", "A", _hm(d), _hm(d + 10), "A", 10),
     ("run", "A", "B", _hm(d + 10), _hm(d + 70), "1337"),
     ("footpath", "B", "B", _hm(d + 70), _hm(d + 72), 2),
     ("run", "B", "C", _hm(d + 85), _hm(d + 145), "7331"),
     ("offset", "C", "START", _hm(d + 145), _hm(d + 160), "C", 15)]
    for d in range(20, 20 + 8 * 30, 30)
]
