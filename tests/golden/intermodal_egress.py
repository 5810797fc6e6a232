"""test/routing/gpu_intermodal_egress_test.cc:21-119 as data.

GTFS timetable (times UTC, service MON = 2024-06-17 only, date range
[2024-06-16, 2024-06-19)):
  T1: A 10:00 -> B 10:30
  T2: B 10:40 -> C 11:00
  T4: A2 09:50 -> C 11:00
Query: forward, intermodal start {A 5 min, A2 5 min}, intermodal destination
{C 0 min}, start interval [09:45, 10:45], min_connection_count 1, no
start footpaths.  The journey via A and B (leaving the start at 09:55,
reaching END at 11:00 with 1 transfer) must be found.
"""

DATE = (2024, 6, 17)
RANGE_FROM = (2024, 6, 16)
RANGE_DAYS = 3
TRIPS = {
    "T1": (["A", "B"], [600, 630]),
    "T2": (["B", "C"], [640, 660]),
    "T4": (["A2", "C"], [590, 660]),
}
START = [("A", 5), ("A2", 5)]
DEST = [("C", 0)]
INTERVAL = ("09:45", "10:45")
MIN_CONNECTIONS = 1
MUST_FIND = ("09:55", "11:00", 1)
