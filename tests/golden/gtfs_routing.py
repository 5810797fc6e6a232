"""GTFS feeds of nigiri's routing tests and the journeys those tests print,
transcribed as data.

TRANSFER_TIME_FEED: test/routing/transfer_time_settings_test.cc:33-74
(date range [2019-05-01, 2019-05-02), Europe/Berlin = UTC+2 that day).
  T1: A 10:00 -> B 10:10, T2: B 10:13 -> C 10:22, T3: B 10:20 -> C 10:30,
  T4: A 09:55 -> C 10:35 (local); transfer time 2 min at every stop.
Queries A -> C (forward) and C -> A (backward) over the whole date range as
the interval, use_start_footpaths false (query default), with the
transfer_time_settings and limits of :184-340.  Expected journeys are the
printed ones (:76-137), as legs in travel order with UTC times on
2019-05-01: ("run", from, to, dep, arr, trip_id) and
("footpath", from, to, dep, arr, duration).

MAX_INTERVAL_FEED: test/routing/max_interval_test.cc:19-55 (date range
[2019-03-25, 2019-11-01)); query :72-91, expected journeys :57-70.
"""

TRANSFER_TIME_FEED = r"""
# agency.txt
agency_id,agency_name,agency_url,agency_timezone
DB,Deutsche Bahn,https://deutschebahn.com,Europe/Berlin

# stops.txt
stop_id,stop_name,stop_desc,stop_lat,stop_lon,stop_url,location_type,parent_station
A,A,,0.0,1.0,,
B,B,,2.0,3.0,,
C,C,,4.0,5.0,,

# routes.txt
route_id,agency_id,route_short_name,route_long_name,route_desc,route_type
R1,DB,1,,,3
R2,DB,2,,,3
R3,DB,3,,,3

# trips.txt
route_id,service_id,trip_id,trip_headsign,block_id
R1,S1,T1,,
R2,S1,T2,,
R2,S1,T3,,
R3,S1,T4,,

# stop_times.txt
trip_id,arrival_time,departure_time,stop_id,stop_sequence
T1,10:00:00,10:00:00,A,0
T1,10:10:00,10:10:00,B,1
T2,10:13:00,10:13:00,B,0
T2,10:22:00,10:22:00,C,1
T3,10:20:00,10:20:00,B,0
T3,10:30:00,10:30:00,C,1
T4,09:55:00,09:55:00,A,0
T4,10:35:00,10:35:00,C,1

# calendar_dates.txt
service_id,date,exception_type
S1,20190501,1
"""
TRANSFER_TIME_RANGE = ((2019, 5, 1), (2019, 5, 2))
DATE = (2019, 5, 1)

DEFAULT = [("run", "A", "B", "08:00", "08:10", "T1"), ("footpath", "B", "B", "08:10", "08:12", 2),
           ("run", "B", "C", "08:13", "08:22", "T2")]
DUR10 = [("run", "A", "B", "08:00", "08:10", "T1"), ("footpath", "B", "B", "08:10", "08:20", 10),
         ("run", "B", "C", "08:20", "08:30", "T3")]
DUR3 = [("run", "A", "B", "08:00", "08:10", "T1"), ("footpath", "B", "B", "08:10", "08:13", 3),
        ("run", "B", "C", "08:13", "08:22", "T2")]
DUR4 = [("run", "A", "B", "08:00", "08:10", "T1"), ("footpath", "B", "B", "08:10", "08:14", 4),
        ("run", "B", "C", "08:20", "08:30", "T3")]
DIRECT = [("run", "A", "C", "07:55", "08:35", "T4")]

# (query fields, expected journeys forward, expected journeys backward);
# the settings are transfer_time_settings{default_ = false, ...} unless empty
TRANSFER_TIME_CASES = [
    (dict(max_travel_time=30), [DEFAULT], [DEFAULT]),
    (dict(max_travel_time=30, tts_min_transfer_time=10), [DUR10], [DUR10]),
    (dict(max_travel_time=30, tts_factor=1.5), [DUR3], [DUR3]),
    (dict(max_travel_time=30, tts_factor=2.0), [DUR4], [DUR4]),
    (dict(max_travel_time=30, tts_min_transfer_time=10, tts_factor=2.0), [DUR10], [DUR10]),
    (dict(max_travel_time=30, tts_min_transfer_time=3, tts_factor=2.0), [DUR4], [DUR4]),
    (dict(max_travel_time=30, tts_additional_time=2), [DUR4], [DUR4]),
    (dict(max_travel_time=30, tts_additional_time=1, tts_factor=1.5), [DUR4], [DUR4]),
    (dict(max_travel_time=30, tts_min_transfer_time=3, tts_additional_time=1), [DUR4], [DUR4]),
    (dict(max_travel_time=30, tts_min_transfer_time=3, tts_additional_time=5, tts_factor=2.5),
     [DUR10], [DUR10]),
    # default max_travel_time: the direct trip too (printed in result order)
    (dict(), [DIRECT, DEFAULT], [DEFAULT, DIRECT]),
]
# forward only (:306-340); the allowed_claszes = {} case (no transit) is not
# here: this build has no class filter (every class is allowed)
TRANSFER_TIME_FWD_ONLY = [
    (dict(max_travel_time=29, tts_factor=2.0), []),
    (dict(max_travel_time=-1440), [DIRECT, DEFAULT]),  # invalid: sanitized to 7200
    (dict(max_transfers=0), [DIRECT]),
    (dict(max_transfers=0, max_travel_time=35), []),
]

MAX_INTERVAL_FEED = r"""
# agency.txt
agency_id,agency_name,agency_url,agency_timezone
DB,Deutsche Bahn,https://deutschebahn.com,Europe/Berlin

# stops.txt
stop_id,stop_name,stop_desc,stop_lat,stop_lon,stop_url,location_type,parent_station
A,A,,0.0,1.0,,
B,B,,2.0,3.0,,

# calendar_dates.txt
service_id,date,exception_type
S_RE1,20190501,1

# routes.txt
route_id,agency_id,route_short_name,route_long_name,route_desc,route_type
R_RE1,DB,RE 1,,,3

# trips.txt
route_id,service_id,trip_id,trip_headsign,block_id
R_RE1,S_RE1,T_RE1,RE 1,
R_RE1,S_RE1,T_RE2,RE 1,
R_RE1,S_RE1,T_RE3,RE 1,

# stop_times.txt
trip_id,arrival_time,departure_time,stop_id,stop_sequence,pickup_type,drop_off_type
T_RE1,12:00:00,12:00:00,A,1,0,0
T_RE1,13:00:00,13:00:00,B,2,0,0
T_RE2,13:00:00,13:00:00,A,1,0,0
T_RE2,14:00:00,14:00:00,B,2,0,0
T_RE3,14:00:00,14:00:00,A,1,0,0
T_RE3,15:00:00,15:00:00,B,2,0,0
"""
MAX_INTERVAL_RANGE = ((2019, 3, 25), (2019, 11, 1))
# interval [10:00, 11:00) UTC on 2019-05-01, min_connection_count 3, extended
# later, max_interval [10:00, 12:00)
MAX_INTERVAL_QUERY = dict(interval=("10:00", "11:00"), min_connection_count=3,
                          extend_interval_later=1, max_interval=("10:00", "12:00"))
MAX_INTERVAL_EXPECTED = [[("run", "A", "B", "10:00", "11:00", "T_RE1")],
                         [("run", "A", "B", "11:00", "12:00", "T_RE2")]]
