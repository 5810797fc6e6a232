"""GTFS feeds of nigiri's routing tests, loaded by the product's GTFS loader
(nigiri_amd_tt_load_gtfs), searched by the CPU oracle and compared with the
journeys nigiri's tests print (tests/golden/gtfs_routing.py); plus the
loader's local -> UTC conversion against Python's zoneinfo."""
import datetime
import zoneinfo

import numpy as np
import pytest

import nigiri_amd as na
import oracle
from golden import gtfs_routing as G

EPOCH = datetime.date(1970, 1, 1)


def day_minutes(ymd):
    return (datetime.date(*ymd) - EPOCH).days * 1440


def hm(s):
    h, m = s.split(":")
    return int(h) * 60 + int(m)


def fmt(t, day0):
    return f"{(t - day0) // 60:02d}:{(t - day0) % 60:02d}"


def named_legs(tt, legs, day0):
    """Legs as the fixtures write them: runs by trip_id, footpaths by duration."""
    out = []
    for kind, fr, to, dep, arr, what in legs:
        if kind == na.LEG_RUN:
            out.append(("run", tt.location_id(fr), tt.location_id(to), fmt(dep, day0), fmt(arr, day0),
                        tt.trip_id(what[0])))
        elif kind == na.LEG_FOOTPATH:
            out.append(("footpath", tt.location_id(fr), tt.location_id(to), fmt(dep, day0), fmt(arr, day0),
                        what[1]))
        else:
            out.append(("offset", tt.location_id(fr), tt.location_id(to), fmt(dep, day0), fmt(arr, day0),
                        tt.location_id(what[0]), what[1]))
    return out


_TT = {}


def transfer_time_tt():
    if "tts" not in _TT:
        _TT["tts"] = na.Timetable.load_gtfs(G.TRANSFER_TIME_FEED, *G.TRANSFER_TIME_RANGE)
    return _TT["tts"]


def max_interval_tt():
    if "mi" not in _TT:
        _TT["mi"] = na.Timetable.load_gtfs(G.MAX_INTERVAL_FEED, *G.MAX_INTERVAL_RANGE)
    return _TT["mi"]


def tts_query(tt, fields, fwd):
    """transfer_time_settings_test.cc:155-163: A -> C forward, C -> A backward, the
    date range as the interval; settings given means default_ = false."""
    a, c = tt.location("A"), tt.location("C")
    d0, d1 = (day_minutes(x) for x in G.TRANSFER_TIME_RANGE)
    spec = dict(start=[(a if fwd else c, 0)], dest=[(c if fwd else a, 0)], interval=(d0, d1),
                dir=na.DIR_FWD if fwd else na.DIR_BWD, use_start_footpaths=0)
    spec.update(fields)
    if any(k.startswith("tts_") for k in fields):
        spec["tts_default"] = 0
    return spec


def tts_cases():
    for fields, fwd, bwd in G.TRANSFER_TIME_CASES:
        yield fields, True, fwd
        yield fields, False, bwd
    for fields, fwd in G.TRANSFER_TIME_FWD_ONLY:
        yield fields, True, fwd


TTS_CASES = list(tts_cases())


@pytest.mark.parametrize("fields,fwd,expected", TTS_CASES,
                         ids=[f"{'fwd' if f else 'bwd'}-{i}" for i, (_, f, _) in enumerate(TTS_CASES)])
def test_transfer_time_settings_golden(fields, fwd, expected):
    tt = transfer_time_tt()
    qs = na.Queries.build([tts_query(tt, fields, fwd)])
    out = oracle.search(tt, qs[0])
    day0 = day_minutes(G.DATE)
    assert [named_legs(tt, l, day0) for l in out["legs"]] == expected
    # journeys agree with their legs: (departure, arrival, transfers)
    for (start, dest, _, transfers), legs in zip(out["journeys"], expected):
        assert (min(start, dest), max(start, dest)) == (day0 + hm(legs[0][3]), day0 + hm(legs[-1][4]))
        assert transfers == sum(1 for x in legs if x[0] == "run") - 1


@pytest.mark.parametrize("fields,fwd,expected", [c for c in TTS_CASES if c[2]][:12])
def test_transfer_time_settings_pong_equals_range(fields, fwd, expected):
    # test/raptor_search.cc:269-299: pong finds the interval's journeys too
    tt = transfer_time_tt()
    qs = na.Queries.build([tts_query(tt, fields, fwd)])
    rng = oracle.search(tt, qs[0])
    pg = oracle.pong(tt, qs[0])
    day0 = day_minutes(G.DATE)
    key = lambda out: sorted(  # noqa: E731
        (min(j[0], j[1]), max(j[0], j[1]), j[3], tuple(named_legs(tt, l, day0)))
        for j, l in zip(out["journeys"], out["legs"]))
    assert key(pg) == key(rng)


def max_interval_query(tt, **over):
    d = day_minutes(G.DATE)
    q = dict(G.MAX_INTERVAL_QUERY)
    q.update(over)
    spec = dict(start=[(tt.location("A"), 0)], dest=[(tt.location("B"), 0)], use_start_footpaths=0,
                interval=tuple(d + hm(x) for x in q.pop("interval")))
    if "max_interval" in q:
        mi = q.pop("max_interval")
        if mi is not None:
            spec["max_interval"] = tuple(d + hm(x) for x in mi)
    spec.update(q)
    return spec


def test_max_interval_golden():
    tt = max_interval_tt()
    qs = na.Queries.build([max_interval_query(tt)])
    out = oracle.search(tt, qs[0])
    day0 = day_minutes(G.DATE)
    assert [named_legs(tt, l, day0) for l in out["legs"]] == G.MAX_INTERVAL_EXPECTED


def test_max_interval_unset_finds_the_third_trip():
    # without max_interval_ the initial widening (interval_estimate.h:40-80:
    # +1 h + 3 h later) reaches T_RE3 at 12:00 UTC
    tt = max_interval_tt()
    qs = na.Queries.build([max_interval_query(tt, max_interval=None)])
    out = oracle.search(tt, qs[0])
    day0 = day_minutes(G.DATE)
    got = [named_legs(tt, l, day0) for l in out["legs"]]
    assert got == G.MAX_INTERVAL_EXPECTED + [[("run", "A", "B", "12:00", "13:00", "T_RE3")]]


def test_timeout_zero_stops_interval_extension():
    # min_connection_count 4 with three trips: the initial widening searches
    # [08:00, 14:00) (+1 h + 4 h later, interval_estimate.h:40-80); finding 3 < 4,
    # the search extends later until the interval cannot grow, unless the
    # timeout is reached (search.h:240-244), which timeout 0 always is
    tt = max_interval_tt()
    d = day_minutes(G.DATE)
    base = dict(interval=("08:00", "09:00"), min_connection_count=4, max_interval=None)
    no_to = oracle.search(tt, na.Queries.build([max_interval_query(tt, **base)])[0])
    to0 = oracle.search(tt, na.Queries.build([max_interval_query(tt, **base, timeout_s=0)])[0])
    to_big = oracle.search(tt, na.Queries.build([max_interval_query(tt, **base, timeout_s=3600)])[0])
    assert len(no_to["journeys"]) == len(to0["journeys"]) == 3
    assert to0["interval"] == (d + hm("08:00"), d + hm("14:00"))
    assert no_to["interval"][1] > d + hm("14:00")
    assert to_big["interval"] == no_to["interval"] and to_big["journeys"] == no_to["journeys"]


def test_pong_timeout_zero_runs_nothing():
    # pong.cc:242-245: the ping loop does not start once the timeout is reached
    tt = transfer_time_tt()
    qs = na.Queries.build([dict(tts_query(tt, {}, True), timeout_s=0)])
    assert oracle.pong(tt, qs[0])["journeys"] == []
    qs = na.Queries.build([tts_query(tt, {}, True)])
    assert oracle.pong(tt, qs[0])["journeys"] != []


# ---- local -> UTC (noon offsets, local_to_utc.h) against zoneinfo ------------

def daily_feed(zone, dep, arr):
    return f"""
# agency.txt
agency_id,agency_name,agency_url,agency_timezone
X,X,https://example.org,{zone}

# stops.txt
stop_id,stop_name,stop_desc,stop_lat,stop_lon,stop_url,location_type,parent_station
A,A,,0.0,1.0,,
B,B,,2.0,3.0,,

# calendar.txt
service_id,monday,tuesday,wednesday,thursday,friday,saturday,sunday,start_date,end_date
S,1,1,1,1,1,1,1,20190101,20191231

# routes.txt
route_id,agency_id,route_short_name,route_long_name,route_desc,route_type
R,X,1,,,3

# trips.txt
route_id,service_id,trip_id,trip_headsign,block_id
R,S,T,,

# stop_times.txt
trip_id,arrival_time,departure_time,stop_id,stop_sequence,pickup_type,drop_off_type
T,{dep},{dep},A,1,0,0
T,{arr},{arr},B,2,0,0
"""


def utc_departures(tt):
    """All UTC departure minutes at the first stop, over every transport and traffic day."""
    v = tt.view
    bf = tt.array("bitfields").reshape(-1, 8)
    tdays = tt.array("transport_traffic_days")
    rtf = tt.array("route_transport_from")
    rstf = tt.array("route_stop_time_from")
    ev = tt.array("event_times")
    out = set()
    for r in range(v.n_routes):
        n_t = int(rtf[r + 1] - rtf[r])
        for k in range(n_t):
            e = int(ev[rstf[r] + k])
            mins = (e & 31) * 1440 + (e >> 5)
            words = bf[tdays[rtf[r] + k]]
            for day in range(512):
                if (int(words[day // 64]) >> (day % 64)) & 1:
                    out.add(v.internal_from + day * 1440 + mins)
    return out


@pytest.mark.parametrize("zone", ["Europe/Berlin", "America/New_York", "Australia/Sydney",
                                  "Asia/Kolkata", "America/Sao_Paulo", "UTC"])
@pytest.mark.parametrize("dep,arr", [("12:00:00", "12:30:00"), ("23:40:00", "24:20:00"),
                                     ("25:10:00", "25:50:00"), ("00:20:00", "01:00:00")])
def test_local_to_utc_matches_zoneinfo(zone, dep, arr):
    tt = na.Timetable.load_gtfs(daily_feed(zone, dep, arr), (2019, 1, 1), (2020, 1, 1))
    z = zoneinfo.ZoneInfo(zone)
    h, m, _ = (int(x) for x in dep.split(":"))
    last_day_offset = int(arr.split(":")[0]) // 24
    want = set()
    for d in range(365):
        day = datetime.date(2019, 1, 1) + datetime.timedelta(days=d)
        # local_to_utc.h:201-206: a service day counts when the days its trip
        # runs on overlap the date range
        if not (day + datetime.timedelta(days=h // 24) < datetime.date(2020, 1, 1)
                and datetime.date(2019, 1, 1) < day + datetime.timedelta(days=last_day_offset + 1)):
            continue
        # nigiri converts with the UTC offset at local noon of the service day
        noon = datetime.datetime(day.year, day.month, day.day, 12, tzinfo=z)
        off = int(noon.utcoffset().total_seconds() // 60)
        want.add((day - EPOCH).days * 1440 + h * 60 + m - off)
    assert utc_departures(tt) == want
    # routes keep FIFO order modulo 1440 (the loader's grouping) and pass validation
    assert tt.view.n_transports >= 1


def test_noon_offset_on_dst_switch_days():
    # Europe/Berlin switches at 02:00 on 2019-03-31 and 03:00 on 2019-10-27:
    # noon of both days already has the new offset (+120, then +60)
    tt = na.Timetable.load_gtfs(daily_feed("Europe/Berlin", "12:00:00", "12:30:00"), (2019, 1, 1),
                                (2020, 1, 1))
    deps = utc_departures(tt)
    for ymd, off in [((2019, 3, 30), 60), ((2019, 3, 31), 120), ((2019, 10, 26), 120),
                     ((2019, 10, 27), 60)]:
        assert day_minutes(ymd) + 720 - off in deps


def test_gtfs_loader_refuses_unsupported_features():
    feed = na.parse_mem_dir(daily_feed("UTC", "12:00:00", "12:30:00"))
    feed["frequencies.txt"] = "trip_id,start_time,end_time,headway_secs\nT,06:00:00,08:00:00,600\n"
    with pytest.raises(na.NigiriError) as e:
        na.Timetable.load_gtfs(feed, (2019, 1, 1), (2019, 2, 1))
    assert e.value.code == na.EUNSUPPORTED
    feed = na.parse_mem_dir(daily_feed("Nowhere/Zone", "12:00:00", "12:30:00"))
    with pytest.raises(na.NigiriError):
        na.Timetable.load_gtfs(feed, (2019, 1, 1), (2019, 2, 1))


def test_gtfs_transfers_become_symmetric_footpaths():
    feed = na.parse_mem_dir(G.TRANSFER_TIME_FEED)
    feed["transfers.txt"] = ("from_stop_id,to_stop_id,transfer_type,min_transfer_time\n"
                             "A,C,2,300\nB,B,2,240\n")
    tt = na.Timetable.load_gtfs(feed, *G.TRANSFER_TIME_RANGE)
    a, b, c = tt.location("A"), tt.location("B"), tt.location("C")
    off, fp = tt.array("fp_out_offsets"), tt.array("fp_out")
    out = {(l, int(x) & ((1 << 23) - 1), int(x) >> 23) for l in range(tt.n_locations)
           for x in fp[off[l]:off[l + 1]]}
    assert out == {(a, c, 5), (c, a, 5)}
    assert int(tt.array("transfer_time")[b]) == 4  # same-stop transfer: transfer_time_
