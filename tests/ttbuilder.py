"""Tiny-timetable builder for tests: lines with explicit trips (UTC minutes),
grouped into nigiri routes the way the GTFS/HRD loaders do (first departure
normalised below 24:00 with the traffic days shifted, FIFO order modulo
1440 per route — loader/get_index.h:8-32)."""
import numpy as np

import nigiri_amd as na

N_SPECIAL = 9


def day_number(y, m, d):
    """Days since 1970-01-01 (proleptic Gregorian)."""
    import datetime
    return (datetime.date(y, m, d) - datetime.date(1970, 1, 1)).days


class Builder:
    def __init__(self, first_day, n_ext_days, n_stops):
        """first_day: days since epoch of date_range.from; stops are 9..9+n-1."""
        self.first_day = first_day
        self.n_ext = n_ext_days
        self.n_loc = N_SPECIAL + n_stops
        self.transfer = np.full(self.n_loc, 2, np.uint8)
        self.transfer[:N_SPECIAL] = 0
        self.lines = []  # (stops, [(times, days)])
        self.fps = []

    @property
    def internal_from(self):
        return (self.first_day - 5) * 1440

    def internal_day(self, day_since_epoch):
        return day_since_epoch - (self.first_day - 5)

    def add_line(self, stops, trips, in_out=None):
        """trips: list of (times, service_days) where times = [dep0, arr1, dep1, ...,
        arr_n-1] in minutes after midnight UTC of the service day (may exceed 1440)
        and service_days = list of days since epoch."""
        self.lines.append((list(stops), list(trips), in_out))

    def add_footpath(self, a, b, dur, both=True):
        self.fps.append((a, b, dur))
        if both:
            self.fps.append((b, a, dur))

    def build(self):
        route_stops, rso = [], [0]
        rtf, rstf = [0], [0]
        events, tdays = [], []
        bitfields = {}

        def bf_index(bits):
            key = tuple(bits)
            if key not in bitfields:
                bitfields[key] = len(bitfields)
            return bitfields[key]

        for stops, trips, in_out in self.lines:
            expanded = []
            for times, days in trips:
                shift = times[0] // 1440
                t = [x - shift * 1440 for x in times]
                words = [0] * 8
                for d in days:
                    bit = self.internal_day(d) + shift
                    assert 0 <= bit < 512
                    words[bit // 64] |= 1 << (bit % 64)
                expanded.append((t, words))
            groups = []
            for tr in expanded:
                placed = False
                for g in groups:
                    key0 = tr[0][0] % 1440
                    idx = 0
                    while idx < len(g) and g[idx][0][0] % 1440 < key0:
                        idx += 1
                    ok = True
                    for i, v in enumerate(tr[0]):
                        if idx > 0 and v % 1440 < g[idx - 1][0][i] % 1440:
                            ok = False
                        if idx < len(g) and v % 1440 > g[idx][0][i] % 1440:
                            ok = False
                    if ok:
                        g.insert(idx, tr)
                        placed = True
                        break
                if not placed:
                    groups.append([tr])
            for g in groups:
                for j, s in enumerate(stops):
                    io = in_out[j] if in_out else (1, 1)
                    route_stops.append((s & ((1 << 28) - 1)) | (io[0] << 28) | (io[1] << 29)
                                       | (io[0] << 30) | (io[1] << 31))
                rso.append(len(route_stops))
                for t, words in g:
                    tdays.append(bf_index(words))
                rtf.append(len(tdays))
                for ev in range(2 * (len(stops) - 1)):
                    for t, _ in g:
                        m = t[ev]
                        events.append(((m % 1440) << 5) | (m // 1440))
                rstf.append(len(events))
        bf = np.zeros((max(len(bitfields), 1), 8), np.uint64)
        for key, i in bitfields.items():
            bf[i] = np.array(key, np.uint64)
        fo = [[] for _ in range(self.n_loc)]
        for a, b, d in self.fps:
            fo[a].append(b | (d << 23))
        fpo = [0]
        fp = []
        for l in range(self.n_loc):
            fp += fo[l]
            fpo.append(len(fp))
        return na.Timetable.from_arrays(
            n_locations=self.n_loc, route_stop_offsets=rso, route_stops=route_stops,
            route_transport_from=rtf, route_stop_time_from=rstf, event_times=events,
            transport_traffic_days=tdays, bitfields=bf.reshape(-1), transfer_time=self.transfer,
            fp_out_offsets=fpo, fp_out=fp,
            date_range_from=self.first_day * 1440,
            date_range_to=(self.first_day + self.n_ext) * 1440,
            internal_from=self.internal_from, n_days=self.n_ext + 6)
