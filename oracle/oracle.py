"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU
baseline leg of bench.py.  The product package never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")


class Stats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "n_routes_visited", "n_route_stops_visited", "n_earliest_trip_calls", "n_events_scanned",
        "n_route_day_tests", "n_transport_day_tests", "n_marked_stations",
        "n_location_routes_read", "n_transfer_stations", "n_footpaths_visited",
        "n_label_writes", "n_rounds", "n_epochs", "alg_bytes")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


def build(arch_flags=""):
    """Compiles the oracle with its Makefile (host C++ only)."""
    cmd = ["make", "-C", HERE]
    if arch_flags:
        cmd.append(f"ARCHFLAGS={arch_flags}")
    subprocess.run(cmd, check=True, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    return LIB_PATH


BASELINE_PATH = os.path.join(HERE, "_build", "liboracle_baseline.so")


def build_baseline():
    """Compiles the timed CPU baseline (-O3 -march=native, no byte counters) for
    the machine it runs on."""
    subprocess.run(["make", "-C", HERE, "baseline"], check=True, stdout=subprocess.PIPE,
                   stderr=subprocess.STDOUT)
    return BASELINE_PATH


_lib = None
_baseline = None


def baseline_lib():
    """The baseline build (search_many only is meant to be called on it)."""
    global _baseline
    if _baseline is None:
        h = C.CDLL(build_baseline())
        _bind(h)
        _baseline = h
    return _baseline


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _bind(C.CDLL(LIB_PATH))
    return _lib


def _bind(h):
    """Declares the entry points' signatures on a loaded library."""
    from nigiri_amd import TtView, Query, Journey, Leg  # struct layouts only
    P = C.POINTER
    h.oracle_search.restype = C.c_int
    h.oracle_search.argtypes = [P(TtView), P(Query), P(Journey), C.c_uint32, P(C.c_uint32),
                                P(Leg), C.c_uint32, P(C.c_uint32),
                                P(C.c_int16), P(C.c_int16), P(C.c_int16), P(C.c_int16),
                                P(C.c_uint16), P(C.c_int32), P(Stats)]
    h.oracle_pong.restype = C.c_int
    h.oracle_pong.argtypes = [P(TtView), P(Query), P(Journey), C.c_uint32, P(C.c_uint32),
                              P(Leg), C.c_uint32, P(C.c_uint32), P(C.c_int32), P(Stats)]
    h.oracle_optimize_footpaths.restype = C.c_int
    h.oracle_optimize_footpaths.argtypes = [P(TtView), P(Query), P(Leg), C.c_uint32]
    h.oracle_lb.restype = C.c_int
    h.oracle_lb.argtypes = [P(TtView), P(Query), P(C.c_uint16)]
    h.oracle_fastest_direct.restype = C.c_int
    h.oracle_fastest_direct.argtypes = [P(TtView), P(Query), P(C.c_int16)]
    h.oracle_starts.restype = C.c_int
    h.oracle_starts.argtypes = [P(TtView), P(Query), P(C.c_int32), P(C.c_int32),
                                P(C.c_uint32), C.c_uint32, P(C.c_uint32)]
    h.oracle_search_many.restype = C.c_int
    h.oracle_search_many.argtypes = [P(TtView), P(Query), C.c_uint32, C.c_int,
                                     P(C.c_double), P(C.c_double), P(C.c_uint32), P(Stats)]
    return h


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def search(tt, q, state=False, cap=4096, leg_cap=1 << 15):
    """Runs one query; returns journeys (start, dest_time, dest, transfers), their legs
    (one list of Leg.as_tuple() per journey) and optionally the state."""
    from nigiri_amd import Journey, Leg
    L = tt.n_locations
    jr = (Journey * cap)()
    lg = (Leg * leg_cap)()
    n = C.c_uint32()
    n_legs = C.c_uint32()
    info = np.zeros(4, np.int32)
    stats = Stats()
    rt = best = tmp = tad = lb = None
    args = [None] * 5
    if state:
        rt = np.zeros((16, L), np.int16)
        best = np.zeros(L, np.int16)
        tmp = np.zeros(L, np.int16)
        tad = np.zeros(16, np.int16)
        lb = np.zeros(L, np.uint16)
        args = [_p(rt, C.c_int16), _p(best, C.c_int16), _p(tmp, C.c_int16), _p(tad, C.c_int16),
                _p(lb, C.c_uint16)]
    rc = lib().oracle_search(C.byref(tt.view), C.byref(q), jr, cap, C.byref(n), lg, leg_cap,
                             C.byref(n_legs), *args, _p(info, C.c_int32), C.byref(stats))
    if rc != 0:
        raise RuntimeError("oracle_search failed")
    if n.value > cap or n_legs.value > leg_cap:
        raise RuntimeError("oracle_search: output capacity exceeded")
    k = n.value
    js = [(jr[i].start_time, jr[i].dest_time, jr[i].dest, jr[i].transfers) for i in range(k)]
    legs = [[lg[jr[i].leg_offset + x].as_tuple() for x in range(jr[i].n_legs)]
            for i in range(k)]
    out = {"journeys": js, "legs": legs, "base_day": int(info[0]),
           "fastest_direct": int(info[1]), "interval": (int(info[2]), int(info[3])),
           "stats": stats.as_dict()}
    if state:
        out.update(round_times=rt, best=best, tmp=tmp, time_at_dest=tad, lb=lb)
    return out


def pong(tt, q, cap=4096, leg_cap=1 << 15):
    """Range search by pong (ping/pong pairs); same result layout as search().
    Raises RuntimeError where the reference throws."""
    from nigiri_amd import Journey, Leg
    jr = (Journey * cap)()
    lg = (Leg * leg_cap)()
    n = C.c_uint32()
    n_legs = C.c_uint32()
    info = np.zeros(4, np.int32)
    stats = Stats()
    rc = lib().oracle_pong(C.byref(tt.view), C.byref(q), jr, cap, C.byref(n), lg, leg_cap,
                           C.byref(n_legs), _p(info, C.c_int32), C.byref(stats))
    if rc != 0:
        raise RuntimeError("oracle_pong failed")
    if n.value > cap or n_legs.value > leg_cap:
        raise RuntimeError("oracle_pong: output capacity exceeded")
    k = n.value
    js = [(jr[i].start_time, jr[i].dest_time, jr[i].dest, jr[i].transfers) for i in range(k)]
    legs = [[lg[jr[i].leg_offset + x].as_tuple() for x in range(jr[i].n_legs)]
            for i in range(k)]
    return {"journeys": js, "legs": legs, "base_day": int(info[0]),
            "fastest_direct": int(info[1]), "interval": (int(info[2]), int(info[3])),
            "stats": stats.as_dict()}


def optimize_footpaths(tt, q, legs):
    """Runs optimize_footpaths on a list of Leg structures (travel order); returns the
    optimised legs as Leg.as_tuple() tuples."""
    from nigiri_amd import Leg
    arr = (Leg * max(len(legs), 1))(*legs)
    if lib().oracle_optimize_footpaths(C.byref(tt.view), C.byref(q), arr, len(legs)) != 0:
        raise RuntimeError("oracle_optimize_footpaths failed")
    return [arr[i].as_tuple() for i in range(len(legs))]


def lower_bounds(tt, q):
    lb = np.zeros(tt.n_locations, np.uint16)
    if lib().oracle_lb(C.byref(tt.view), C.byref(q), _p(lb, C.c_uint16)) != 0:
        raise RuntimeError("oracle_lb failed")
    return lb


def fastest_direct(tt, q):
    out = C.c_int16()
    if lib().oracle_fastest_direct(C.byref(tt.view), C.byref(q), C.byref(out)) != 0:
        raise RuntimeError("oracle_fastest_direct failed")
    return out.value


def starts(tt, q, cap=1 << 16):
    a = np.zeros(cap, np.int32)
    b = np.zeros(cap, np.int32)
    s = np.zeros(cap, np.uint32)
    n = C.c_uint32()
    if lib().oracle_starts(C.byref(tt.view), C.byref(q), _p(a, C.c_int32), _p(b, C.c_int32),
                           _p(s, C.c_uint32), cap, C.byref(n)) != 0:
        raise RuntimeError("oracle_starts failed")
    k = min(n.value, cap)
    return list(zip(a[:k].tolist(), b[:k].tolist(), s[:k].tolist()))


def search_many(tt, queries, threads=1, baseline=False):
    """Throughput run (benchmark.cc:210-270 style); baseline=True times the
    -O3 -march=native build without byte counters."""
    n = len(queries)
    lat = np.zeros(max(n, 1), np.float64)
    nj = np.zeros(max(n, 1), np.uint32)
    wall = C.c_double()
    stats = Stats()
    h = baseline_lib() if baseline else lib()
    rc = h.oracle_search_many(C.byref(tt.view), queries.q, n, threads, _p(lat, C.c_double),
                                  C.byref(wall), _p(nj, C.c_uint32), C.byref(stats))
    if rc != 0:
        raise RuntimeError("oracle_search_many failed")
    return {"latency_ms": lat[:n], "wall_ms": wall.value, "n_journeys": nj[:n],
            "stats": stats.as_dict()}
