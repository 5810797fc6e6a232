"""nigiri_amd — MI355X RAPTOR for nigiri's earliest-arrival and range queries.

Python mirror of the C ABI in include/nigiri_amd.h (ctypes).  The compute
path is libnigiri_amd.so (HIP kernels for gfx950 + host C++); this module
only marshals arguments.  Importing it does not touch the GPU; creating an
Image does.  If the shared library is missing every entry point raises.
"""
import atexit
import ctypes as C
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# NIGIRI_AMD_LIB selects another build of the library (e.g. to compare builds)
LIB_PATH = os.environ.get("NIGIRI_AMD_LIB") or os.path.join(_HERE, "libnigiri_amd.so")

OK, EINVAL, ENOMEM, EDEVICE, ECAPACITY, EUNSUPPORTED, ERECONSTRUCT = 0, -1, -2, -3, -4, -5, -6
DIR_FWD, DIR_BWD = 0, 1
MATCH_EXACT, MATCH_INTERMODAL = 0, 1
N_SPECIAL = 9
END = 1


class NigiriError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"nigiri_amd error {code}: {msg}")
        self.code = code


class TtView(C.Structure):
    _fields_ = [
        ("n_locations", C.c_uint32), ("n_routes", C.c_uint32), ("n_transports", C.c_uint32),
        ("n_bitfields", C.c_uint32), ("n_route_stops", C.c_uint32), ("n_event_times", C.c_uint32),
        ("n_location_routes", C.c_uint32), ("n_footpaths", C.c_uint32),
        ("n_lb_fwd", C.c_uint32), ("n_lb_bwd", C.c_uint32),
        ("date_range_from", C.c_int32), ("date_range_to", C.c_int32),
        ("internal_from", C.c_int32), ("n_days", C.c_uint32),
        ("route_stop_offsets", C.POINTER(C.c_uint32)), ("route_stops", C.POINTER(C.c_uint32)),
        ("route_transport_from", C.POINTER(C.c_uint32)),
        ("route_stop_time_from", C.POINTER(C.c_uint32)),
        ("event_times", C.POINTER(C.c_uint16)),
        ("transport_traffic_days", C.POINTER(C.c_uint32)),
        ("route_traffic_days", C.POINTER(C.c_uint32)), ("bitfields", C.POINTER(C.c_uint64)),
        ("location_route_offsets", C.POINTER(C.c_uint32)),
        ("location_routes", C.POINTER(C.c_uint32)),
        ("transfer_time", C.POINTER(C.c_uint8)),
        ("fp_out_offsets", C.POINTER(C.c_uint32)), ("fp_out", C.POINTER(C.c_uint32)),
        ("fp_in_offsets", C.POINTER(C.c_uint32)), ("fp_in", C.POINTER(C.c_uint32)),
        ("lb_fwd_offsets", C.POINTER(C.c_uint32)), ("lb_fwd", C.POINTER(C.c_uint32)),
        ("lb_bwd_offsets", C.POINTER(C.c_uint32)), ("lb_bwd", C.POINTER(C.c_uint32)),
        ("coord_x", C.POINTER(C.c_float)), ("coord_y", C.POINTER(C.c_float)),
    ]


class Query(C.Structure):
    _fields_ = [
        ("dir", C.c_int32), ("start_match_mode", C.c_int32), ("dest_match_mode", C.c_int32),
        ("use_start_footpaths", C.c_int32), ("is_interval", C.c_int32),
        ("start_time", C.c_int32), ("interval_from", C.c_int32), ("interval_to", C.c_int32),
        ("n_start", C.c_uint32), ("n_dest", C.c_uint32),
        ("start_loc", C.POINTER(C.c_uint32)), ("start_dur", C.POINTER(C.c_int16)),
        ("dest_loc", C.POINTER(C.c_uint32)), ("dest_dur", C.POINTER(C.c_int16)),
        ("max_start_offset", C.c_int16), ("max_travel_time", C.c_int16),
        ("max_transfers", C.c_uint8), ("pad_", C.c_uint8 * 3),
        ("min_connection_count", C.c_uint32), ("extend_interval_earlier", C.c_int32),
        ("extend_interval_later", C.c_int32), ("prf_idx", C.c_int32),
        ("tts_default", C.c_int32), ("tts_min_transfer_time", C.c_int16),
        ("tts_additional_time", C.c_int16), ("tts_factor", C.c_float),
        ("has_fastest_direct", C.c_int32), ("fastest_direct", C.c_int16), ("pad2_", C.c_int16),
        ("fastest_direct_factor", C.c_double),
        ("has_max_interval", C.c_int32), ("max_interval_from", C.c_int32),
        ("max_interval_to", C.c_int32), ("has_timeout", C.c_int32), ("timeout_s", C.c_int32),
        ("pad3_", C.c_int32),
    ]


class Journey(C.Structure):
    _fields_ = [("start_time", C.c_int32), ("dest_time", C.c_int32), ("dest", C.c_uint32),
                ("transfers", C.c_uint8), ("pad_", C.c_uint8 * 3),
                ("leg_offset", C.c_uint32), ("n_legs", C.c_uint32)]


LEG_RUN, LEG_FOOTPATH, LEG_OFFSET = 0, 1, 2


class Leg(C.Structure):
    _fields_ = [("kind", C.c_int32), ("from_", C.c_uint32), ("to", C.c_uint32),
                ("dep_time", C.c_int32), ("arr_time", C.c_int32), ("transport", C.c_uint32),
                ("route", C.c_uint32), ("day", C.c_uint16), ("stop_from", C.c_uint16),
                ("stop_to", C.c_uint16), ("duration", C.c_int16), ("target", C.c_uint32)]

    def as_tuple(self):
        """(kind, from, to, dep, arr, what): what is (transport, route, day, stop_from, stop_to)
        for a run and (target, duration) for a footpath or an offset."""
        if self.kind == LEG_RUN:
            what = (self.transport, self.route, self.day, self.stop_from, self.stop_to)
        else:
            what = (self.target, self.duration)
        return (self.kind, self.from_, self.to, self.dep_time, self.arr_time, what)


class Result(C.Structure):
    _fields_ = [("status", C.c_int32), ("n_journeys", C.c_uint32), ("journey_offset", C.c_uint32),
                ("n_epochs", C.c_uint32), ("n_rounds", C.c_uint32), ("slot", C.c_int32),
                ("interval_from", C.c_int32), ("interval_to", C.c_int32),
                ("fastest_direct", C.c_int16), ("base_day", C.c_int16), ("latency_ms", C.c_double),
                ("queue_ms", C.c_double)]


class AlgoInit(C.Structure):
    _fields_ = [("dir", C.c_int32), ("base_day", C.c_int32), ("n_dest", C.c_uint32),
                ("dest", C.POINTER(C.c_uint32)), ("intermodal", C.c_int32),
                ("n_dist", C.c_uint32), ("dist_loc", C.POINTER(C.c_uint32)),
                ("dist_min", C.POINTER(C.c_uint16)), ("lb", C.POINTER(C.c_uint16)),
                ("tts_default", C.c_int32), ("tts_min_transfer_time", C.c_int16),
                ("tts_additional_time", C.c_int16), ("tts_factor", C.c_float)]


class EngineStats(C.Structure):
    _fields_ = [("epochs", C.c_uint64), ("rounds", C.c_uint64),
                ("route_scan_launches", C.c_uint64), ("route_scan_ms", C.c_double),
                ("transfer_launches", C.c_uint64), ("transfer_ms", C.c_double),
                ("work", C.c_uint64 * 16), ("reconstructed", C.c_uint64),
                ("reconstruct_ms", C.c_double), ("legs_ms", C.c_double),
                ("results_ms", C.c_double), ("graph_rounds", C.c_uint64),
                ("graph_captures", C.c_uint64)]


class EngineConfig(C.Structure):
    _fields_ = [("admission_wait_ms", C.c_double), ("grid_waves", C.c_uint32),
                ("eager_rounds", C.c_uint32), ("route_claim", C.c_uint32),
                ("transfer_claim", C.c_uint32), ("pack_max_entries", C.c_uint32),
                ("admit_when_idle", C.c_uint32),
                ("transfer_pack_max_entries", C.c_uint32), ("state_budget_bytes", C.c_uint64),
                ("loose_lanes", C.c_uint32), ("pad_", C.c_uint32)]


class TuneTrial(C.Structure):
    _fields_ = [("knob", C.c_uint32), ("value", C.c_uint32), ("ms", C.c_double)]


TUNE_KNOBS = ["grid_waves", "route_claim", "transfer_claim", "pack_max_entries",
              "transfer_pack_max_entries", "loose_lanes"]


WORK_NAMES = ["route_stops", "events", "route_day_tests", "transport_day_tests", "tmp_writes",
              "transfer_stations", "footpaths", "marked_stations", "location_routes",
              "wave_stops", "transfer_items", "hazard_items", "hazard_lanes", "loose_wave_stops",
              "loose_items", "skipped_stops"]


# every entry point of include/nigiri_amd.h: name -> (restype, argtypes)
_P = C.c_void_p
SIGNATURES = {
    "nigiri_amd_last_error": (C.c_char_p, []),
    "nigiri_amd_version": (C.c_int, []),
    "nigiri_amd_tt_generate": (C.c_int, [C.c_int, C.c_uint64, C.c_double, C.POINTER(_P)]),
    "nigiri_amd_tt_from_arrays": (C.c_int, [C.POINTER(TtView), C.POINTER(_P)]),
    "nigiri_amd_tt_view": (C.c_int, [_P, C.POINTER(TtView)]),
    "nigiri_amd_tt_load_gtfs": (C.c_int, [C.POINTER(C.c_char_p), C.POINTER(C.c_char_p), C.c_uint32,
                                          C.c_int32, C.c_int32, C.c_char_p, C.POINTER(_P)]),
    "nigiri_amd_tt_location_id": (C.c_int, [_P, C.c_uint32, C.c_char_p, C.c_uint32]),
    "nigiri_amd_tt_lower_bounds": (C.c_int, [_P, C.c_int32, C.c_uint32, C.POINTER(C.c_uint32),
                                             C.POINTER(C.c_uint16), C.POINTER(C.c_uint16)]),
    "nigiri_amd_tt_find_location": (C.c_int, [_P, C.c_char_p, C.POINTER(C.c_uint32)]),
    "nigiri_amd_tt_transport_trip_id": (C.c_int, [_P, C.c_uint32, C.c_char_p, C.c_uint32]),
    "nigiri_amd_tt_free": (None, [_P]),
    "nigiri_amd_queries_generate": (C.c_int, [_P, C.c_int, C.c_uint64, C.c_uint32,
                                              C.POINTER(Query), C.POINTER(C.c_uint32),
                                              C.POINTER(C.c_int16), C.POINTER(C.c_uint32)]),
    "nigiri_amd_image_create": (C.c_int, [_P, C.c_int, C.POINTER(_P)]),
    "nigiri_amd_image_bytes": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "nigiri_amd_image_destroy": (None, [_P]),
    "nigiri_amd_lower_bounds": (C.c_int, [_P, C.c_int32, C.c_uint32, C.POINTER(C.c_uint32),
                                          C.POINTER(C.c_uint16), C.c_void_p]),
    "nigiri_amd_lower_bounds_noalloc": (C.c_int, [_P, C.c_int32, C.c_uint32, C.POINTER(C.c_uint32),
                                                  C.POINTER(C.c_uint16), C.c_void_p]),
    "nigiri_amd_fastest_direct": (C.c_int, [_P, C.POINTER(Query), C.c_uint32, C.POINTER(C.c_int16)]),
    "nigiri_amd_engine_create": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(_P)]),
    "nigiri_amd_engine_state_bytes": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "nigiri_amd_engine_destroy": (None, [_P]),
    "nigiri_amd_search": (C.c_int, [_P, C.POINTER(Query), C.c_uint32, C.POINTER(Result),
                                    C.POINTER(Journey), C.c_uint32, C.POINTER(Leg), C.c_uint32]),
    "nigiri_amd_algo_init": (C.c_int, [_P, C.c_uint32, C.POINTER(AlgoInit)]),
    "nigiri_amd_algo_reset_arrivals": (C.c_int, [_P, C.c_uint32]),
    "nigiri_amd_algo_next_start_time": (C.c_int, [_P, C.c_uint32]),
    "nigiri_amd_algo_add_start": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_int32]),
    "nigiri_amd_algo_execute": (C.c_int, [_P, C.c_uint32, C.c_int32, C.c_uint8, C.c_int32,
                                          C.POINTER(C.c_int16), C.POINTER(C.c_uint32)]),
    "nigiri_amd_algo_reconstruct": (C.c_int, [_P, C.c_uint32, C.POINTER(Query),
                                              C.POINTER(Journey), C.POINTER(Leg), C.c_uint32]),
    "nigiri_amd_algo_fill_bounds": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_uint32]),
    "nigiri_amd_algo_set_bounds": (C.c_int, [_P, C.c_uint32, C.c_uint32]),
    "nigiri_amd_pong": (C.c_int, [_P, C.POINTER(Query), C.c_uint32, C.POINTER(Result),
                                  C.POINTER(Journey), C.c_uint32, C.POINTER(Leg), C.c_uint32]),
    "nigiri_amd_algo_download": (C.c_int, [_P, C.c_uint32, C.POINTER(C.c_int16),
                                           C.POINTER(C.c_int16), C.POINTER(C.c_int16),
                                           C.POINTER(C.c_int16), C.POINTER(C.c_uint16)]),
    "nigiri_amd_engine_configure": (C.c_int, [_P, C.POINTER(EngineConfig)]),
    "nigiri_amd_engine_get_config": (C.c_int, [_P, C.POINTER(EngineConfig)]),
    "nigiri_amd_engine_autotune": (C.c_int, [_P, C.POINTER(Query), C.c_uint32, C.c_uint32,
                                             C.POINTER(EngineConfig), C.POINTER(TuneTrial),
                                             C.c_uint32, C.POINTER(C.c_uint32)]),
    "nigiri_amd_engine_profile": (C.c_int, [_P, C.c_int]),
    "nigiri_amd_engine_stats": (C.c_int, [_P, C.POINTER(EngineStats), C.c_int]),
    "nigiri_amd_serve": (C.c_int, [_P, C.POINTER(Query), C.POINTER(C.c_double), C.c_uint32,
                                   C.POINTER(Result), C.POINTER(Journey), C.c_uint32,
                                   C.POINTER(Leg), C.c_uint32, C.POINTER(C.c_double)]),
}

_lib_handle = None


def lib():
    """Loads libnigiri_amd.so (raises if it has not been built)."""
    global _lib_handle
    if _lib_handle is None:
        if not os.path.exists(LIB_PATH):
            raise NigiriError(EINVAL, f"{LIB_PATH} missing: run nigiri_amd.build.build()")
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib_handle = h
    return _lib_handle


def _check(code):
    if code != OK:
        raise NigiriError(code, lib().nigiri_amd_last_error().decode())


def ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


_VIEW_ARRAYS = {
    # name: (dtype, length field expression)
    "route_stop_offsets": (np.uint32, lambda v: v.n_routes + 1),
    "route_stops": (np.uint32, lambda v: v.n_route_stops),
    "route_transport_from": (np.uint32, lambda v: v.n_routes + 1),
    "route_stop_time_from": (np.uint32, lambda v: v.n_routes + 1),
    "event_times": (np.uint16, lambda v: v.n_event_times),
    "transport_traffic_days": (np.uint32, lambda v: v.n_transports),
    "route_traffic_days": (np.uint32, lambda v: v.n_routes),
    "bitfields": (np.uint64, lambda v: v.n_bitfields * 8),
    "location_route_offsets": (np.uint32, lambda v: v.n_locations + 1),
    "location_routes": (np.uint32, lambda v: v.n_location_routes),
    "transfer_time": (np.uint8, lambda v: v.n_locations),
    "fp_out_offsets": (np.uint32, lambda v: v.n_locations + 1),
    "fp_out": (np.uint32, lambda v: v.n_footpaths),
    "fp_in_offsets": (np.uint32, lambda v: v.n_locations + 1),
    "fp_in": (np.uint32, lambda v: v.n_footpaths),
    "lb_fwd_offsets": (np.uint32, lambda v: v.n_locations + 1),
    "lb_fwd": (np.uint32, lambda v: v.n_lb_fwd),
    "lb_bwd_offsets": (np.uint32, lambda v: v.n_locations + 1),
    "lb_bwd": (np.uint32, lambda v: v.n_lb_bwd),
}


def parse_mem_dir(text):
    """Splits text in the format of nigiri's loader::mem_dir::read into {file name: content}:
    a line "# name" starts a file, the lines up to the next such line are its content."""
    files, cur = {}, None
    for line in text.splitlines():
        if line.startswith("# "):
            cur = line[2:].strip()
            files[cur] = []
        elif cur is not None:
            files[cur].append(line)
    return {k: "\n".join(v).strip("\n") + "\n" for k, v in files.items()}


class Timetable:
    """Host timetable (nigiri's RAPTOR arrays as CSR)."""

    def __init__(self, handle):
        self.handle = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle
        self.view = TtView()
        _check(lib().nigiri_amd_tt_view(self.handle, C.byref(self.view)))

    @classmethod
    def generate(cls, cfg, seed=None, scale=1.0):
        """Synthetic timetable of workload `cfg` (1..5); default seed 0x6E69 + cfg."""
        h = C.c_void_p()
        _check(lib().nigiri_amd_tt_generate(cfg, 0x6E69 + cfg if seed is None else seed,
                                            float(scale), C.byref(h)))
        return cls(h)

    @classmethod
    def from_arrays(cls, **a):
        """Builds a timetable from plain arrays (see nigiri_amd_tt_from_arrays)."""
        v = TtView()
        keep = []

        def arr(name, dtype):
            x = np.ascontiguousarray(a[name], dtype=dtype)
            keep.append(x)
            return x

        n_loc = int(a["n_locations"])
        rso = arr("route_stop_offsets", np.uint32)
        v.n_locations = n_loc
        v.n_routes = len(rso) - 1
        v.route_stop_offsets = ptr(rso, C.c_uint32)
        v.route_stops = ptr(arr("route_stops", np.uint32), C.c_uint32)
        v.route_transport_from = ptr(arr("route_transport_from", np.uint32), C.c_uint32)
        v.route_stop_time_from = ptr(arr("route_stop_time_from", np.uint32), C.c_uint32)
        v.event_times = ptr(arr("event_times", np.uint16), C.c_uint16)
        v.transport_traffic_days = ptr(arr("transport_traffic_days", np.uint32), C.c_uint32)
        bf = arr("bitfields", np.uint64)
        v.n_bitfields = len(bf) // 8
        v.bitfields = ptr(bf, C.c_uint64)
        v.transfer_time = ptr(arr("transfer_time", np.uint8), C.c_uint8)
        v.fp_out_offsets = ptr(arr("fp_out_offsets", np.uint32), C.c_uint32)
        v.fp_out = ptr(arr("fp_out", np.uint32), C.c_uint32)
        v.date_range_from = int(a["date_range_from"])
        v.date_range_to = int(a["date_range_to"])
        v.internal_from = int(a["internal_from"])
        v.n_days = int(a["n_days"])
        h = C.c_void_p()
        _check(lib().nigiri_amd_tt_from_arrays(C.byref(v), C.byref(h)))
        return cls(h)

    @classmethod
    def load_gtfs(cls, files, date_from, date_to, tz_dir=None):
        """Loads a GTFS feed (see nigiri_amd_tt_load_gtfs).  files: dict of file name ->
        text, or a string in the format of nigiri's loader::mem_dir::read ("# name.txt"
        lines starting each file).  date_from / date_to: (y, m, d) of
        timetable::date_range_ ([from, to)).  tz_dir defaults to the tzdata package's
        zoneinfo directory when it is installed."""
        import datetime
        if isinstance(files, str):
            files = parse_mem_dir(files)
        if tz_dir is None:
            try:
                import tzdata
                tz_dir = os.path.join(os.path.dirname(tzdata.__file__), "zoneinfo")
            except ImportError:
                tz_dir = ""
        names = (C.c_char_p * max(len(files), 1))(*[k.encode() for k in files])
        texts = (C.c_char_p * max(len(files), 1))(*[v.encode() for v in files.values()])
        epoch = datetime.date(1970, 1, 1)
        d0 = (datetime.date(*date_from) - epoch).days
        d1 = (datetime.date(*date_to) - epoch).days
        h = C.c_void_p()
        _check(lib().nigiri_amd_tt_load_gtfs(names, texts, len(files), d0, d1, tz_dir.encode(),
                                             C.byref(h)))
        return cls(h)

    def lower_bounds(self, seeds, direction=0):
        """Lower bounds (nigiri_amd_tt_lower_bounds) for seeds [(location, minutes)]."""
        loc = np.array([l for l, _ in seeds], np.uint32)
        mins = np.array([d for _, d in seeds], np.uint16)
        out = np.zeros(self.n_locations, np.uint16)
        _check(lib().nigiri_amd_tt_lower_bounds(self.handle, direction, len(seeds), ptr(loc, C.c_uint32),
                                                ptr(mins, C.c_uint16), ptr(out, C.c_uint16)))
        return out

    def location(self, stop_id):
        """Location index of a GTFS stop_id (loaded feeds only)."""
        out = C.c_uint32()
        _check(lib().nigiri_amd_tt_find_location(self.handle, stop_id.encode(), C.byref(out)))
        return out.value

    def location_id(self, l):
        buf = C.create_string_buffer(256)
        _check(lib().nigiri_amd_tt_location_id(self.handle, l, buf, 256))
        return buf.value.decode()

    def trip_id(self, t):
        """GTFS trip_id of transport t (loaded feeds only)."""
        buf = C.create_string_buffer(256)
        _check(lib().nigiri_amd_tt_transport_trip_id(self.handle, t, buf, 256))
        return buf.value.decode()

    def array(self, name):
        dtype, n = _VIEW_ARRAYS[name]
        ctype = np.ctypeslib.as_ctypes_type(dtype)
        p = getattr(self.view, name)
        count = n(self.view)
        if count == 0:
            return np.zeros(0, dtype)
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(ctype)), shape=(count,)).copy()

    @property
    def n_locations(self):
        return self.view.n_locations

    def close(self):
        if self.handle:
            lib().nigiri_amd_tt_free(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Queries:
    """A contiguous array of nigiri_query structs plus the offset storage they point into."""

    def __init__(self, n):
        self.q = (Query * max(n, 1))()
        self.n = n
        self._keep = []

    def __len__(self):
        return self.n

    def __getitem__(self, i):
        return self.q[i]

    @classmethod
    def generate(cls, tt, cfg, n, seed=None):
        seed = 0x7172 + cfg if seed is None else seed
        need = C.c_uint32(0)
        _check(lib().nigiri_amd_queries_generate(tt.handle, cfg, seed, n, None, None, None,
                                                 C.byref(need)))
        self = cls(n)
        loc = np.zeros(max(need.value, 1), np.uint32)
        dur = np.zeros(max(need.value, 1), np.int16)
        self._keep += [loc, dur]
        _check(lib().nigiri_amd_queries_generate(tt.handle, cfg, seed, n, self.q,
                                                 ptr(loc, C.c_uint32), ptr(dur, C.c_int16),
                                                 C.byref(need)))
        return self

    @classmethod
    def build(cls, specs):
        """specs: list of dicts with keys start=[(loc, dur)], dest=[(loc, dur)] and any Query field."""
        self = cls(len(specs))
        for i, s in enumerate(specs):
            q = self.q[i]
            q.dir = s.get("dir", DIR_FWD)
            q.start_match_mode = s.get("start_match_mode", MATCH_EXACT)
            q.dest_match_mode = s.get("dest_match_mode", MATCH_EXACT)
            q.use_start_footpaths = int(s.get("use_start_footpaths", 0))
            q.max_start_offset = s.get("max_start_offset", 7200)
            q.max_travel_time = s.get("max_travel_time", 7200)
            q.max_transfers = s.get("max_transfers", 14)
            q.tts_default = int(s.get("tts_default", 1))
            q.tts_min_transfer_time = s.get("tts_min_transfer_time", 0)
            q.tts_additional_time = s.get("tts_additional_time", 0)
            q.tts_factor = s.get("tts_factor", 1.0)
            q.fastest_direct_factor = s.get("fastest_direct_factor", 1.0)
            q.min_connection_count = s.get("min_connection_count", 0)
            q.extend_interval_earlier = int(s.get("extend_interval_earlier", 0))
            q.extend_interval_later = int(s.get("extend_interval_later", 0))
            if "fastest_direct" in s:
                q.has_fastest_direct = 1
                q.fastest_direct = s["fastest_direct"]
            if "max_interval" in s:
                q.has_max_interval = 1
                q.max_interval_from, q.max_interval_to = s["max_interval"]
            if "timeout_s" in s:
                q.has_timeout = 1
                q.timeout_s = s["timeout_s"]
            if "interval" in s:
                q.is_interval = 1
                q.interval_from, q.interval_to = s["interval"]
            else:
                q.start_time = s["time"]
            for which in ("start", "dest"):
                offs = s[which]
                loc = np.array([o[0] for o in offs], np.uint32)
                dur = np.array([o[1] for o in offs], np.int16)
                self._keep += [loc, dur]
                setattr(q, "n_" + which, len(offs))
                setattr(q, which + "_loc", ptr(loc, C.c_uint32))
                setattr(q, which + "_dur", ptr(dur, C.c_int16))
        return self

    def subset(self, idx):
        out = Queries(len(idx))
        out._keep = [self]
        for j, i in enumerate(idx):
            out.q[j] = self.q[i]
        return out


# device objects still open at interpreter exit are closed by an exit hook,
# engines before images, while the HIP runtime is still whole (objects left
# to garbage collection at exit would be released in no particular order,
# possibly after a profiler or the runtime has shut down)
_open_engines = weakref.WeakSet()
_open_images = weakref.WeakSet()


@atexit.register
def _close_open_objects():
    for e in list(_open_engines):
        e.close()
    for i in list(_open_images):
        i.close()


class Image:
    """Device-resident timetable (one per GPU)."""

    def lower_bounds(self, seeds, direction, lb_ptr, noalloc=False):
        """Lower bounds for seeds [(location, minutes)] into the device buffer at
        lb_ptr (n_locations uint16): nigiri_amd_lower_bounds, or its variant that
        allocates nothing (noalloc=True)."""
        loc = np.array([l for l, _ in seeds], np.uint32)
        mins = np.array([d for _, d in seeds], np.uint16)
        fn = lib().nigiri_amd_lower_bounds_noalloc if noalloc else lib().nigiri_amd_lower_bounds
        _check(fn(self.handle, direction, len(seeds), ptr(loc, C.c_uint32), ptr(mins, C.c_uint16),
                  C.c_void_p(lb_ptr)))

    def fastest_direct(self, queries):
        """get_fastest_direct of every query on the device: minutes (32767:
        none)."""
        out = np.zeros(max(len(queries), 1), np.int16)
        _check(lib().nigiri_amd_fastest_direct(self.handle, queries.q, len(queries), ptr(out, C.c_int16)))
        return out[:len(queries)]

    def __init__(self, tt, device=0):
        self.tt = tt  # the image borrows the host timetable
        self.handle = C.c_void_p()
        _check(lib().nigiri_amd_image_create(tt.handle, device, C.byref(self.handle)))
        _open_images.add(self)

    def bytes(self):
        b = C.c_uint64()
        _check(lib().nigiri_amd_image_bytes(self.handle, C.byref(b)))
        return b.value

    def close(self):
        if self.handle:
            lib().nigiri_amd_image_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _collect(res, jr, n, lg=None):
    out = []
    for i in range(n):
        r = res[i]
        js, legs = [], []
        for x in range(r.n_journeys):
            j = jr[r.journey_offset + x]
            js.append((j.start_time, j.dest_time, j.dest, j.transfers))
            if lg is not None:
                legs.append([lg[j.leg_offset + y].as_tuple() for y in range(j.n_legs)])
        out.append({"status": r.status, "journeys": js, "legs": legs, "n_epochs": r.n_epochs,
                    "n_rounds": r.n_rounds, "base_day": r.base_day, "fastest_direct": r.fastest_direct,
                    "slot": r.slot,
                    "latency_ms": r.latency_ms, "queue_ms": r.queue_ms,
                    "interval_from": r.interval_from, "interval_to": r.interval_to})
    return out


class Engine:
    """Device query state: lanes=64 (batched) or 1 (one query per launch)."""

    def __init__(self, image, lanes=64, groups=1, max_k=16):
        self.image = image
        self.lanes, self.groups, self.max_k = lanes, groups, max_k
        self.handle = C.c_void_p()
        _check(lib().nigiri_amd_engine_create(image.handle, lanes, groups, max_k,
                                              C.byref(self.handle)))
        _open_engines.add(self)

    @property
    def slots(self):
        return self.lanes * self.groups

    def state_bytes(self):
        b = C.c_uint64()
        _check(lib().nigiri_amd_engine_state_bytes(self.handle, C.byref(b)))
        return b.value

    def configure(self, admission_wait_ms=2.0, grid_waves=0, graphs=True, admit_when_idle=False,
                  state_budget_bytes=0, **knobs):
        """Sets the engine configuration; launch-shape knobs (route_claim,
        transfer_claim, pack_max_entries, transfer_pack_max_entries,
        loose_lanes) left out keep their values."""
        c = EngineConfig()
        c.admission_wait_ms = admission_wait_ms
        c.grid_waves = grid_waves
        c.eager_rounds = 0 if graphs else 1
        c.admit_when_idle = 1 if admit_when_idle else 0
        c.state_budget_bytes = state_budget_bytes
        for k, v in knobs.items():
            if k not in TUNE_KNOBS[1:]:
                raise TypeError(f"unknown engine knob {k}")
            setattr(c, k, v)
        _check(lib().nigiri_amd_engine_configure(self.handle, C.byref(c)))

    def get_config(self):
        c = EngineConfig()
        _check(lib().nigiri_amd_engine_get_config(self.handle, C.byref(c)))
        return {n: getattr(c, n) for n, _ in EngineConfig._fields_}

    def autotune(self, queries, repeats=1):
        """Chooses the launch-shape knobs on a sample of queries; returns the
        chosen configuration and the trials [(knob, value, ms)]."""
        cfg = EngineConfig()
        trials = (TuneTrial * 64)()
        n_trials = C.c_uint32()
        _check(lib().nigiri_amd_engine_autotune(self.handle, queries.q, len(queries), repeats,
                                                C.byref(cfg), trials, 64, C.byref(n_trials)))
        chosen = {n: getattr(cfg, n) for n in TUNE_KNOBS}
        return chosen, [(TUNE_KNOBS[t.knob], t.value, round(t.ms, 3))
                        for t in trials[:min(64, n_trials.value)]]

    def profile(self, enable=True):
        _check(lib().nigiri_amd_engine_profile(self.handle, int(enable)))

    def stats(self, reset=False):
        st = EngineStats()
        _check(lib().nigiri_amd_engine_stats(self.handle, C.byref(st), int(reset)))
        d = {n: getattr(st, n) for n, _ in EngineStats._fields_ if n != "work"}
        d["work"] = {n: st.work[i] for i, n in enumerate(WORK_NAMES)}
        return d

    def search(self, queries, journey_capacity=None, legs=True, leg_capacity=None):
        """Runs the queries; each result holds its journeys and (legs=True) their legs as
        Leg.as_tuple() lists.  legs="raw" has the library write the legs but leaves them in
        the ctypes array (returned as the "legs" entry of the first result)."""
        n = len(queries)
        cap = journey_capacity or max(64, 32 * n)
        res = (Result * max(n, 1))()
        jr = (Journey * cap)()
        leg_cap = (leg_capacity or 24 * cap) if legs else 0
        lg = (Leg * max(leg_cap, 1))() if legs else None
        _check(lib().nigiri_amd_search(self.handle, queries.q, n, res, jr, cap, lg, leg_cap))
        return _collect(res, jr, n, lg if legs is True else None)

    def serve(self, queries, arrival_ms, journey_capacity=None, legs=True, leg_capacity=None):
        n = len(queries)
        cap = journey_capacity or max(64, 32 * n)
        res = (Result * max(n, 1))()
        jr = (Journey * cap)()
        leg_cap = (leg_capacity or 24 * cap) if legs else 0
        lg = (Leg * max(leg_cap, 1))() if legs else None
        arr = np.ascontiguousarray(arrival_ms, np.float64)
        wall = C.c_double()
        _check(lib().nigiri_amd_serve(self.handle, queries.q, ptr(arr, C.c_double), n, res, jr,
                                      cap, lg, leg_cap, C.byref(wall)))
        return _collect(res, jr, n, lg if legs is True else None), wall.value

    def pong(self, query, ping_slot=0, journey_capacity=4096, leg_capacity=1 << 15):
        """Range search of one query by pong (nigiri's pong_search) on a plain engine
        (lanes == 1); query is a Query struct.  Returns the result dict of search()."""
        res = (Result * 1)()
        jr = (Journey * journey_capacity)()
        lg = (Leg * leg_capacity)()
        _check(lib().nigiri_amd_pong(self.handle, C.byref(query), ping_slot, res, jr,
                                     journey_capacity, lg, leg_capacity))
        out = _collect(res, jr, 1, lg)[0]
        out["interval"] = (res[0].interval_from, res[0].interval_to)
        return out

    def reconstruct(self, slot, query, journey, leg_capacity=64):
        """Legs of journey (start, dest_time, dest, transfers) from the slot's state as its
        last epoch left it (nigiri's raptor::reconstruct); query is a Query struct."""
        j = Journey(journey[0], journey[1], journey[2], journey[3])
        lg = (Leg * leg_capacity)()
        _check(lib().nigiri_amd_algo_reconstruct(self.handle, slot, C.byref(query), C.byref(j),
                                                 lg, leg_capacity))
        return [lg[i].as_tuple() for i in range(j.n_legs)]

    def download(self, slot):
        L = self.image.tt.n_locations
        rt = np.zeros((self.max_k, L), np.int16)
        best = np.zeros(L, np.int16)
        tmp = np.zeros(L, np.int16)
        tad = np.zeros(16, np.int16)
        lb = np.zeros(L, np.uint16)
        _check(lib().nigiri_amd_algo_download(self.handle, slot, ptr(rt, C.c_int16),
                                              ptr(best, C.c_int16), ptr(tmp, C.c_int16),
                                              ptr(tad, C.c_int16), ptr(lb, C.c_uint16)))
        return {"round_times": rt, "best": best, "tmp": tmp, "time_at_dest": tad, "lb": lb}

    def close(self):
        if self.handle:
            lib().nigiri_amd_engine_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
