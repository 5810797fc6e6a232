/*
 * nigiri_amd.h — C ABI of the MI355X RAPTOR library (libnigiri_amd.so).
 *
 * Plain C: pointers, sizes and status codes only.  No exception, no C++ or
 * torch type crosses this boundary.  Every function returns NIGIRI_AMD_OK (0)
 * or a negative status; nigiri_amd_last_error() gives the message of the
 * last failure on the calling thread.
 *
 * Times crossing the boundary are minutes since the unix epoch (int32),
 * exactly nigiri's unixtime_t (include/nigiri/types.h:360).  Durations are
 * minutes (int16, nigiri's duration_t, types.h:359).  Location indices follow
 * nigiri's location_idx_t, including the 9 special stations at 0..8
 * (include/nigiri/special_stations.h:7-22).
 */
#ifndef NIGIRI_AMD_H
#define NIGIRI_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NIGIRI_AMD_OK 0
#define NIGIRI_AMD_EINVAL (-1)   /* bad argument / malformed input          */
#define NIGIRI_AMD_ENOMEM (-2)   /* device or host allocation failed        */
#define NIGIRI_AMD_EDEVICE (-3)  /* HIP runtime error                        */
#define NIGIRI_AMD_ECAPACITY (-4) /* request exceeds slots / state memory   */
#define NIGIRI_AMD_EUNSUPPORTED (-5) /* query feature not supported here    */
#define NIGIRI_AMD_ERECONSTRUCT (-6) /* no legs for the journey: nigiri's
                                        reconstruct throws (reconstruct.cc:
                                        159, 738-741, 794-796)              */

#define NIGIRI_AMD_DIR_FWD 0
#define NIGIRI_AMD_DIR_BWD 1

#define NIGIRI_AMD_MATCH_EXACT 0      /* location_match_mode::kExact      */
#define NIGIRI_AMD_MATCH_INTERMODAL 1 /* location_match_mode::kIntermodal */

/* -------------------------------------------------------------------------
 * Timetable view: the arrays of nigiri::timetable the RAPTOR path reads
 * (include/nigiri/timetable.h:342-398, 89-98, 421-422), flattened to CSR.
 * All pointers are borrowed; the view never owns memory.
 * ---------------------------------------------------------------------- */
typedef struct nigiri_tt_view {
  uint32_t n_locations;  /* including the 9 special stations               */
  uint32_t n_routes;
  uint32_t n_transports;
  uint32_t n_bitfields;  /* each bitfield is 8 x uint64 (bitset<512>)      */
  uint32_t n_route_stops;    /* size of route_stops                         */
  uint32_t n_event_times;    /* size of event_times                         */
  uint32_t n_location_routes;
  uint32_t n_footpaths;      /* size of fp_out == size of fp_in             */
  uint32_t n_lb_fwd, n_lb_bwd;
  /* calendar: nigiri's date_range_ and internal interval (timetable.h:232-252)
   * internal day 0 starts at internal_from (= date_range.from - 5 days).   */
  int32_t date_range_from; /* unix minutes, midnight                        */
  int32_t date_range_to;   /* unix minutes, midnight (exclusive)            */
  int32_t internal_from;   /* unix minutes of internal day 0                */
  uint32_t n_days;         /* internal_interval_days().size()               */

  /* T1 route_location_seq_ (vecvec<route, stop>): packed stop words
   *   bits 0..27 location, 28 in_allowed, 29 out_allowed,
   *   30 in_allowed_wheelchair, 31 out_allowed_wheelchair (stop.h:85-89)   */
  const uint32_t* route_stop_offsets;  /* [n_routes + 1]                   */
  const uint32_t* route_stops;         /* [n_route_stops]                  */
  /* T2 route_transport_ranges_, route_stop_time_ranges_, route_stop_times_
   *   transports of route r are [route_transport_from[r], ...[r+1])
   *   event block of route r starts at route_stop_time_from[r]; stop-event
   *   block j (= 2*stop_idx - is_arrival) holds one delta per transport.
   *   delta packing (types.h:630-631): bits 0..4 days, bits 5..15 mam.     */
  const uint32_t* route_transport_from; /* [n_routes + 1]                  */
  const uint32_t* route_stop_time_from; /* [n_routes + 1]                  */
  const uint16_t* event_times;          /* [n_event_times]                 */
  /* T3 traffic days: bitfield index per transport / route, 512-bit sets    */
  const uint32_t* transport_traffic_days; /* [n_transports]                */
  const uint32_t* route_traffic_days;     /* [n_routes]                    */
  const uint64_t* bitfields;              /* [n_bitfields * 8]             */
  /* T4 location_routes_                                                    */
  const uint32_t* location_route_offsets; /* [n_locations + 1]             */
  const uint32_t* location_routes;        /* [n_location_routes]           */
  /* T5 transfer_time_ (u8 minutes) and footpaths_out_/in_[profile 0];
   *   footpath packing (footpath.h:65-66): bits 0..22 target, 23..31 min.  */
  const uint8_t* transfer_time;           /* [n_locations]                 */
  const uint32_t* fp_out_offsets;         /* [n_locations + 1]             */
  const uint32_t* fp_out;                 /* [n_footpaths]                 */
  const uint32_t* fp_in_offsets;          /* [n_locations + 1]             */
  const uint32_t* fp_in;                  /* [n_footpaths]                 */
  /* lower-bound graphs fwd_/bwd_search_lb_graph_ (timetable.h:421-422),
   * same footpath packing                                                  */
  const uint32_t* lb_fwd_offsets;         /* [n_locations + 1]             */
  const uint32_t* lb_fwd;                 /* [n_lb_fwd]                    */
  const uint32_t* lb_bwd_offsets;         /* [n_locations + 1]             */
  const uint32_t* lb_bwd;                 /* [n_lb_bwd]                    */
  /* planar stop coordinates in metres (synthetic timetables only; used to
   * draw intermodal offsets).  May be NULL.                                */
  const float* coord_x;                   /* [n_locations]                 */
  const float* coord_y;                   /* [n_locations]                 */
} nigiri_tt_view;

/* -------------------------------------------------------------------------
 * Query: the subset of nigiri::routing::query (include/nigiri/routing/
 * query.h:70-101) that the static RAPTOR path supports (no vias, no
 * td offsets, no real-time).  Offsets are (location, duration) pairs.
 * ---------------------------------------------------------------------- */
typedef struct nigiri_query {
  int32_t dir;               /* NIGIRI_AMD_DIR_FWD / _BWD                    */
  int32_t start_match_mode;  /* NIGIRI_AMD_MATCH_*                           */
  int32_t dest_match_mode;
  int32_t use_start_footpaths;
  int32_t is_interval;       /* 0: start_time_ is a point (ontrip)           */
  int32_t start_time;        /* unix minutes (ontrip)                        */
  int32_t interval_from;     /* unix minutes, [from, to) (pretrip)           */
  int32_t interval_to;
  uint32_t n_start;
  uint32_t n_dest;
  const uint32_t* start_loc;
  const int16_t* start_dur;
  const uint32_t* dest_loc;
  const int16_t* dest_dur;
  int16_t max_start_offset;  /* default 7200 (kMaxTravelTime)               */
  int16_t max_travel_time;   /* default 7200                                */
  uint8_t max_transfers;     /* default 14 (kMaxTransfers)                  */
  uint8_t pad_[3];
  uint32_t min_connection_count; /* widens the interval (search.h:237-317) */
  int32_t extend_interval_earlier;
  int32_t extend_interval_later;
  int32_t prf_idx;               /* only 0 is supported                     */
  int32_t tts_default;           /* transfer_time_settings::default_        */
  int16_t tts_min_transfer_time;
  int16_t tts_additional_time;
  float tts_factor;
  int32_t has_fastest_direct;    /* query::fastest_direct_ (optional)       */
  int16_t fastest_direct;
  int16_t pad2_;
  double fastest_direct_factor;  /* default 1.0                             */
  /* query::max_interval_ (query.h:89, optional): the search interval, as
   * widened for min_connection_count, is clamped into [from, to)
   * (interval_estimate.h:132-145)                                           */
  int32_t has_max_interval;
  int32_t max_interval_from;     /* unix minutes                            */
  int32_t max_interval_to;
  /* the timeout argument of raptor_search / pong_search (search.h:179,
   * 227-244; pong.cc:68, 235-245), optional, in seconds: once it has passed
   * since the first start time began, the interval is not extended further
   * (search) or no further ping starts (pong)                               */
  int32_t has_timeout;
  int32_t timeout_s;
  int32_t pad3_;
} nigiri_query;

/* A journey of the result Pareto set (journey.h:104-111).  Its legs, in
 * travel order, are legs[leg_offset .. leg_offset + n_legs) of the legs
 * array passed with it (n_legs = 0 when no legs array was given). */
typedef struct nigiri_journey {
  int32_t start_time;  /* unix minutes */
  int32_t dest_time;   /* unix minutes */
  uint32_t dest;       /* location index (END = 1 for intermodal)           */
  uint8_t transfers;
  uint8_t pad_[3];
  uint32_t leg_offset;
  uint32_t n_legs;
} nigiri_journey;

/* journey::leg (journey.h:44-75); `kind` names the alternative of uses_. */
#define NIGIRI_AMD_LEG_RUN 0      /* journey::run_enter_exit               */
#define NIGIRI_AMD_LEG_FOOTPATH 1 /* footpath                              */
#define NIGIRI_AMD_LEG_OFFSET 2   /* routing::offset (intermodal)          */
typedef struct nigiri_leg {
  int32_t kind;
  uint32_t from, to;           /* locations (START = 0, END = 1)          */
  int32_t dep_time, arr_time;  /* unix minutes                            */
  uint32_t transport;          /* RUN: transport index                    */
  uint32_t route;              /* RUN: route of the transport             */
  uint16_t day;                /* RUN: traffic day (day_idx_t)            */
  uint16_t stop_from;          /* RUN: stop range [stop_from, stop_to) of */
  uint16_t stop_to;            /*      the route                          */
  int16_t duration;            /* FOOTPATH / OFFSET: minutes              */
  uint32_t target;             /* FOOTPATH / OFFSET: target location      */
} nigiri_leg;

/* Per-query outcome of a batched search. */
typedef struct nigiri_result {
  int32_t status;          /* NIGIRI_AMD_OK or a negative status            */
  uint32_t n_journeys;     /* journeys written (<= journey_capacity)        */
  uint32_t journey_offset; /* index of the first journey in the out array   */
  uint32_t n_epochs;       /* start-time groups executed                    */
  uint32_t n_rounds;       /* RAPTOR rounds executed, summed over epochs    */
  int32_t slot;            /* engine slot that ran the query (-1: none);
                              its state stays readable with
                              nigiri_amd_algo_download until the slot is
                              reused                                        */
  int32_t interval_from;   /* search interval actually searched             */
  int32_t interval_to;
  int16_t fastest_direct;
  int16_t base_day;
  double latency_ms;       /* arrival -> completion (served path)           */
  double queue_ms;         /* arrival -> first round on the device (served
                              path: preparation, lower bounds, free slot)   */
} nigiri_result;

typedef struct nigiri_amd_tt nigiri_amd_tt;         /* host timetable       */
typedef struct nigiri_amd_image nigiri_amd_image;   /* device timetable     */
typedef struct nigiri_amd_engine nigiri_amd_engine; /* device query state   */

const char* nigiri_amd_last_error(void);
int nigiri_amd_version(void);

/* ---- host timetable --------------------------------------------------- */

/* Seeded synthetic timetable for workload `cfg` (1..5).  `scale` in (0, 1]
 * shrinks stop/route/transport counts for quick tests (1.0 = full size).
 * Replaces reading nigiri's tt.bin (src/timetable.cc:275-281) for the
 * benchmark workloads; invariants follow the loader (FIFO routes,
 * gtfs/load_timetable.cc:291-301; route traffic days, init_finish.cc:
 * 168-198; footpath order, build_footpaths.cc:404-414; lb graph,
 * loader/build_lb_graph.h:11-99). */
int nigiri_amd_tt_generate(int cfg, uint64_t seed, double scale,
                           nigiri_amd_tt** out);
/* Builds a timetable from caller arrays.  Only the route, transport,
 * bitfield, transfer-time and footpath arrays of `v` are read; location
 * routes, route traffic days, footpaths_in and the lb graphs are rebuilt as
 * nigiri's finalize() does (src/loader/init_finish.cc:200-234). */
int nigiri_amd_tt_from_arrays(const nigiri_tt_view* v, nigiri_amd_tt** out);
/* Loads a GTFS feed given as files in memory, as nigiri's loader::mem_dir
 * holds them (file_names[i] such as "stops.txt", file_contents[i] its text),
 * for the date range [date_from_day, date_to_day) in days since 1970-01-01
 * (timetable::date_range_), the way nigiri's GTFS loader and finalize() do
 * (replaces loader::gtfs::load_timetable, src/loader/gtfs/load_timetable.cc:
 * 84-545, and loader::finalize, src/loader/init_finish.cc:200-234).
 * tz_dir: directory of TZif time-zone files (NULL or "": $TZDIR, else
 * /usr/share/zoneinfo).  Feeds with frequencies, stay-seated transfers,
 * block_id trips, flex or stop times that need interpolation are refused
 * (NIGIRI_AMD_EUNSUPPORTED).  Footpaths come from transfers.txt only. */
int nigiri_amd_tt_load_gtfs(const char* const* file_names,
                            const char* const* file_contents, uint32_t n_files,
                            int32_t date_from_day, int32_t date_to_day,
                            const char* tz_dir, nigiri_amd_tt** out);
/* Feed ids of a loaded feed: the stop_id of location l (special stations:
 * their nigiri names, special_stations.h:7-22) and the trip_id of transport
 * t, NUL-terminated into buf[0 .. cap); the location of a stop_id
 * (timetable::locations_.location_id_to_idx_).  NIGIRI_AMD_EINVAL for a
 * timetable without ids or an unknown id. */
int nigiri_amd_tt_location_id(const nigiri_amd_tt* tt, uint32_t l, char* buf,
                              uint32_t cap);
int nigiri_amd_tt_find_location(const nigiri_amd_tt* tt, const char* stop_id,
                                uint32_t* l);
int nigiri_amd_tt_transport_trip_id(const nigiri_amd_tt* tt, uint32_t t,
                                    char* buf, uint32_t cap);
int nigiri_amd_tt_view(const nigiri_amd_tt* tt, nigiri_tt_view* out);

/* Lower bounds of a search on the host (replaces routing::dijkstra,
 * src/routing/dijkstra.cc:18-67, as search::init calls it, search.h:
 * 114-131): the shortest travel time from every location to the seeds
 * (location, minutes: the destination offsets) on the lower-bound graph of
 * direction `dir`, 0xFFFF where it is 7201 minutes or more.  Computed with
 * the timetable's contraction hierarchy (upward search, then a sweep over
 * all locations); the values equal the dial-bucket Dijkstra's.
 * lb: [n_locations]. */
int nigiri_amd_tt_lower_bounds(const nigiri_amd_tt* tt, int32_t dir,
                               uint32_t n_seeds, const uint32_t* seed_loc,
                               const uint16_t* seed_min, uint16_t* lb);
void nigiri_amd_tt_free(nigiri_amd_tt* tt);

/* Seeded query workload for `cfg` (point / range / intermodal mix as the
 * workload defines).  Offsets are written into caller arrays sized by
 * nigiri_amd_queries_generate(...) with out == NULL (returns the number of
 * offset entries needed in *n_offsets). */
int nigiri_amd_queries_generate(const nigiri_amd_tt* tt, int cfg,
                                uint64_t seed, uint32_t n_queries,
                                nigiri_query* out, uint32_t* offset_loc,
                                int16_t* offset_dur, uint32_t* n_offsets);

/* ---- device image (replaces gpu_timetable, gpu/raptor.h:27-33) -------- */

int nigiri_amd_image_create(const nigiri_amd_tt* tt, int device,
                            nigiri_amd_image** out);
int nigiri_amd_image_bytes(const nigiri_amd_image* img, uint64_t* bytes);
void nigiri_amd_image_destroy(nigiri_amd_image* img);

/* Lower bounds of a search on the device (replaces routing::dijkstra,
 * src/routing/dijkstra.cc:18-67, as search::init calls it, search.h:
 * 114-131): lb[l] = shortest travel time from location l to the seeds
 * (location, minutes) on the lower-bound graph of `dir`, 0xFFFF where it is
 * 7201 minutes or more; the values equal the dial-bucket Dijkstra's.  An
 * upward search of the image's contraction hierarchy on the host, then the
 * PHAST sweep over all locations on the device.  lb: device memory,
 * [n_locations] uint16, written before the call returns. */
int nigiri_amd_lower_bounds(nigiri_amd_image* img, int32_t dir,
                            uint32_t n_seeds, const uint32_t* seed_loc,
                            const uint16_t* seed_min, uint16_t* lb);
/* Fastest direct connection of each query on the device (replaces
 * routing::get_fastest_direct, src/routing/get_fastest_direct.cc:12-100, as
 * search::search calls it, search.h:193): the shortest walk from a start
 * offset into a destination offset, following footpaths (footpaths_out_
 * forward, footpaths_in_ backward) only into destinations, with the query's
 * transfer-time settings as given, times fastest_direct_factor rounded up;
 * 32767 (kUnreachableDirect) when there is none.  A query that sets
 * has_fastest_direct gets its own value back.  out: [n] minutes (host).
 * NIGIRI_AMD_EINVAL for a location out of range, NIGIRI_AMD_ECAPACITY for a
 * query with more than 1024 start and destination locations.
 * nigiri_amd_search runs the same kernel for its queries. */
int nigiri_amd_fastest_direct(nigiri_amd_image* img, const nigiri_query* qs,
                              uint32_t n, int16_t* out);
/* nigiri_amd_lower_bounds without any device allocation.  No counterpart in
 * the reference: for callers that cannot allocate device memory (code under
 * graph capture set-up, tight serving loops, harnesses that forbid
 * hipMalloc).  The upward search's labels are staged in a static device
 * array of 6 MiB + 256 B per GPU (held until the process ends), so a call
 * takes at most 2^20 labels (NIGIRI_AMD_ECAPACITY beyond; the searches here
 * settle a few thousand).  All calls on one GPU share that array: calls that
 * overlap on one GPU race.  Same arguments and results as
 * nigiri_amd_lower_bounds. */
int nigiri_amd_lower_bounds_noalloc(nigiri_amd_image* img, int32_t dir,
                                    uint32_t n_seeds, const uint32_t* seed_loc,
                                    const uint16_t* seed_min, uint16_t* lb);

/* ---- engine: device query state (replaces gpu_raptor_state,
 * gpu/raptor.h:47-53, and raptor_state, raptor/raptor_state.h:22-111) ----
 *   lanes      : 64 for the batched engine (64 queries share a wavefront),
 *                1 for the plain one-query-per-launch engine
 *   n_groups   : number of 64-query groups (lanes == 64) or of independent
 *                single-query states (lanes == 1)
 *   max_k      : rows of round_times (end_k = min(max_transfers,14)+2 <=
 *                max_k <= 16)                                              */
int nigiri_amd_engine_create(nigiri_amd_image* img, uint32_t lanes,
                             uint32_t n_groups, uint32_t max_k,
                             nigiri_amd_engine** out);
int nigiri_amd_engine_state_bytes(const nigiri_amd_engine* e, uint64_t* b);
void nigiri_amd_engine_destroy(nigiri_amd_engine* e);

/* Full search of n queries (replaces routing::raptor_search,
 * include/nigiri/routing/raptor_search.h:9-17, and search<>::execute,
 * search.h:204-345), including the reconstruction of every journey's legs
 * (search.h:457-475; journeys without legs are dropped as nigiri drops
 * them, search.h:335-336).  Journeys of query i are written to
 * journeys[res[i].journey_offset ...], their legs to legs[journey.leg_offset
 * ...]; journey_capacity and leg_capacity bound the totals (legs may be
 * NULL with leg_capacity 0: legs are then not written).  Queries run
 * batched: up to lanes * n_groups at once. */
int nigiri_amd_search(nigiri_amd_engine* e, const nigiri_query* qs,
                      uint32_t n, nigiri_result* res, nigiri_journey* journeys,
                      uint32_t journey_capacity, nigiri_leg* legs,
                      uint32_t leg_capacity);

/* Range (or ontrip) search by pong (replaces routing::pong_search,
 * include/nigiri/routing/raptor/pong.h:22-30; src/routing/raptor/pong.cc:
 * 62-428) for one query on a plain engine (lanes == 1): ping searches on
 * slot ping_slot, pong searches on ping_slot + 1, bounded by the ping's
 * round times.  Journeys (start and destination times in travel direction;
 * dest is the pong's destination, i.e. the query's start location, as in
 * the reference) and their legs go to journeys / legs as in
 * nigiri_amd_search, res->n_epochs counts executes.  Not done: the leg
 * rewrite by earlier alternatives (pong.cc:430-470). */
int nigiri_amd_pong(nigiri_amd_engine* e, const nigiri_query* q,
                    uint32_t ping_slot, nigiri_result* res,
                    nigiri_journey* journeys, uint32_t journey_capacity,
                    nigiri_leg* legs, uint32_t leg_capacity);

/* ---- Algo-level entry points for one slot: the method set of
 * raptor<> (raptor.h:174-414) / gpu_raptor<> (gpu/raptor.h:83-99) that
 * search<Dir, Algo> drives.  `slot` < lanes * n_groups. ---------------- */
typedef struct nigiri_algo_init {
  int32_t dir;
  int32_t base_day;          /* day_idx_t base (search.h:158-164)           */
  uint32_t n_dest;           /* is_dest_ set bits (non-intermodal)          */
  const uint32_t* dest;      /* ascending                                   */
  int32_t intermodal;        /* dist_to_dest non-empty                      */
  uint32_t n_dist;           /* intermodal: reachable (location, minutes)   */
  const uint32_t* dist_loc;  /* ascending                                   */
  const uint16_t* dist_min;
  const uint16_t* lb;        /* [n_locations] travel-time lower bounds      */
  int32_t tts_default;
  int16_t tts_min_transfer_time;
  int16_t tts_additional_time;
  float tts_factor;
} nigiri_algo_init;

int nigiri_amd_algo_init(nigiri_amd_engine* e, uint32_t slot,
                         const nigiri_algo_init* init);
int nigiri_amd_algo_reset_arrivals(nigiri_amd_engine* e, uint32_t slot);
int nigiri_amd_algo_next_start_time(nigiri_amd_engine* e, uint32_t slot);
int nigiri_amd_algo_add_start(nigiri_amd_engine* e, uint32_t slot,
                              uint32_t location, int32_t unixtime);
/* Runs the rounds of one epoch; writes dest round times into
 * dest_times[n_dest_bits * max_k] (row k of dest j at [j*max_k + k]) as
 * deltas; journey extraction (raptor.h:393-413) happens in the caller. */
int nigiri_amd_algo_execute(nigiri_amd_engine* e, uint32_t slot,
                            int32_t start_time, uint8_t max_transfers,
                            int32_t worst_time_at_dest, int16_t* dest_times,
                            uint32_t* n_rounds);
/* Bounds of a pong search (raptor<>::fill_bounds / set_bounds,
 * raptor.h:117-172, with within_bounds, raptor.h:458-489).  fill_bounds
 * writes bounds rows 0 .. n_rows-1 of slot dst_slot from the round times of
 * slot src_slot as its last execute left them (row k: the best of rounds
 * 0..k); in the reference ping and pong share one state, here the pong's
 * slot holds the bounds.  set_bounds(slot, last_round) makes the slot's
 * next executes prune labels by bounds row last_round - k (0 switches
 * bounds off; an execute with more rounds than last_round is refused).
 * Plain engine only (lanes == 1): NIGIRI_AMD_EUNSUPPORTED otherwise. */
int nigiri_amd_algo_fill_bounds(nigiri_amd_engine* e, uint32_t src_slot,
                                uint32_t dst_slot, uint32_t n_rows);
int nigiri_amd_algo_set_bounds(nigiri_amd_engine* e, uint32_t slot,
                               uint32_t last_round);
/* Reconstructs the legs of journey *j (its start_time, dest_time, dest and
 * transfers) from the slot's state as the last execute left it (replaces
 * raptor<>::reconstruct, raptor.h:416-423, i.e. reconstruct_journey,
 * reconstruct.cc:879-902, for queries without vias).  q is the search's
 * query (its offsets, match modes and is_interval are read).  Legs go to
 * legs[0 .. j->n_legs); returns NIGIRI_AMD_ERECONSTRUCT where nigiri's
 * reconstruct throws. */
int nigiri_amd_algo_reconstruct(nigiri_amd_engine* e, uint32_t slot,
                                const nigiri_query* q, nigiri_journey* j,
                                nigiri_leg* legs, uint32_t leg_capacity);
/* Copies the slot's state to host arrays (any pointer may be NULL):
 * round_times[max_k][n_locations], best[n_locations], tmp[n_locations],
 * time_at_dest[max_k], lb[n_locations] (the lower bounds the search used;
 * the raptor reads them from the caller's lb vector, raptor.h:430-446). */
int nigiri_amd_algo_download(nigiri_amd_engine* e, uint32_t slot,
                             int16_t* round_times, int16_t* best,
                             int16_t* tmp, int16_t* time_at_dest,
                             uint16_t* lb);

/* ---- instrumentation (no counterpart in the reference) -------------------
 * Work counters of the algorithmic byte model, accumulated on the device
 * (units: update_route stop iterations, events examined, traffic-day tests,
 * ... as the oracle counts them), and, when profiling is enabled, HIP-event
 * timings of the route-scan and transfer kernels.                          */
typedef struct nigiri_engine_stats {
  uint64_t epochs;
  uint64_t rounds;
  uint64_t route_scan_launches;
  double route_scan_ms;
  uint64_t transfer_launches;
  double transfer_ms;
  uint64_t work[16];
  uint64_t reconstructed;   /* journeys handed to the reconstruction        */
  double reconstruct_ms;    /* host wall time of the reconstruction steps   */
  double legs_ms;           /*   of which writing the legs on the host      */
  double results_ms;        /* host wall time of the journey extraction     */
  uint64_t graph_rounds;    /* rounds replayed from captured hipGraphs      */
  uint64_t graph_captures;  /* round graphs captured (engine lifetime)      */
} nigiri_engine_stats;

int nigiri_amd_engine_profile(nigiri_amd_engine* e, int enable);
int nigiri_amd_engine_stats(nigiri_amd_engine* e, nigiri_engine_stats* out,
                            int reset);

/* ---- engine configuration (no counterpart in the reference) -------------
 * admission_wait_ms: serving loop; at an epoch boundary, wait up to this
 *   long for queries that have arrived but are still being prepared, so
 *   that free slots start the epoch filled (default 2 ms).
 * grid_waves: wavefronts per persistent-kernel launch (0 = 32 per CU).
 * eager_rounds: 0 = each round's kernel sequence is captured once as a
 *   hipGraph and replayed (default); 1 = kernels are launched one by one.
 * Launch-shape knobs (0 = keep the current value; results never depend on
 * them, only speed does; nigiri_amd_engine_autotune sets them):
 * route_claim: packed route wavefronts a wavefront claims per atomic (1).
 * transfer_claim: stations a wavefront claims per atomic in the transfer
 *   and footpath kernel (16).
 * pack_max_entries: (route, group) entries, i.e. 64-query groups, that one
 *   packed route wavefront may mix (1-64; 64).
 * transfer_pack_max_entries: (station, group) entries, i.e. 64-query
 *   groups, that one packed transfer wavefront may mix (1-64; 64).
 * loose_lanes: packed route wavefronts of fewer queries than this are
 *   scanned loose instead, one (route, query) per lane, so that sparse
 *   routes do not each hold a whole wavefront (1 = none loose; 1-33; 17).
 * admit_when_idle: serving loop; 1 = new queries are admitted only when no
 *   slot is busy, so that a batch starts together while the preparation and
 *   lower bounds of the next batch overlap the current batch's rounds
 *   (0 = a query takes any free slot at an epoch boundary; default).
 * state_budget_bytes: admission by device state memory.  A query takes
 *   n_locations * (2 * end_k + 9) bytes while it runs (end_k rows of int16
 *   round_times plus its best, tmp, lower-bound and mark columns) and is
 *   admitted only while the admitted queries' total stays within the
 *   budget; it waits otherwise.  A query that can never fit (more than the
 *   budget, or end_k above the engine's max_k) fails alone with
 *   NIGIRI_AMD_ECAPACITY.  0 = the memory of all slots at max_k rows.
 */
typedef struct nigiri_engine_config {
  double admission_wait_ms;
  uint32_t grid_waves;
  uint32_t eager_rounds;
  uint32_t route_claim;
  uint32_t transfer_claim;
  uint32_t pack_max_entries;
  uint32_t admit_when_idle;
  uint32_t transfer_pack_max_entries;
  uint64_t state_budget_bytes;
  uint32_t loose_lanes;
  uint32_t pad_;
} nigiri_engine_config;

int nigiri_amd_engine_configure(nigiri_amd_engine* e,
                                const nigiri_engine_config* c);
/* The engine's current configuration. */
int nigiri_amd_engine_get_config(const nigiri_amd_engine* e,
                                 nigiri_engine_config* out);

/* ---- autotuner (no counterpart in the reference) ------------------------
 * Chooses the launch-shape knobs of nigiri_engine_config (grid_waves,
 * route_claim, transfer_claim, pack_max_entries,
 * transfer_pack_max_entries, loose_lanes) for this engine and timetable on a sample of
 * queries (ideally one batch of the real
 * workload): one pass of coordinate descent in which every candidate value
 * of a knob runs the sample through the engine `repeats` times and the value
 * with the smallest median engine time is kept.  The chosen values stay set
 * on the engine and are written to *out (may be NULL).  Each trial is
 * reported as (knob, value, ms) in trials[0 .. min(n, trial_capacity)),
 * with the total count in *n_trials (may be NULL).  Knob ids: 0 grid_waves,
 * 1 route_claim, 2 transfer_claim, 3 pack_max_entries,
 * 4 transfer_pack_max_entries, 5 loose_lanes.
 * Results never depend on the knobs. */
typedef struct nigiri_tune_trial {
  uint32_t knob;
  uint32_t value;
  double ms;
} nigiri_tune_trial;

int nigiri_amd_engine_autotune(nigiri_amd_engine* e, const nigiri_query* qs,
                               uint32_t n, uint32_t repeats,
                               nigiri_engine_config* out,
                               nigiri_tune_trial* trials,
                               uint32_t trial_capacity, uint32_t* n_trials);

/* ---- serving loop ------------------------------------------------------ */
/* Runs a continuous-batching server over the engine: query i becomes
 * visible at arrival_ms[i] (relative to the start), is admitted when a slot
 * and state memory are free, and completes with res[i].latency_ms set.
 * Reports through res/journeys exactly like nigiri_amd_search. */
int nigiri_amd_serve(nigiri_amd_engine* e, const nigiri_query* qs,
                     const double* arrival_ms, uint32_t n,
                     nigiri_result* res, nigiri_journey* journeys,
                     uint32_t journey_capacity, nigiri_leg* legs,
                     uint32_t leg_capacity, double* wall_ms);

#ifdef __cplusplus
}
#endif

#endif /* NIGIRI_AMD_H */
