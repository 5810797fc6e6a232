// Device-side data structures shared by the kernels and the engine host code.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace nigiri_amd {

constexpr int kRowsMax = 16;          // round_times rows (kMaxTransfers + 2)
constexpr std::uint32_t kMaxDestBits = 64U;  // destination stations per query
constexpr std::uint32_t kMaxDist = 256U;     // intermodal egress offsets
// relaxations into a query's destinations per round (its destinations'
// transfers and incoming footpaths, or its egress offsets; checked when a
// query is prepared), and time_at_dest changes per phase (at most one per
// such relaxation)
constexpr std::uint32_t kMaxDestEvents = 1024U;
constexpr std::uint32_t kMaxChanges = kMaxDestEvents;
constexpr std::uint32_t kPosTransfer = 0U << 30;
constexpr std::uint32_t kPosIntermodal = 1U << 30;
constexpr std::uint32_t kPosFootpath = 2U << 30;
constexpr std::uint16_t kUnreachableLb = 0xFFFFU;

// work counters: units of the algorithmic byte model
enum : int {
  kWorkRouteStops = 0,     // update_route stop iterations (per query)
  kWorkEvents = 1,         // events examined by get_earliest_transport
  kWorkRouteDayTests = 2,  // route traffic-day tests
  kWorkTransportTests = 3, // transport traffic-day tests
  kWorkTmpWrites = 4,      // tmp improvements in the route phase
  kWorkTransferStations = 5,
  kWorkFootpaths = 6,
  kWorkMarkedStations = 7,
  kWorkLocationRoutes = 8,
  kWorkWaveStops = 9,      // wavefront stop iterations of the batched route scan
  kWorkTransferItems = 10,  // (group, station) items of the batched transfer kernel
  kWorkHazardItems = 11,    // (group, station) items of the hazard replay
  kWorkHazardLanes = 12,    // queries replayed there
  kWorkLooseWaveStops = 13,  // the part of kWorkWaveStops spent on loose route items
  kWorkLooseItems = 14,      // loose (route, query) items scanned
  kWorkSkippedStops = 15,    // the part of kWorkRouteStops before a scan's start (not read)
  kWorkCount = 16
};

// A relaxation into a destination of one query in the transfer phase, at
// its position in nigiri's order of relaxations.
struct dest_event {
  std::uint32_t pos;
  std::uint32_t target;
  std::int16_t x;
  std::uint8_t kind;  // 0 transfer/footpath, 1 intermodal egress
};

// Downward sweep of PHAST for one search direction (the contraction
// hierarchy of the lower-bound graph): locations in level order, each with
// its edges from higher-ranked locations.
struct dev_sweep {
  std::uint32_t n_nodes;
  std::uint32_t const* order;
  std::uint32_t const* edge_off;  // [n_nodes + 1]
  std::uint32_t const* src;
  std::uint16_t const* w;         // minutes, at most 7201
  std::uint32_t const* level_off;  // [n_levels + 1] into order
};

// Read-only timetable image in HBM (one allocation per array).
struct dev_image {
  std::uint32_t L, R, T, NB;
  std::uint32_t const* route_stop_offsets;
  std::uint32_t const* route_stops;
  std::uint32_t const* route_transport_from;
  std::uint32_t const* route_stop_time_from;
  std::uint16_t const* event_times;
  std::uint32_t const* transport_traffic_days;
  std::uint32_t const* route_traffic_days;
  std::uint64_t const* bitfields;
  std::uint32_t const* location_route_offsets;
  std::uint32_t const* location_routes;
  // work order of the batched route phase: routes sorted by the position of
  // their stops (route_work_order), the rank of each location_routes entry's
  // route in it (lr_rank) and the route of each rank (rank_route)
  std::uint32_t const* lr_rank;
  std::uint32_t const* rank_route;
  std::uint8_t const* transfer_time;
  std::uint32_t const* fp_out_offsets;
  std::uint32_t const* fp_out;
  std::uint32_t const* fp_in_offsets;
  std::uint32_t const* fp_in;
  std::uint32_t const* fp_in_to_out;  // in-edge index -> out-edge index
  std::uint32_t const* fp_out_to_in;  // out-edge index -> in-edge index
  // the stops of every location's routes: for location l, in the order of
  // location_routes[l] and then of the stops, each (route, stop index) where
  // the route visits l (journey reconstruction, reconstruct.cc:456-519)
  std::uint32_t const* lrs_offsets;
  std::uint32_t const* lrs_route;
  std::uint16_t const* lrs_stop;
  // lower-bound graphs (loader/build_lb_graph.h), packed like footpaths
  std::uint32_t const* lb_fwd_offsets;
  std::uint32_t const* lb_fwd;
  std::uint32_t const* lb_bwd_offsets;
  std::uint32_t const* lb_bwd;
  dev_sweep sweep[2];  // [0]: forward search's lower bounds, [1]: backward
};

// Per-slot parameters written by the host at admission.
struct slot_param {
  std::int16_t tts_min, tts_add;
  float tts_factor;
  std::uint16_t base;
  std::uint16_t n_dest;  // non-intermodal destination bits
  std::uint16_t n_dist;  // intermodal egress offsets
  std::uint8_t end_k;
  std::uint8_t tts_default;
  std::uint8_t intermodal;
  std::uint8_t bounds_last_k;  // plain engine: pong bounds (0 = none)
};

// Device query state of an engine.  `W` lanes share one group: 64 for the
// batched engine (one query per lane of a wavefront), 1 for the plain
// single-query engine.  Layouts:
//   round_times [G][K][L][W], best/tmp/lb [G][L][W]  (int16 / uint16)
//   station marks / dest masks: W == 64 -> one uint64 lane mask per (g, l);
//                               W == 1  -> bit l % 64 of word g*LW + l/64
//   route marks likewise over R.
// 64-bit words covering n flag bytes (flag arrays are allocated and read in
// whole words)
inline constexpr std::size_t flag_words(std::size_t n) { return (n + 7U) / 8U; }

struct dev_state {
  std::uint32_t W, G, K, L, R;
  std::uint32_t LW, RW;  // words per group for W == 1 bitsets
  std::int16_t* round_times;
  std::int16_t* bounds;    // plain engine: [G][K][L], the other search's
                           // best round times of rounds 0..k (else null)
  std::int16_t* best;
  std::int32_t* tmp;       // int32 so that route scans update it with
                           // non-returning atomicMin / atomicMax
  std::uint16_t* lb;
  std::uint64_t* mark_t;   // marks of the transfer phase / starts
  std::uint64_t* mark_r;   // marks of the route phase
  std::uint64_t* route_mark;  // plain engine only
  std::uint64_t* dest_mask;
  std::uint64_t* grp_active;
  std::uint8_t* any_route;  // per slot: a marked station had routes
  std::uint8_t* any_tmp;    // per slot: the route phase improved tmp
  // work lists: packed (g << 32 | index)
  std::uint64_t* list_t;  // stations marked in mark_t
  std::uint64_t* list_r;  // stations marked in mark_r
  std::uint64_t* list_route;
  std::uint64_t* list_hz;  // stations with hazard lanes (g << 32 | location)
  std::uint64_t* hz_mask;  // hazard lanes per station (same layout as marks)
  // batched engine (W == 64): one flag byte per (group, location / route),
  // stored (no atomic) next to the lane-mask words and compacted into the
  // work lists
  std::uint8_t* flag_t;
  std::uint8_t* flag_r;
  std::uint8_t* flag_route;
  std::uint8_t* flag_hz;
  // batched engine: route-phase marks, one byte per (g, l, lane) [G][L][64]
  std::uint8_t* mark_rb;
  // batched engine: lanes that may hold a label, so that the best sweep and
  // the lane resets touch only those.  rt_valid [G][K][L]: lanes whose
  // round_times[k][l] may be valid; bt_valid [G][L]: lanes whose best[l] or
  // tmp[l] may be valid, apart from the current round's station marks
  // (mark_t, mark_rb), which are folded in when they are cleared.  A bit is
  // set wherever a label is written and cleared only by a reset.
  std::uint64_t* rt_valid;
  std::uint64_t* bt_valid;
  // batched engine: packed route wavefronts (route << 32 | start << 16 |
  // n_lanes; packed transfer wavefronts: location << 32 | n_lanes) and
  // their lanes (g << 6 | lane, 64 per wavefront); start is the scan index
  // of the route's first stop marked for one of the wavefront's queries
  std::uint64_t* packed_info;
  std::uint16_t* packed_lanes;
  std::uint32_t* counts;   // [0] list_t, [1] list_r, [2] list_route, [3] hz, [4] overflow flags,
                           // [5] packed route wavefronts, [6] / [7] work cursors of
                           // the route scan / transfer kernels, [8] loose route
                           // items, [9] their cursor (batched engine); 16 entries
  // batched engine: loose route items (route << 32 | start << 16 | g << 6 |
  // lane; g < 1024): the queries of packed wavefronts too sparse to share a
  // route's staging, scanned one (route, query) per lane
  std::uint64_t* loose;
  std::size_t cap_loose;
  // algorithmic work counters (see kWork* indices)
  unsigned long long* work;
  std::size_t cap_t, cap_r, cap_route, cap_hz;
  std::size_t cap_pack;  // packed route / transfer wavefronts (batched engine)
  // per slot
  slot_param* param;
  std::uint8_t* kslot;     // [S] round of the slot's running epoch
  std::int16_t* tad;       // [S][16]
  dest_event* dest_ev;     // [S][kMaxDestEvents] destination relaxations of the phase
  std::int16_t* tad0;      // [S]
  std::uint32_t* n_chg;    // [S]
  std::uint32_t* chg_pos;  // [S][kMaxChanges]
  std::int16_t* chg_tad;   // [S][kMaxChanges]
  std::uint32_t* dest;     // [S][kMaxDestBits]
  std::uint32_t* dist_loc; // [S][kMaxDist]
  std::uint16_t* dist_min; // [S][kMaxDist]
  std::int16_t* dest_rows; // [S][kMaxDestBits][K]
};

}  // namespace nigiri_amd
