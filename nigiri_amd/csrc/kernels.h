// gfx950 round kernels of the RAPTOR engine.
//
// One RAPTOR round (raptor.h:218-387) for every active query slot is a
// fixed sequence of persistent, grid-stride kernels that read their work
// counts from device memory, so a round is enqueued without host syncs:
//
//   round_begin   time_at_dest from destination labels  (raptor.h:219-226)
//   best_sweep    best = min(round_times[k], best)       (only for slots in
//                 a later start-time epoch; earlier rows are empty otherwise)
//   mark_routes   station marks -> route marks           (raptor.h:228-246)
//   route_scan    loop_routes / update_route / get_earliest_transport
//                 (raptor.h:496-575, 1181-1442): batched engine = one
//                 wavefront per (64-query group, route), one query per lane;
//                 plain engine = one route per lane
//   clear marks, group_update (the two early exits, raptor.h:243-246, 370-373)
//   dest_pass     destination events of update_transfers /
//                 update_intermodal_footpaths / update_footpaths, simulated
//                 in list order per query: yields the exact time_at_dest
//                 trajectory of the phase
//   transfer_fp   all other transfer and footpath relaxations, each checked
//                 against the time_at_dest value in force at its position
//   rt_fixup      round_times[k] = best for stations improved in the phase
//   hazards       exact replay for the rare stations where time_at_dest fell
//                 between a passing and a later pruned relaxation
//
// The integer semantics (loose <= pruning, pre-writes of round_times, clamp
// to int16, uint16 day arithmetic) follow raptor.h line by line; every
// comparison is templated on the search direction.
#pragma once

#include <cstdint>

#include "device.h"
#include "rc_record.h"

namespace nigiri_amd {

// Launch-shape knobs chosen by the autotuner (runtime values; the results
// do not depend on them).
struct round_tune {
  std::uint32_t route_claim;       // packed route wavefronts claimed per atomic
  std::uint32_t transfer_claim;    // stations claimed per atomic (transfers)
  std::uint32_t pack_max_entries;  // (route, group) entries per packed wavefront
  std::uint32_t transfer_pack_max_entries;  // (station, group) entries per packed transfer wavefront
  std::uint32_t loose_lanes;  // packed route wavefronts of fewer lanes go loose (1: none; <= 33)
};

struct round_args {
  dev_image im;
  dev_state st;
  round_tune tune;
  std::int16_t invalid;
  // optional timing events around the route scan / transfer kernels
  hipEvent_t ev[4];
};

// Device lower bounds by PHAST (the exact shortest travel times of
// dijkstra.cc:18-67, cut off at 7201 minutes; see ch.h): every selected
// bound is set to unreachable, the upward search's labels are written, and
// the downward sweep runs level by level.  W = 64: the selected bounds are
// lanes masks[i] of group sel[i] in dist[G][L][64] (the 64 queries of a
// group share each location's 128 B row); W = 1: rows sel[i] of dist[*][L].
// A seed is an index into dist and its minutes.
struct lb_phast_args {
  dev_sweep sw;
  std::uint16_t* dist;
  std::uint32_t L;
  std::uint32_t const* sel;
  std::uint64_t const* masks;
  std::uint32_t n_sel;
  std::uint32_t const* seed_idx;
  std::uint16_t const* seed_min;
  std::uint32_t n_seeds;
};

// level_off: the sweep's level boundaries (host copy); the sweep runs
// levels [first_level, n_levels) (the ones before are swept on the host and
// arrive as seeds), a launch per large level
void launch_lb_phast(std::uint32_t W, lb_phast_args const& a, std::uint32_t const* level_off,
                     std::uint32_t first_level, std::uint32_t n_levels, hipStream_t stream);
// get_fastest_direct (get_fastest_direct.cc:12-100) for several queries:
// one wavefront per query runs the reference's Dijkstra over the query's
// node set (its start and destination locations, ascending), following
// footpaths (footpaths_out forward, footpaths_in backward) only into
// destinations.  Labels leave the queue in (minutes, insertion) order, as
// the reference's bucket queue, so that the stop at special_station::kEnd
// cuts the search where the reference cuts it.
constexpr std::uint32_t kFdMaxNodes = 1024U;
struct fd_node {
  std::uint32_t loc;
  std::uint16_t start;  // min(32767, shortest start offset) if is_start
  std::uint16_t off;    // shortest destination offset if is_dest
  std::uint8_t is_start, is_dest;
};
struct fd_query {
  std::uint32_t node_begin, node_end;  // into the node array
  double factor;                       // fastest_direct_factor
  float tts_factor;                    // transfer_time_settings as given
  std::int16_t tts_min, tts_add;
  std::uint8_t fwd, tts_default;
};
struct fd_args {
  dev_image im;
  fd_query const* q;
  fd_node const* nodes;
  std::int16_t* out;  // minutes, 32767 (kUnreachableDirect) if none
  std::uint32_t n;
};
void launch_fastest_direct(fd_args const& a, hipStream_t stream);

// PHAST rows [n_rows][L] (device) into the lower-bound lanes of the slots
// slots[0 .. n_rows) of the batched engine
void launch_lb_scatter64(dev_state const& st, std::uint16_t const* rows, std::uint32_t const* slots,
                         std::uint32_t n_rows, hipStream_t stream);
// lower bounds given as a row [L] (device) into the column of one slot
void launch_lb_put(std::uint32_t W, dev_state const& st, std::uint32_t slot, std::uint16_t const* row,
                   hipStream_t stream);

// The batched route scan stages the events of a stop in LDS for routes of at
// most this many transports; larger routes are scanned lane by lane from
// global memory.
constexpr std::uint32_t kBatchedMaxRouteTransports = 256U;

// Launch helpers (defined in kernels.hip), all asynchronous on `stream`.
// One round for all running slots.  `sweep` is device memory: sweep[0] =
// number of groups whose best row is swept, sweep[1..] = those groups; it is
// read only when has_sweep is set.
void launch_round(bool fwd, std::uint32_t W, round_args const& a,
                  std::uint32_t const* sweep, bool has_sweep,
                  std::uint32_t n_slots, int grid_waves, hipStream_t stream);

// The same round replayed from hipGraphs captured on first use (per
// direction, lanes, sweep on/off, grid and round_args; recaptured when the
// arguments change).  When a.ev is set the round is held as one graph per
// segment and the events are recorded between the graphs.
struct round_graphs;
round_graphs* round_graphs_create();
void round_graphs_destroy(round_graphs* g);
std::uint64_t round_graphs_captures(round_graphs const* g);
hipError_t launch_round_graph(round_graphs& g, bool fwd, std::uint32_t W, round_args const& a,
                              std::uint32_t const* sweep, bool has_sweep, std::uint32_t n_slots,
                              int grid_waves, hipStream_t stream);

struct epoch_args {
  dev_state st;
  std::int16_t invalid;
  // lane reset
  std::uint32_t const* reset_groups;  // groups with any lane to reset
  std::uint32_t n_reset_groups;
  std::uint64_t const* reset_rt_mask;  // [G] lanes whose round_times reset
  std::uint64_t const* reset_bt_mask;  // [G] lanes whose best/tmp reset
  // [G] batched engine: lanes reset densely (every element of their state,
  // not only those the validity bitmaps list)
  std::uint64_t const* reset_full_mask;
  // destination mask ops: (slot << 32 | location); dest_set[i] != 0 sets
  std::uint64_t const* dest_ops;
  std::uint8_t const* dest_set;
  std::uint32_t n_dest_ops;
  // starts: slot, location, delta
  std::uint32_t const* start_slot;
  std::uint32_t const* start_loc;
  std::int16_t const* start_delta;
  std::uint32_t n_starts;
  // per slot init: flags bit0 = running this epoch, bit1 = new query
  std::uint8_t const* slot_flags;
  std::int16_t const* slot_worst;  // d_worst_at_dest per slot
  std::uint32_t n_slots;
};

// lane resets, destination masks, slot init, starts (lower bounds go
// through launch_lb_phast)
void launch_epoch_begin(bool fwd, std::uint32_t W, epoch_args const& a,
                        hipStream_t stream);
// state download: round_times [K][L], best [L], tmp [L] of one slot
void launch_gather_slot(std::uint32_t W, dev_state const& st, std::uint32_t slot,
                        std::int16_t* rt, std::int16_t* best, std::int16_t* tmp, std::uint16_t* lb,
                        hipStream_t stream);
// round_times[k][dest bit] of the listed slots -> st.dest_rows (journey
// extraction input, raptor.h:393-413)
// Plain engine: bounds of slot dst from the round times of slot src (rows
// 0 .. n_rows-1, n_rows <= K); fwd is the direction of src's search.
void launch_fill_bounds(bool fwd, dev_state const& st, std::uint32_t src, std::uint32_t dst,
                        std::uint32_t n_rows, hipStream_t stream);
void launch_extract(std::uint32_t W, dev_state const& st, std::uint32_t const* slots,
                    std::uint32_t n_slots, hipStream_t stream);

// Journey reconstruction (reconstruct.cc:162-877, no vias, no real-time
// data, no time-dependent offsets) against a slot's round_times, right after
// the epoch that found the journey.  One wavefront per job walks the rounds
// k = transfers + 1 .. 1 back from the destination; per round it records the
// walk (transfer, footpath or egress offset) and the trip that reach the
// round's location, searching routes, trips and boarding stops 64 at a time
// in the reference's order.  The host turns the records into legs.
static_assert(kRcMaxSteps == kRowsMax - 1, "one step per round k = 1 .. 15");

struct rc_args {
  dev_image im;
  dev_state st;
  rc_job const* jobs;
  std::uint32_t n_jobs;
  std::uint32_t const* off_loc;
  std::int16_t const* off_dur;
  rc_out* out;
};

void launch_reconstruct(bool fwd, std::uint32_t W, rc_args const& a, hipStream_t stream);

}  // namespace nigiri_amd
