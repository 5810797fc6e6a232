// Host half of the search driver (nigiri's search<Dir, Algo>, include/
// nigiri/routing/search.h:78-505).  It prepares everything a query needs
// before its rounds run on the device — destinations, lower bounds, fastest
// direct connection, base day, start-label groups — and turns the
// destination round times of each epoch into the final Pareto set.
#pragma once

#include <chrono>
#include <cstdint>
#include <utility>
#include <vector>

#include "timetable.h"

namespace nigiri_amd {

// Per-query limits of the device engines: destination stations of a
// non-intermodal query, egress offsets of an intermodal one, and the
// relaxations into its destinations one round can make (each destination's
// transfer and incoming footpaths, or each egress offset).
constexpr std::size_t kMaxQueryDests = 64U;
constexpr std::size_t kMaxQueryEgress = 256U;
constexpr std::size_t kMaxQueryDestEvents = 1024U;

// Relaxations into the destinations one round can make (see above).
std::size_t dest_relaxations(timetable const& tt, bool fwd, bool intermodal,
                             std::vector<std::uint32_t> const& dests);

struct host_tts {
  bool default_{true};
  std::int16_t min_transfer_time_{0}, additional_time_{0};
  float factor_{1.0F};
};

struct host_journey {
  std::int32_t start_time_, dest_time_;
  std::uint32_t dest_;
  std::uint8_t transfers_;
  bool reconstructed_{false};
  bool error_{false};  // reconstruction found no legs (search.h:463-473)
  std::vector<nigiri_leg> legs_{};  // travel order
  bool dominates(host_journey const& o) const {  // journey.h:80-88
    if (start_time_ <= dest_time_) {
      return transfers_ <= o.transfers_ && start_time_ >= o.start_time_ &&
             dest_time_ <= o.dest_time_;
    }
    return transfers_ <= o.transfers_ && start_time_ <= o.start_time_ &&
           dest_time_ >= o.dest_time_;
  }
  std::int16_t travel_time() const {
    return static_cast<std::int16_t>(std::abs(dest_time_ - start_time_));
  }
};

// One start-time group (search.h:424-454): the labels added before one
// execute() call.
struct epoch {
  std::int32_t start_time_;
  std::int32_t worst_time_at_dest_;
  bool reset_ = false;  // reset_arrivals before this epoch (search.h:299-311)
  std::vector<std::uint32_t> loc_;
  std::vector<std::int32_t> time_;  // time at stop (unix minutes)
};

struct prepared_query {
  bool fwd_{true};
  bool ontrip_{true};
  bool intermodal_dest_{false};
  bool overlap_{false};  // start/destination overlap: empty result
  std::uint8_t max_transfers_{14};
  std::uint8_t end_k_{16};
  std::int16_t fastest_direct_{32767};
  std::int16_t max_travel_time_{7200};
  std::uint16_t base_{0};
  std::int32_t iv_from_{0}, iv_to_{0};  // the search interval (grows, search.h:237-317)
  // interval extension (search.h:237-317, interval_estimate.h)
  std::uint32_t min_connection_count_{0U};
  bool extend_earlier_{false}, extend_later_{false};
  std::int32_t ext_from_{0}, ext_to_{0};  // timetable's external interval
  std::int32_t max_from_{0}, max_to_{0};  // interval the int16 deltas can hold
  bool has_max_interval_{false};          // query::max_interval_ (query.h:89)
  std::int32_t max_iv_from_{0}, max_iv_to_{0};
  // timeout (search.h:179, 227-244): measured from the first epoch's start
  bool has_timeout_{false};
  std::int32_t timeout_s_{0};
  bool exec_started_{false};
  std::chrono::steady_clock::time_point exec_start_{};
  std::int16_t max_start_offset_{7200};
  // shortest start offset per location (get_starts, start_times.cc:251-271)
  std::vector<std::pair<std::uint32_t, std::int16_t>> start_map_;
  host_tts tts_;
  bool start_intermodal_{false};
  // query offsets in sanitized order (query.cc:59-67: sorted by duration)
  std::vector<std::pair<std::uint32_t, std::int16_t>> start_offsets_, dest_offsets_;
  std::vector<std::uint32_t> dests_;  // ascending (non-intermodal)
  std::vector<std::uint32_t> dist_loc_;  // ascending (intermodal)
  std::vector<std::uint16_t> dist_min_;
  std::vector<std::uint16_t> lb_;  // [n_locations] (host lower bounds only)
  // shortest-path sources of the lower bounds: (location, offset minutes)
  std::vector<std::pair<std::uint32_t, std::uint16_t>> lb_seeds_;
  // PHAST upward search from them: (location, minutes), for the device sweep
  std::vector<std::pair<std::uint32_t, std::uint16_t>> lb_up_;
  std::vector<epoch> epochs_;
  std::uint64_t locality_key_{0U};  // Morton code of the first start stop
  std::int32_t slot_{-1};           // engine slot that ran the query
  std::uint32_t rounds_{0U};        // rounds the engine ran for it, over its epochs
  std::uint64_t state_bytes_{0U};   // device state charged at admission
  std::vector<host_journey> results_;  // running Pareto set

  std::int32_t base_minutes(timetable const& tt) const {
    return tt.internal_from + static_cast<std::int32_t>(base_) * 1440;
  }
  // number of destination rows reported per epoch
  std::uint32_t n_dest_bits() const {
    return intermodal_dest_ ? 1U : static_cast<std::uint32_t>(dests_.size());
  }
  std::uint32_t dest_bit(std::uint32_t j) const {
    return intermodal_dest_ ? kEnd : dests_[j];
  }
};

// Device state memory of one query: end_k rows of round_times (int16) over
// all locations, plus its best (int16), tmp (int32), lower-bound (uint16)
// and route-phase mark (one byte) columns.
inline std::uint64_t query_state_bytes(std::uint32_t n_locations, std::uint32_t end_k) {
  return static_cast<std::uint64_t>(n_locations) * (2U * end_k + 9U);
}

// Scratch reused across queries by one host worker (Dijkstra buckets,
// upward searches).
struct host_scratch {
  std::vector<std::vector<std::uint32_t>> buckets_;
  ch_search_scratch ch_;
};

// Throws status_error for unsupported or malformed queries.  With
// host_lb == true the lower bounds are computed on the host (lb_); else the
// device sweeps them from lb_up_ (the upward search's labels with the host
// levels of the sweep already done; lb_seeds_ holds their sources either
// way).
void prepare(timetable const& tt, nigiri_query const& q, prepared_query& p,
             host_scratch& s, bool host_lb = true, std::int16_t const* fastest_direct = nullptr);

// get_fastest_direct's node set of one query (get_fastest_direct.cc:19-46):
// its start and destination locations in ascending order, each with the
// shortest start offset (32767 at most) and destination offset.  Throws on
// a location out of range.
struct fd_node_host {
  std::uint32_t loc;
  std::uint16_t start;  // min(32767, shortest start offset) if is_start
  std::uint16_t off;    // shortest destination offset if is_dest
  std::uint8_t is_start, is_dest;
};
void fastest_direct_nodes(timetable const& tt, nigiri_query const& q, std::vector<fd_node_host>& out);

// Journey extraction of one epoch (raptor.h:393-413).
// dest_rows[j * max_k + k] is round_times[k][dest bit j] (delta).
void add_epoch_results(timetable const& tt, prepared_query& p,
                       std::size_t epoch_idx, std::int16_t const* dest_rows,
                       std::uint32_t max_k);

// The reconstruction gate of search_interval() (search.h:457-475): indices
// of the journeys of p.results_ to reconstruct after the current epoch.
void reconstruction_candidates(prepared_query const& p, std::vector<std::size_t>& out);

// Early termination of search_interval() (search.h:477-491) after epoch
// epoch_idx: with enough journeys and a one-sided extension the remaining
// epochs are dropped and the search interval shrinks to the epoch's start.
void early_termination(prepared_query& p, std::size_t epoch_idx);

// Marks the start of search::execute's processing (search.h:226): the
// point the timeout is measured from.  Only the first call counts.
void mark_execute_start(prepared_query& p);

// is_timeout_reached (search.h:227-235).
bool timeout_reached(prepared_query const& p);

// Whether execute() (search.h:240-270) goes on after the last epoch:
// a pretrip search with fewer than min_connection_count journeys in it,
// before its timeout.
bool wants_extension(prepared_query const& p);

// The interval extension step of execute() (search.h:282-317): widens the
// search interval and appends the epochs of its new parts (the first one
// with reset_ where the reference resets the arrivals), removing journeys
// outside the interval when it resets.  Repeats while a step adds no
// epoch.  Returns whether epochs were appended; false ends the search.
bool extend_interval(timetable const& tt, prepared_query& p);

// Final filtering of search::execute (search.h:319-336).
void finish(prepared_query& p);

}  // namespace nigiri_amd
