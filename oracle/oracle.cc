// CPU oracle: a line-faithful restatement of nigiri's CPU RAPTOR search path.
//
// TEST INFRASTRUCTURE ONLY — loaded by tests/, the smoke check and the CPU
// baseline leg of bench.py, never by the product library.
//
// Restated files (reference paths relative to the nigiri repository):
//   include/nigiri/routing/raptor/raptor.h          raptor<Dir, false, 0, kOneToOne>
//   include/nigiri/routing/raptor/raptor_state.h    raptor_state
//   include/nigiri/routing/search.h                 search<Dir, Algo>
//   src/routing/raptor_search.cc                    raptor_search_with_dir
//   src/routing/start_times.cc                      get_starts, collect_destinations
//   src/routing/query.cc                            query::sanitize
//   src/routing/dijkstra.cc, routing/dijkstra.h     lower bounds
//   include/nigiri/common/dial.h                    bucket queue
//   src/routing/get_fastest_direct.cc               fastest direct connection
//   include/nigiri/routing/interval_estimate.h      initial() / extension()
//   src/routing/raptor/reconstruct.cc               reconstruct_journey
//   src/routing/raptor/optimize_footpaths.cc        optimize_footpaths
//   include/nigiri/routing/pareto_set.h, journey.h  result set
//   include/nigiri/common/delta_t.h                 clamp, split_day_mam, ...
//   include/nigiri/routing/transfer_time_settings.h adjusted_transfer_time
//
// How each part is pinned (tests/test_oracle_*.py):
//   raptor execute / search driver: golden journeys of the reference tests
//     (test/routing/raptor_test.cc:18-112, test/routing/
//     raptor_intermodal_test.cc:16-251 including the backward search with
//     interval extension, test/routing/gpu_intermodal_egress_test.cc,
//     transcribed into our builder), closed forms on periodic line
//     networks, and brute force (time-expanded exhaustive search) on tiny
//     random networks.
//   reconstruct_journey / optimize_footpaths: the printed legs of the same
//     golden journeys; every reconstructed journey is a feasible itinerary
//     of the timetable; optimize_footpaths on the scenarios of
//     test/routing/optmize_footpaths_test.cc without their via stops, with
//     expected legs derived by hand.
//   dijkstra lb: brute-force Bellman-Ford on random graphs, and lb <= the
//     travel time of every journey.  (test/routing/dijkstra_test.cc needs
//     parent-station equivalence, which this timetable model does not have.)
//   get_fastest_direct: brute force over footpath chains.
//   get_starts: the printed start-label list of
//     test/routing/start_times_test.cc:26-308 (files_simple timetable).
//   Parity unpinned (documented): the exact pruned contents of round_times /
//     best / tmp beyond what journeys expose, cista::bitset::test for indices
//     >= 512 (treated as "not set"), and the iteration order of the start
//     hash map in get_starts (taken ascending by location).
//
// Semantics that are easy to get wrong and are kept on purpose:
//   * loose (<=) pruning against time_at_dest (raptor.h:454-456);
//   * round_times[k] is pre-written before the time_at_dest / lb checks in
//     update_transfers / update_footpaths (raptor.h:709-713, 790-794), and
//     overwritten (not min'd) when the check passes (raptor.h:734, 827);
//   * the intermodal egress overwrites round_times[k][END] (raptor.h:1019);
//   * time_at_dest is a running minimum updated during the footpath phases
//     and read by later relaxations in ascending station / list order.
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <map>
#include <optional>
#include <stdexcept>
#include <thread>
#include <vector>

namespace oracle {

using delta_t = std::int16_t;
constexpr int kMaxTransfers = 14;
constexpr int kRows = kMaxTransfers + 2;
constexpr int kMaxTravelTime = 7200;
constexpr std::uint16_t kUnreachable = 0xFFFFU;
constexpr std::uint32_t kEnd = 1U;
constexpr std::int16_t kFpMaxDuration = 511;  // footpath::kMaxDuration
constexpr std::int16_t kUnreachableDirect = 32767;

// ---------------------------------------------------------------------------
// timetable accessors (timetable.h:147-177)
struct tt_t {
  explicit tt_t(nigiri_tt_view const& v) : v_{v} {}
  nigiri_tt_view const& v_;

  std::uint32_t n_locations() const { return v_.n_locations; }
  std::uint32_t n_stops(std::uint32_t r) const {
    return v_.route_stop_offsets[r + 1] - v_.route_stop_offsets[r];
  }
  std::uint32_t stop(std::uint32_t r, std::uint32_t i) const {
    return v_.route_stops[v_.route_stop_offsets[r] + i];
  }
  std::uint32_t n_transports(std::uint32_t r) const {
    return v_.route_transport_from[r + 1] - v_.route_transport_from[r];
  }
  // event_times_at_stop(r, stop_idx, ev_type)[t_offset]
  std::uint16_t event(std::uint32_t r, std::uint32_t t_offset,
                      std::uint32_t stop_idx, bool arr) const {
    auto const n_t = n_transports(r);
    return v_.event_times[v_.route_stop_time_from[r] +
                          n_t * (stop_idx * 2U - (arr ? 1U : 0U)) + t_offset];
  }
  // bitset<512>::test; indices >= 512 read as "not set" (parity unpinned)
  bool test(std::uint32_t bf, std::uint32_t day) const {
    if (day >= 512U) return false;
    return ((v_.bitfields[bf * 8U + day / 64U] >> (day % 64U)) & 1U) != 0U;
  }
  bool is_transport_active(std::uint32_t t, std::uint32_t day) const {
    return test(v_.transport_traffic_days[t], day);
  }
  bool is_route_active(std::uint32_t r, std::uint32_t day) const {
    return test(v_.route_traffic_days[r], day);
  }
  std::int32_t internal_from() const { return v_.internal_from; }
};

inline int delta_days(std::uint16_t d) { return d & 31; }
inline int delta_mam(std::uint16_t d) { return d >> 5; }
inline int delta_count(std::uint16_t d) { return delta_days(d) * 1440 + delta_mam(d); }
inline std::uint32_t stop_loc(std::uint32_t s) { return s & ((1U << 28) - 1U); }
inline bool stop_in(std::uint32_t s) { return ((s >> 28) & 1U) != 0U; }
inline bool stop_out(std::uint32_t s) { return ((s >> 29) & 1U) != 0U; }
inline bool stop_in_wc(std::uint32_t s) { return ((s >> 30) & 1U) != 0U; }
inline bool stop_out_wc(std::uint32_t s) { return ((s >> 31) & 1U) != 0U; }
inline std::uint32_t fp_target(std::uint32_t fp) { return fp & ((1U << 23) - 1U); }
inline std::int16_t fp_duration(std::uint32_t fp) {
  return static_cast<std::int16_t>(fp >> 23);
}

// delta_t.h:19-33
inline delta_t clamp(long long t) {
  return static_cast<delta_t>(std::clamp(t, static_cast<long long>(SHRT_MIN),
                                         static_cast<long long>(SHRT_MAX)));
}

// ---------------------------------------------------------------------------
// transfer_time_settings.h:21-47
struct tts_t {
  bool default_{true};
  std::int16_t min_transfer_time_{0};
  std::int16_t additional_time_{0};
  float factor_{1.0F};
};

template <typename T>
T adjusted(tts_t const& s, T const duration) {
  if (s.default_) return duration;
  return static_cast<T>(static_cast<T>(s.additional_time_) +
                        std::max(static_cast<T>(s.min_transfer_time_),
                                 static_cast<T>(static_cast<float>(duration) *
                                                s.factor_)));
}

// ---------------------------------------------------------------------------
struct offset_t {
  std::uint32_t target_;
  std::int16_t duration_;
};

struct query_t {
  bool fwd_{true};
  bool is_interval_{false};
  std::int32_t start_time_{0};
  std::int32_t iv_from_{0}, iv_to_{0};
  bool start_intermodal_{false}, dest_intermodal_{false};
  bool use_start_footpaths_{false};
  std::vector<offset_t> start_, destination_;
  std::int16_t max_start_offset_{kMaxTravelTime};
  std::uint8_t max_transfers_{kMaxTransfers};
  std::int16_t max_travel_time_{kMaxTravelTime};
  std::uint32_t min_connection_count_{0U};
  bool extend_earlier_{false}, extend_later_{false};
  tts_t tts_;
  bool has_fastest_direct_{false};
  std::int16_t fastest_direct_{0};
  double fastest_direct_factor_{1.0};
  bool has_max_interval_{false};  // query::max_interval_ (query.h:89)
  std::int32_t max_iv_from_{0}, max_iv_to_{0};
  bool has_timeout_{false};       // raptor_search / pong_search timeout
  std::int32_t timeout_s_{0};
};

query_t from_c(nigiri_query const& c) {
  query_t q;
  q.fwd_ = c.dir == NIGIRI_AMD_DIR_FWD;
  q.is_interval_ = c.is_interval != 0;
  q.start_time_ = c.start_time;
  q.iv_from_ = c.interval_from;
  q.iv_to_ = c.interval_to;
  q.start_intermodal_ = c.start_match_mode == NIGIRI_AMD_MATCH_INTERMODAL;
  q.dest_intermodal_ = c.dest_match_mode == NIGIRI_AMD_MATCH_INTERMODAL;
  q.use_start_footpaths_ = c.use_start_footpaths != 0;
  for (auto i = 0U; i != c.n_start; ++i) q.start_.push_back({c.start_loc[i], c.start_dur[i]});
  for (auto i = 0U; i != c.n_dest; ++i) q.destination_.push_back({c.dest_loc[i], c.dest_dur[i]});
  q.max_start_offset_ = c.max_start_offset;
  q.max_transfers_ = c.max_transfers;
  q.max_travel_time_ = c.max_travel_time;
  q.min_connection_count_ = c.min_connection_count;
  q.extend_earlier_ = c.extend_interval_earlier != 0;
  q.extend_later_ = c.extend_interval_later != 0;
  q.tts_.default_ = c.tts_default != 0;
  q.tts_.min_transfer_time_ = c.tts_min_transfer_time;
  q.tts_.additional_time_ = c.tts_additional_time;
  q.tts_.factor_ = c.tts_factor;
  q.has_fastest_direct_ = c.has_fastest_direct != 0;
  q.fastest_direct_ = c.fastest_direct;
  q.fastest_direct_factor_ = c.fastest_direct_factor;
  q.has_max_interval_ = c.has_max_interval != 0;
  q.max_iv_from_ = c.max_interval_from;
  q.max_iv_to_ = c.max_interval_to;
  q.has_timeout_ = c.has_timeout != 0;
  q.timeout_s_ = c.timeout_s;
  return q;
}

// query.cc:27-31, 84-91 (no vias: sanitize_via_stops is a no-op)
void sanitize(query_t& q) {
  if (q.start_intermodal_ && q.use_start_footpaths_) {
    throw std::invalid_argument(
        "intermodal start is incompatible with use_start_footpaths");
  }
  if (q.max_travel_time_ < 0 || q.max_travel_time_ > kMaxTravelTime) {
    q.max_travel_time_ = kMaxTravelTime;
  }
  auto const by_dur = [](offset_t const& a, offset_t const& b) {
    return a.duration_ < b.duration_;
  };
  std::stable_sort(q.start_.begin(), q.start_.end(), by_dur);
  std::stable_sort(q.destination_.begin(), q.destination_.end(), by_dur);
}

// ---------------------------------------------------------------------------
// journey::leg (journey.h:44-75): kind_ names the alternative held in uses_.
struct leg_t {
  int kind_{NIGIRI_AMD_LEG_RUN};
  std::uint32_t from_{0U}, to_{0U};
  std::int32_t dep_time_{0}, arr_time_{0};
  // run_enter_exit (journey.h:24-42): transport, traffic day, stop range
  std::uint32_t transport_{0U}, route_{0U};
  std::uint16_t day_{0U}, stop_from_{0U}, stop_to_{0U};
  // footpath / offset: target and duration
  std::uint32_t target_{0U};
  std::int16_t duration_{0};
};

// the leg constructor (journey.h:47-57): a and b are in search direction
leg_t make_leg(bool fwd, std::uint32_t a, std::uint32_t b, std::int32_t time_a,
               std::int32_t time_b, int kind) {
  leg_t l;
  l.kind_ = kind;
  l.from_ = fwd ? a : b;
  l.to_ = fwd ? b : a;
  l.dep_time_ = fwd ? time_a : time_b;
  l.arr_time_ = fwd ? time_b : time_a;
  return l;
}

// journey.h:80-88, pareto_set.h:21-37
struct journey_t {
  std::int32_t start_time_, dest_time_;
  std::uint32_t dest_;
  std::uint8_t transfers_;
  bool error_{false};
  bool is_reconstructed_{false};
  std::vector<leg_t> legs_{};

  bool dominates(journey_t const& o) const {
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

struct pareto_set {
  bool add(journey_t const& el) {
    std::size_t n_removed = 0U;
    for (std::size_t i = 0U; i < els_.size(); ++i) {
      if (els_[i].dominates(el)) return false;
      if (el.dominates(els_[i])) {
        ++n_removed;
        continue;
      }
      els_[i - n_removed] = els_[i];
    }
    els_.resize(els_.size() - n_removed + 1U);
    els_.back() = el;
    return true;
  }
  std::vector<journey_t> els_;
};

// ---------------------------------------------------------------------------
// raptor_state.h:22-111 (Vias = 0 view)
struct raptor_state {
  std::uint32_t n_locations_{0U};
  std::vector<delta_t> tmp_, best_, round_times_;  // round_times [kRows][L]
  std::vector<delta_t> bounds_;  // [kRows][L], written by fill_bounds
  std::vector<bool> station_mark_, prev_station_mark_, route_mark_;
  void resize(std::uint32_t n_locations, std::uint32_t n_routes) {
    n_locations_ = n_locations;
    tmp_.resize(n_locations);
    best_.resize(n_locations);
    round_times_.resize(static_cast<std::size_t>(n_locations) * kRows);
    bounds_.resize(static_cast<std::size_t>(n_locations) * kRows);
    station_mark_.resize(n_locations);
    prev_station_mark_.resize(n_locations);
    route_mark_.resize(n_routes);
  }
  delta_t& rt(std::uint32_t k, std::uint32_t l) {
    return round_times_[static_cast<std::size_t>(k) * n_locations_ + l];
  }
  delta_t& bound(std::uint32_t k, std::uint32_t l) {
    return bounds_[static_cast<std::size_t>(k) * n_locations_ + l];
  }
};

struct counters {
  oracle_stats s{};
};

// The algorithmic-byte counter of the measurement model; compiled out for
// the timed CPU baseline (-DORACLE_NO_BYTE_COUNTERS), which then keeps only
// the work counters nigiri's raptor_stats keeps too (raptor_stats.h:9-52).
#ifdef ORACLE_NO_BYTE_COUNTERS
#define ORACLE_ALG_BYTES(x) ((void)0)
#else
#define ORACLE_ALG_BYTES(x) (c_.s.alg_bytes += (x))
#endif

// ---------------------------------------------------------------------------
// raptor.h:25-1535 with Rt = false, Vias = 0, SearchMode = kOneToOne.
template <bool Fwd>
struct raptor {
  static constexpr delta_t kInvalid = Fwd ? SHRT_MAX : SHRT_MIN;
  static bool is_better(int a, int b) { return Fwd ? a < b : a > b; }
  static bool is_better_or_eq(int a, int b) { return Fwd ? a <= b : a >= b; }
  static delta_t get_best(delta_t a, delta_t b) { return is_better(a, b) ? a : b; }
  static int dir(int a) { return (Fwd ? 1 : -1) * a; }

  struct transport {
    std::uint32_t t_idx_{~0U};
    std::uint16_t day_{0xFFFFU};
    bool is_valid() const { return day_ != 0xFFFFU; }
  };

  raptor(tt_t const& tt, raptor_state& state, std::vector<bool> const& is_dest,
         std::vector<std::uint16_t> const& dist_to_dest,
         std::vector<std::uint16_t> const& lb, std::uint16_t base,
         tts_t const& tts, counters& c)
      : tt_{tt},
        n_locations_{tt.n_locations()},
        n_routes_{tt.v_.n_routes},
        state_{state},
        is_dest_{is_dest},
        dist_to_end_{dist_to_dest},
        lb_{lb},
        base_{base},
        tts_{tts},
        c_{c} {
    state_.resize(n_locations_, n_routes_);
    reset_arrivals();
    if (!dist_to_end_.empty()) {
      end_reachable_.assign(n_locations_, false);
      for (auto i = 0U; i != dist_to_end_.size(); ++i) {
        if (dist_to_end_[i] != kUnreachable) end_reachable_[i] = true;
      }
    }
  }

  // raptor.h:174-177
  void reset_arrivals() {
    std::fill(std::begin(time_at_dest_), std::end(time_at_dest_), kInvalid);
    std::fill(state_.round_times_.begin(), state_.round_times_.end(), kInvalid);
  }

  // raptor.h:179-188
  void next_start_time() {
    std::fill(state_.best_.begin(), state_.best_.end(), kInvalid);
    std::fill(state_.tmp_.begin(), state_.tmp_.end(), kInvalid);
    std::fill(state_.prev_station_mark_.begin(), state_.prev_station_mark_.end(), false);
    std::fill(state_.station_mark_.begin(), state_.station_mark_.end(), false);
    std::fill(state_.route_mark_.begin(), state_.route_mark_.end(), false);
  }

  // raptor.h:190-203
  void add_start(std::uint32_t l, std::int32_t t) {
    auto const d = unix_to_delta(t);
    state_.best_[l] = get_best(d, state_.best_[l]);
    state_.rt(0, l) = get_best(d, state_.rt(0, l));
    state_.station_mark_[l] = true;
  }

  // raptor.h:117-170 (Vias = 0, no td footpaths): bounds row k is the best
  // round time of rounds 0..k
  void fill_bounds(std::size_t n_rows) {
    for (auto x = 0U; x != n_locations_; ++x) state_.bound(0, x) = state_.rt(0, x);
    for (auto k = 1U; k < n_rows; ++k) {
      for (auto x = 0U; x != n_locations_; ++x) {
        state_.bound(k, x) = Fwd ? std::min(state_.rt(k, x), state_.bound(k - 1U, x))
                                 : std::max(state_.rt(k, x), state_.bound(k - 1U, x));
      }
    }
  }

  // raptor.h:172
  void set_bounds(unsigned last_round) { bounds_last_k_ = last_round; }

  // raptor.h:458-489 (Vias = 0: no stays): t must not be better than the
  // other direction's bound for the remaining rounds plus the transfer time
  bool within_bounds(unsigned k, std::uint32_t l, int t) {
    if (bounds_last_k_ == 0U) return true;
    auto const transfer = dir(adjusted<int>(tts_, static_cast<int>(tt_.v_.transfer_time[l])));
    return is_better_or_eq(t, static_cast<int>(state_.bound(bounds_last_k_ - k, l)) + transfer);
  }

  // raptor.h:205-414
  void execute(std::int32_t start_time, std::uint8_t max_transfers,
               std::int32_t worst_time_at_dest, pareto_set& results) {
    auto const end_k = std::min<unsigned>(max_transfers, kMaxTransfers) + 2U;
    auto const d_worst_at_dest = unix_to_delta(worst_time_at_dest);
    for (auto& t : time_at_dest_) t = get_best(d_worst_at_dest, t);

    auto& best = state_.best_;
    for (auto k = 1U; k != end_k; ++k) {
      ++c_.s.n_rounds;
      for (auto i = 0U; i != n_locations_; ++i) {
        best[i] = get_best(state_.rt(k, i), best[i]);
      }
      for (auto i = 0U; i != n_locations_; ++i) {
        if (is_dest_[i]) update_time_at_dest(k, best[i]);
      }

      auto any_marked = false;
      for (auto i = 0U; i != n_locations_; ++i) {
        if (!state_.station_mark_[i]) continue;
        ++c_.s.n_marked_stations;
        ORACLE_ALG_BYTES(8U);
        for (auto j = tt_.v_.location_route_offsets[i];
             j != tt_.v_.location_route_offsets[i + 1]; ++j) {
          ++c_.s.n_location_routes_read;
          ORACLE_ALG_BYTES(4U);
          any_marked = true;
          state_.route_mark_[tt_.v_.location_routes[j]] = true;
        }
      }
      if (!any_marked) break;

      std::swap(state_.prev_station_mark_, state_.station_mark_);
      std::fill(state_.station_mark_.begin(), state_.station_mark_.end(), false);

      any_marked = loop_routes(k);  // no filters (all claszes allowed)
      if (!any_marked) break;

      std::fill(state_.route_mark_.begin(), state_.route_mark_.end(), false);
      std::swap(state_.prev_station_mark_, state_.station_mark_);
      std::fill(state_.station_mark_.begin(), state_.station_mark_.end(), false);

      update_transfers(k);
      update_intermodal_footpaths(k);
      update_footpaths(k);
    }

    for (auto i = 0U; i != n_locations_; ++i) {
      if (!is_dest_[i]) continue;
      for (auto k = 1U; k != end_k; ++k) {
        auto const dest_time = state_.rt(k, i);
        if (dest_time != kInvalid) {
          results.add(journey_t{start_time, delta_to_unix(dest_time), i,
                                static_cast<std::uint8_t>(k - 1U)});
        }
      }
    }
  }

  std::int32_t base_minutes() const {
    return tt_.internal_from() + static_cast<std::int32_t>(base_) * 1440;
  }
  delta_t unix_to_delta(std::int32_t t) const {
    return clamp(static_cast<long long>(t) - base_minutes());
  }
  std::int32_t delta_to_unix(delta_t d) const { return base_minutes() + d; }

  bool lb_reachable(std::uint32_t i) const { return lb_[i] != kUnreachable; }
  bool is_better_loose(int a, int b) const { return is_better_or_eq(a, b); }

  bool loop_routes(unsigned k) {
    auto any_marked = false;
    for (auto r = 0U; r != n_routes_; ++r) {
      if (!state_.route_mark_[r]) continue;
      ++c_.s.n_routes_visited;
      any_marked |= update_route(k, r);
    }
    return any_marked;
  }

  // raptor.h:680-743
  void update_transfers(unsigned k) {
    auto& best = state_.best_;
    auto& tmp = state_.tmp_;
    for (auto i = 0U; i != n_locations_; ++i) {
      if (!state_.prev_station_mark_[i]) continue;
      ++c_.s.n_transfer_stations;
      ORACLE_ALG_BYTES(2U + 1U + 8U + (dist_to_end_.empty() ? 0U : 2U));
      auto const tmp_time = tmp[i];
      if (tmp_time == kInvalid) continue;
      auto const is_dest = is_dest_[i];
      auto const transfer_time =
          (!is_intermodal_dest() && is_dest)
              ? 0
              : dir(adjusted<std::uint8_t>(tts_, tt_.v_.transfer_time[i]));
      auto const fp_target_time = clamp(static_cast<long long>(tmp_time) + transfer_time);
      if (bounds_last_k_ == 0U && is_better(fp_target_time, best[i])) {
        state_.rt(k, i) = get_best(fp_target_time, state_.rt(k, i));
      }
      if (is_better(fp_target_time, best[i]) &&
          is_better_loose(fp_target_time, time_at_dest_[k])) {
        if (!lb_reachable(i) ||
            !is_better_loose(fp_target_time + dir(lb_[i]), time_at_dest_[k])) {
          continue;
        }
        if (!within_bounds(k, i, fp_target_time)) continue;
        state_.rt(k, i) = fp_target_time;
        best[i] = fp_target_time;
        ORACLE_ALG_BYTES(4U);
        state_.station_mark_[i] = true;
        if (is_dest) update_time_at_dest(k, fp_target_time);
      }
    }
  }

  // raptor.h:745-845
  void update_footpaths(unsigned k) {
    auto& best = state_.best_;
    auto& tmp = state_.tmp_;
    auto const& off = Fwd ? tt_.v_.fp_out_offsets : tt_.v_.fp_in_offsets;
    auto const& fps = Fwd ? tt_.v_.fp_out : tt_.v_.fp_in;
    for (auto i = 0U; i != n_locations_; ++i) {
      if (!state_.prev_station_mark_[i]) continue;
      for (auto e = off[i]; e != off[i + 1]; ++e) {
        ++c_.s.n_footpaths_visited;
        ORACLE_ALG_BYTES(6U);
        auto const fp = fps[e];
        auto const target = fp_target(fp);
        auto const tmp_time = tmp[i];
        if (tmp_time == kInvalid) continue;
        auto const fp_target_time = clamp(
            static_cast<long long>(tmp_time) +
            dir(adjusted<std::int16_t>(tts_, fp_duration(fp)) + 0));
        if (bounds_last_k_ == 0U && is_better(fp_target_time, best[target])) {
          state_.rt(k, target) = get_best(fp_target_time, state_.rt(k, target));
        }
        if (is_better(fp_target_time, best[target]) &&
            is_better_loose(fp_target_time, time_at_dest_[k])) {
          if (!lb_reachable(target) ||
              !is_better_loose(fp_target_time + dir(lb_[target]), time_at_dest_[k])) {
            continue;
          }
          if (!within_bounds(k, target, fp_target_time)) continue;
          state_.rt(k, target) = fp_target_time;
          best[target] = fp_target_time;
          ORACLE_ALG_BYTES(4U);
          state_.station_mark_[target] = true;
          if (is_dest_[target]) update_time_at_dest(k, fp_target_time);
        }
      }
    }
  }

  // raptor.h:954-1062 (no td offsets)
  void update_intermodal_footpaths(unsigned k) {
    if (dist_to_end_.empty()) return;
    auto& best = state_.best_;
    auto& tmp = state_.tmp_;
    for (auto i = 0U; i != n_locations_; ++i) {
      if (!state_.prev_station_mark_[i]) continue;
      if (!end_reachable_[i]) continue;
      if (dist_to_end_[i] != kUnreachable) {
        auto const tmp_time = tmp[i];
        if (tmp_time == kInvalid) continue;
        auto const end_time = clamp(static_cast<long long>(tmp_time) + dir(dist_to_end_[i]));
        if (is_better_loose(end_time, time_at_dest_[k])) {
          state_.rt(k, kEnd) = end_time;
          best[kEnd] = end_time;
          ORACLE_ALG_BYTES(4U);
          update_time_at_dest(k, end_time);
        }
      }
    }
  }

  // raptor.h:1181-1349 (no section filters, Vias = 0)
  bool update_route(unsigned k, std::uint32_t r) {
    auto& best = state_.best_;
    auto& tmp = state_.tmp_;
    bool any_marked = false;
    transport et;
    auto const n_stops = tt_.n_stops(r);
    for (auto i = 0U; i != n_stops; ++i) {
      auto const stop_idx = Fwd ? i : n_stops - i - 1U;
      auto const stp = tt_.stop(r, stop_idx);
      auto const l_idx = stop_loc(stp);
      auto const is_last = i == n_stops - 1U;
      ++c_.s.n_route_stops_visited;
      ORACLE_ALG_BYTES(8U);

      delta_t current_best = kInvalid;
      if (et.is_valid()) {
        if (can_finish(stp)) {
          auto const by_transport = time_at_stop(r, et, stop_idx, Fwd);
          current_best = get_best(get_best(state_.rt(k - 1, l_idx), tmp[l_idx]), best[l_idx]);
          if (is_better_loose(by_transport, time_at_dest_[k]) &&
              lb_reachable(l_idx) &&
              is_better_loose(by_transport + dir(lb_[l_idx]), time_at_dest_[k]) &&
              within_bounds(k, l_idx, by_transport)) {
            tmp[l_idx] = get_best(by_transport, tmp[l_idx]);
            ORACLE_ALG_BYTES(2U);
            state_.station_mark_[l_idx] = true;
            if (is_better(by_transport, current_best)) current_best = by_transport;
            any_marked = true;
          }
        }
      }

      if (is_last || !can_start(stp) || !state_.prev_station_mark_[l_idx]) continue;
      if (!lb_reachable(l_idx)) break;

      auto const et_time_at_stop =
          et.is_valid() ? time_at_stop(r, et, stop_idx, !Fwd) : kInvalid;
      auto const prev_round_time = state_.rt(k - 1, l_idx);
      if (prev_round_time != kInvalid &&
          is_better_or_eq(prev_round_time, et_time_at_stop)) {
        // split_day_mam (delta_t.h:54-59)
        auto const minutes = static_cast<int>(base_) * 1440 + prev_round_time;
        auto const day = static_cast<std::uint16_t>(minutes / 1440);
        auto const mam = static_cast<std::int16_t>(minutes % 1440);
        auto const new_et = get_earliest_transport(k, r, stop_idx, day, mam, l_idx);
        current_best = get_best(get_best(current_best, best[l_idx]), tmp[l_idx]);
        if (new_et.is_valid() &&
            (current_best == kInvalid ||
             is_better_or_eq(time_at_stop(r, new_et, stop_idx, !Fwd), et_time_at_stop))) {
          et = new_et;
        }
      }
    }
    return any_marked;
  }

  // raptor.h:1351-1442
  transport get_earliest_transport(unsigned k, std::uint32_t r, std::uint32_t stop_idx,
                                   std::uint16_t day_at_stop, std::int16_t mam_at_stop,
                                   std::uint32_t l) {
    ++c_.s.n_earliest_trip_calls;
    auto const n_t = tt_.n_transports(r);
    auto const ev_arr = !Fwd;  // kFwd ? kDep : kArr
    // iteration order: forward begin->end, backward rbegin->rend
    auto const ev_at = [&](std::uint32_t pos) {
      auto const t_offset = Fwd ? pos : n_t - 1U - pos;
      return std::make_pair(t_offset, tt_.event(r, t_offset, stop_idx, ev_arr));
    };
    auto const seek_first_day = [&]() {
      auto pos = 0U;
      for (; pos != n_t; ++pos) {
        ++c_.s.n_events_scanned;
        ORACLE_ALG_BYTES(2U);
        if (!is_better(delta_mam(ev_at(pos).second), mam_at_stop)) break;
      }
      return pos;
    };

    constexpr auto n_days_to_iterate = kMaxTravelTime / 1440 + 1;
    for (auto i = 0U; i != static_cast<unsigned>(n_days_to_iterate); ++i) {
      auto const day = static_cast<std::uint16_t>(Fwd ? day_at_stop + i : day_at_stop - i);
      ++c_.s.n_route_day_tests;
      ORACLE_ALG_BYTES(8U);
      if (!tt_.is_route_active(r, day)) continue;
      auto const begin = i == 0U ? seek_first_day() : 0U;
      if (begin == n_t) continue;
      for (auto pos = begin; pos != n_t; ++pos) {
        auto const [t_offset, ev] = ev_at(pos);
        ++c_.s.n_events_scanned;
        ORACLE_ALG_BYTES(2U);
        auto const ev_mam = delta_mam(ev);
        if (!is_better_loose(to_delta(day, ev_mam) + dir(lb_[l]), time_at_dest_[k])) {
          return {~0U, 0xFFFFU};
        }
        auto const t = tt_.v_.route_transport_from[r] + t_offset;
        if (i == 0U && !is_better_or_eq(mam_at_stop, ev_mam)) continue;
        auto const ev_day_offset = delta_days(ev);
        auto const start_day = static_cast<std::uint16_t>(static_cast<int>(day) - ev_day_offset);
        ++c_.s.n_transport_day_tests;
        ORACLE_ALG_BYTES(8U + 4U);
        if (!tt_.is_transport_active(t, start_day)) continue;
        return {t, start_day};
      }
    }
    return {};
  }

  delta_t time_at_stop(std::uint32_t r, transport t, std::uint32_t stop_idx, bool arr) const {
    auto const t_offset = t.t_idx_ - tt_.v_.route_transport_from[r];
    return to_delta(t.day_, delta_count(tt_.event(r, t_offset, stop_idx, arr)));
  }

  delta_t to_delta(std::uint16_t day, int mam) const {
    return clamp((static_cast<long long>(day) - static_cast<long long>(base_)) * 1440 + mam);
  }

  static bool can_start(std::uint32_t stp) { return Fwd ? stop_in(stp) : stop_out(stp); }
  static bool can_finish(std::uint32_t stp) { return Fwd ? stop_out(stp) : stop_in(stp); }

  bool is_intermodal_dest() const { return !dist_to_end_.empty(); }

  // raptor.h:1479-1486
  void update_time_at_dest(unsigned k, delta_t t) {
    for (auto i = k; i != static_cast<unsigned>(kRows); ++i) {
      time_at_dest_[i] = get_best(time_at_dest_[i], t);
    }
  }

  tt_t const& tt_;
  std::uint32_t n_locations_, n_routes_;
  raptor_state& state_;
  std::vector<bool> end_reachable_;
  std::vector<bool> const& is_dest_;
  std::vector<std::uint16_t> const& dist_to_end_;
  std::vector<std::uint16_t> const& lb_;
  delta_t time_at_dest_[kRows];
  unsigned bounds_last_k_{0U};  // 0: no bounds (raptor.h:1527)
  std::uint16_t base_;
  tts_t tts_;
  counters& c_;
};

// ---------------------------------------------------------------------------
// Journey reconstruction: reconstruct_journey_with_vias<Dir, 0>
// (src/routing/raptor/reconstruct.cc:22-877) and optimize_footpaths
// (src/routing/raptor/optimize_footpaths.cc:15-453) for rtt == nullptr, no
// via stops, no time-dependent offsets, all route flags allowed, profile 0.
// This timetable model has no parent stations or equivalences, so every
// location_match_mode matches a location only itself (for_each_meta.h:9-70).
constexpr std::uint32_t kStart = 0U;  // special_station::kStart

// stop word of stop `i` of route `r`, location, times of transport `t` on
// traffic day `day` (timetable.h:185-190)
std::int32_t event_time(tt_t const& tt, std::uint32_t r, std::uint32_t t, std::uint16_t day,
                        std::uint32_t stop_idx, bool arr) {
  return tt.internal_from() + static_cast<std::int32_t>(day) * 1440 +
         delta_count(tt.event(r, t - tt.v_.route_transport_from[r], stop_idx, arr));
}

// location_importance_ as assign_importance computes it
// (src/loader/init_finish.cc:64-101).  Routes carry no transport class in
// this model; every route counts with the priority of class kOther (1).
std::uint32_t importance(tt_t const& tt, std::uint32_t l) {
  auto const& v = tt.v_;
  unsigned count = 0U;  // transport_counts[kOther]
  for (auto j = v.location_route_offsets[l]; j != v.location_route_offsets[l + 1U]; ++j) {
    auto const r = v.location_routes[j];
    for (auto t = v.route_transport_from[r]; t != v.route_transport_from[r + 1U]; ++t) {
      auto const bf = v.transport_traffic_days[t];
      for (auto w = 0U; w != 8U; ++w) {
        count += static_cast<unsigned>(__builtin_popcountll(v.bitfields[bf * 8U + w]));
      }
    }
  }
  std::uint32_t x = 0U;
  x += 1.0F * static_cast<float>(count);  // x += prio[clasz] * t_count
  return x;
}

std::uint32_t max_importance(tt_t const& tt) {
  std::uint32_t m = 0U;
  for (auto l = 0U; l != tt.n_locations(); ++l) m = std::max(m, importance(tt, l));
  return m;
}

// get_penalty (optimize_footpaths.cc:262-285)
double get_penalty(tt_t const& tt, int duration, int buffer, std::uint32_t from,
                   std::uint32_t to) {
  auto weight = static_cast<double>(duration);
  if (from == to) {
    weight -= (static_cast<double>(importance(tt, from)) / max_importance(tt)) * 3U;
  }
  constexpr auto kMaxRewardedBuffer = 15;
  constexpr auto kBufferWeight = 0.2;
  weight -= kBufferWeight * static_cast<double>(std::clamp(buffer, 0, kMaxRewardedBuffer));
  return weight;
}

// optimize_footpaths (optimize_footpaths.cc:15-453) without via stops, on a
// journey whose legs are in travel order; q is the journey-direction query.
void optimize_footpaths(tt_t const& tt, query_t const& q, journey_t& j) {
  auto& legs = j.legs_;
  auto const& v = tt.v_;
  auto const stop_of = [&](leg_t const& run, std::uint32_t i) { return tt.stop(run.route_, i); };
  auto const time_of = [&](leg_t const& run, std::uint32_t i, bool arr) {
    return event_time(tt, run.route_, run.transport_, run.day_, i, arr);
  };

  // optimize_initial_departure (:15-62)
  if (legs.size() > 1U && legs[0].kind_ == NIGIRI_AMD_LEG_OFFSET &&
      legs[1].kind_ == NIGIRI_AMD_LEG_RUN) {
    auto& o_leg = legs[0];
    auto& t_leg = legs[1];
    auto best = o_leg.duration_;
    auto const to = static_cast<std::uint32_t>(t_leg.stop_to_ - 1U);
    for (auto i = 0U; i != to; ++i) {
      auto const stp = stop_of(t_leg, i);
      if (!stop_in(stp)) continue;
      for (auto const& o : q.start_) {
        if (best <= o.duration_ || o.target_ != stop_loc(stp)) continue;
        auto const dep = time_of(t_leg, i, false);
        auto const o_start = dep - o.duration_;
        if (o_leg.dep_time_ <= o_start) {
          o_leg.to_ = stop_loc(stp);
          o_leg.dep_time_ = o_start;
          o_leg.arr_time_ = dep;
          o_leg.target_ = o.target_;
          o_leg.duration_ = o.duration_;
          t_leg.from_ = stop_loc(stp);
          t_leg.dep_time_ = dep;
          t_leg.stop_from_ = static_cast<std::uint16_t>(i);
          best = o.duration_;
        }
      }
    }
  }

  // optimize_initial_start_footpath (:64-127)
  if (legs.size() > 1U && legs[0].kind_ == NIGIRI_AMD_LEG_FOOTPATH &&
      legs[1].kind_ == NIGIRI_AMD_LEG_RUN) {
    auto& fp_leg = legs[0];
    auto& t_leg = legs[1];
    auto const start_loc = fp_leg.from_;
    auto best = adjusted<std::int16_t>(q.tts_, fp_leg.duration_);
    auto const to = static_cast<std::uint32_t>(t_leg.stop_to_ - 1U);
    for (auto i = 0U; i != to; ++i) {
      auto const stp = stop_of(t_leg, i);
      if (!stop_in(stp)) continue;
      for (auto e = v.fp_out_offsets[start_loc]; e != v.fp_out_offsets[start_loc + 1U]; ++e) {
        if (fp_target(v.fp_out[e]) != stop_loc(stp)) continue;
        auto const fp_dur = adjusted<std::int16_t>(q.tts_, fp_duration(v.fp_out[e]));
        if (fp_dur < best) {
          auto const dep = time_of(t_leg, i, false);
          auto const fp_start = dep - fp_dur;
          if (fp_leg.dep_time_ <= fp_start) {
            fp_leg.to_ = stop_loc(stp);
            fp_leg.dep_time_ = fp_start;
            fp_leg.arr_time_ = dep;
            fp_leg.target_ = stop_loc(stp);
            fp_leg.duration_ = fp_dur;
            t_leg.from_ = stop_loc(stp);
            t_leg.dep_time_ = dep;
            t_leg.stop_from_ = static_cast<std::uint16_t>(i);
            best = fp_dur;
          }
        }
        break;
      }
    }
  }

  // optimize_last_arrival (:129-184)
  if (legs.size() > 1U && legs.back().kind_ == NIGIRI_AMD_LEG_OFFSET &&
      legs[legs.size() - 2U].kind_ == NIGIRI_AMD_LEG_RUN) {
    auto& o_leg = legs.back();
    auto& t_leg = legs[legs.size() - 2U];
    auto best = o_leg.duration_;
    auto const n = tt.n_stops(t_leg.route_);
    for (auto i = static_cast<std::uint32_t>(t_leg.stop_from_ + 1U); i < n; ++i) {
      auto const stp = stop_of(t_leg, i);
      if (!stop_out(stp)) continue;
      for (auto const& o : q.destination_) {
        if (best <= o.duration_ || o.target_ != stop_loc(stp)) continue;
        auto const arr = time_of(t_leg, i, true);
        auto const o_end = arr + o.duration_;
        if (o_end <= o_leg.arr_time_) {
          o_leg.from_ = stop_loc(stp);
          o_leg.dep_time_ = arr;
          o_leg.arr_time_ = o_end;
          o_leg.target_ = o.target_;
          o_leg.duration_ = o.duration_;
          t_leg.to_ = stop_loc(stp);
          t_leg.arr_time_ = arr;
          t_leg.stop_to_ = static_cast<std::uint16_t>(i + 1U);
          best = o.duration_;
        }
      }
    }
  }

  // optimize_final_egress_footpath (:186-260)
  if (legs.size() > 1U && legs.back().kind_ == NIGIRI_AMD_LEG_FOOTPATH &&
      legs[legs.size() - 2U].kind_ == NIGIRI_AMD_LEG_RUN) {
    auto& fp_leg = legs.back();
    auto& t_leg = legs[legs.size() - 2U];
    auto const dest_loc = fp_leg.to_;
    auto best = adjusted<std::int16_t>(q.tts_, fp_leg.duration_);
    auto const n = tt.n_stops(t_leg.route_);
    for (auto i = static_cast<std::uint32_t>(t_leg.stop_from_ + 1U); i < n; ++i) {
      auto const stp = stop_of(t_leg, i);
      if (!stop_out(stp)) continue;
      for (auto e = v.fp_in_offsets[dest_loc]; e != v.fp_in_offsets[dest_loc + 1U]; ++e) {
        if (fp_target(v.fp_in[e]) != stop_loc(stp)) continue;
        auto const fp_dur = adjusted<std::int16_t>(q.tts_, fp_duration(v.fp_in[e]));
        if (fp_dur < best) {
          auto const arr = time_of(t_leg, i, true);
          auto const fp_end = arr + fp_dur;
          if (fp_end <= fp_leg.arr_time_) {
            fp_leg.from_ = stop_loc(stp);
            fp_leg.dep_time_ = arr;
            fp_leg.arr_time_ = fp_end;
            fp_leg.target_ = dest_loc;
            fp_leg.duration_ = fp_dur;
            t_leg.to_ = stop_loc(stp);
            t_leg.arr_time_ = arr;
            t_leg.stop_to_ = static_cast<std::uint16_t>(i + 1U);
            best = fp_dur;
          }
        }
        break;
      }
    }
  }

  // optimize_transfers (:287-442)
  for (std::size_t i = 0U; i + 2U < legs.size(); ++i) {
    auto& leg_from = legs[i];
    auto& leg_fp = legs[i + 1U];
    auto& leg_to = legs[i + 2U];
    if (leg_from.kind_ != NIGIRI_AMD_LEG_RUN || leg_fp.kind_ != NIGIRI_AMD_LEG_FOOTPATH ||
        leg_to.kind_ != NIGIRI_AMD_LEG_RUN) {
      continue;
    }
    auto const from_start = static_cast<std::uint32_t>(leg_from.stop_from_ + 1U);
    auto const from_end = tt.n_stops(leg_from.route_);
    auto const to_end = static_cast<std::uint32_t>(leg_to.stop_to_ - 1U);
    auto best = get_penalty(tt, leg_fp.duration_, leg_to.dep_time_ - leg_fp.arr_time_,
                            leg_from.from_, leg_to.to_);
    for (auto a = from_start; a < from_end; ++a) {
      auto const sa = stop_of(leg_from, a);
      if (!stop_out(sa)) continue;
      for (auto b = 0U; b < to_end; ++b) {
        auto const sb = stop_of(leg_to, b);
        if (!stop_in(sb)) continue;
        auto const la = stop_loc(sa), lb = stop_loc(sb);
        for (auto e = v.fp_out_offsets[la]; e != v.fp_out_offsets[la + 1U]; ++e) {
          if (fp_target(v.fp_out[e]) != lb) continue;
          auto const fp_dur = adjusted<std::int16_t>(q.tts_, fp_duration(v.fp_out[e]));
          auto const arr = time_of(leg_from, a, true);
          auto const dep = time_of(leg_to, b, false);
          auto const arr_fp = arr + fp_dur;
          if (arr_fp <= dep) {
            auto const penalty = get_penalty(tt, fp_dur, dep - arr_fp, la, lb);
            if (penalty < best) {
              leg_from.to_ = la;
              leg_from.arr_time_ = arr;
              leg_from.stop_to_ = static_cast<std::uint16_t>(a + 1U);
              leg_to.from_ = lb;
              leg_to.dep_time_ = dep;
              leg_to.stop_from_ = static_cast<std::uint16_t>(b);
              leg_fp.from_ = la;
              leg_fp.to_ = lb;
              leg_fp.dep_time_ = arr;
              leg_fp.arr_time_ = arr_fp;
              leg_fp.target_ = lb;
              leg_fp.duration_ = fp_duration(v.fp_out[e]);
              best = penalty;
            }
          }
          break;
        }
      }
    }
  }
}

template <bool Fwd>
struct reconstructor {
  tt_t const& tt_;
  query_t const& q_;
  raptor_state& rs_;
  std::uint16_t base_day_;
  journey_t& j_;

  static constexpr delta_t kInvalid = Fwd ? SHRT_MAX : SHRT_MIN;
  static int dir(int a) { return Fwd ? a : -a; }
  static bool is_better_or_eq(int a, int b) { return Fwd ? a <= b : a >= b; }
  std::int32_t base() const { return tt_.internal_from() + static_cast<std::int32_t>(base_day_) * 1440; }
  delta_t unix_to_delta(std::int32_t t) const { return clamp(static_cast<long long>(t) - base()); }
  std::int32_t delta_to_unix(delta_t d) const { return base() + d; }

  // reconstruct.cc:22-28
  bool is_journey_start(std::uint32_t l) const {
    return std::any_of(q_.start_.begin(), q_.start_.end(),
                       [&](offset_t const& o) { return o.target_ == l; });
  }
  bool start_matches(int a, int b) const {
    return !q_.is_interval_ ? is_better_or_eq(a, b) : a == b;
  }

  // find_start_in_prev_round (reconstruct.cc:187-293).  The second pass
  // over the finish stop and the k == 1 meta-station case need via stops /
  // location_match_mode::kEquivalent, neither of which occurs here.
  std::optional<leg_t> find_start_in_prev_round(unsigned k, std::uint32_t r, std::uint32_t t,
                                                std::uint16_t day, std::uint32_t finish) const {
    auto const n_stops = Fwd ? finish + 1U : tt_.n_stops(r) - finish;
    for (auto i = 1U; i != n_stops; ++i) {
      auto const stop_idx = Fwd ? finish - i : finish + i;
      auto const stp = tt_.stop(r, stop_idx);
      if ((Fwd && !stop_in(stp)) || (!Fwd && !stop_out(stp))) continue;
      auto const l = stop_loc(stp);
      auto const ev = unix_to_delta(event_time(tt_, r, t, day, stop_idx, !Fwd));
      auto const round_time = rs_.rt(k - 1U, l);
      if (is_better_or_eq(round_time, ev)) {
        auto leg = make_leg(Fwd, l, stop_loc(tt_.stop(r, finish)), delta_to_unix(ev),
                            event_time(tt_, r, t, day, finish, Fwd), NIGIRI_AMD_LEG_RUN);
        leg.transport_ = t;
        leg.route_ = r;
        leg.day_ = day;
        leg.stop_from_ = static_cast<std::uint16_t>(std::min(stop_idx, finish));
        leg.stop_to_ = static_cast<std::uint16_t>(std::max(stop_idx, finish) + 1U);
        return leg;
      }
    }
    return std::nullopt;
  }

  // get_route_transport (reconstruct.cc:304-374), no td footpaths
  std::optional<leg_t> get_route_transport(unsigned k, delta_t time, std::uint32_t r,
                                           std::uint32_t stop_idx) const {
    auto const minutes = static_cast<int>(base_day_) * 1440 + time;  // split_day_mam
    auto const day = minutes / 1440, mam = minutes % 1440;
    auto const& v = tt_.v_;
    for (auto t = v.route_transport_from[r]; t != v.route_transport_from[r + 1U]; ++t) {
      auto const ev = delta_count(tt_.event(r, t - v.route_transport_from[r], stop_idx, Fwd));
      if (ev % 1440 != mam) continue;
      auto const traffic_day = static_cast<std::uint16_t>(day - ev / 1440);
      if (!tt_.is_transport_active(t, traffic_day)) continue;
      if (auto leg = find_start_in_prev_round(k, r, t, traffic_day, stop_idx)) return leg;
    }
    return std::nullopt;
  }

  // get_transport (reconstruct.cc:376-521): static routes of l in order
  std::optional<leg_t> get_transport(unsigned k, std::uint32_t l, delta_t time) const {
    auto const& v = tt_.v_;
    for (auto j = v.location_route_offsets[l]; j != v.location_route_offsets[l + 1U]; ++j) {
      auto const r = v.location_routes[j];
      auto const n = tt_.n_stops(r);
      for (auto i = 0U; i != n; ++i) {
        auto const stp = tt_.stop(r, i);
        if (stop_loc(stp) != l || (Fwd && (i == 0U || !stop_out(stp))) ||
            (!Fwd && (i == n - 1U || !stop_in(stp)))) {
          continue;
        }
        if (auto leg = get_route_transport(k, time, r, i)) return leg;
      }
    }
    return std::nullopt;
  }

  // check_fp (reconstruct.cc:523-649); without vias only the plain attempt
  // can succeed.  `duration` is the footpath's own duration.
  std::optional<std::pair<leg_t, leg_t>> check_fp(unsigned k, std::uint32_t l, delta_t curr_time,
                                                  std::uint32_t target, std::int16_t duration,
                                                  bool adjust) const {
    auto const fp_duration = adjust ? adjusted<std::int16_t>(q_.tts_, duration) : duration;
    auto const fp_start = static_cast<delta_t>(curr_time - dir(fp_duration));
    auto const transport_leg = get_transport(k, target, fp_start);
    if (!transport_leg) return std::nullopt;
    auto fp_leg = make_leg(Fwd, target, l, delta_to_unix(fp_start),
                           delta_to_unix(static_cast<delta_t>(fp_start + dir(fp_duration))),
                           NIGIRI_AMD_LEG_FOOTPATH);
    fp_leg.target_ = target;
    fp_leg.duration_ = duration;
    return std::pair{fp_leg, *transport_leg};
  }

  // get_legs (reconstruct.cc:681-797)
  std::pair<leg_t, leg_t> get_legs(unsigned k, std::uint32_t l) const {
    auto const curr_time = rs_.rt(k, l);
    if (q_.dest_intermodal_ && k == j_.transfers_ + 1U) {
      for (auto const& o : q_.destination_) {  // find_dest_leg (:651-679)
        if (o.duration_ >= kFpMaxDuration) continue;
        if (auto r = check_fp(k, l, rs_.rt(k, l), o.target_, o.duration_, false)) {
          r->first.kind_ = NIGIRI_AMD_LEG_OFFSET;
          return *r;
        }
      }
      throw std::runtime_error("intermodal destination reconstruction failed");
    }
    auto const transfer = static_cast<std::int16_t>(
        k == j_.transfers_ + 1U ? 0 : adjusted<std::uint8_t>(q_.tts_, tt_.v_.transfer_time[l]));
    if (auto r = check_fp(k, l, curr_time, l, transfer, false)) return *r;
    auto const& v = tt_.v_;
    auto const* off = Fwd ? v.fp_in_offsets : v.fp_out_offsets;
    auto const* fps = Fwd ? v.fp_in : v.fp_out;
    for (auto e = off[l]; e != off[l + 1U]; ++e) {
      if (auto r = check_fp(k, l, curr_time, fp_target(fps[e]), fp_duration(fps[e]), true)) {
        return *r;
      }
    }
    throw std::runtime_error("reconstruction failed");
  }

  // find_start_footpath (reconstruct.cc:30-160), no td offsets
  std::optional<leg_t> find_start_footpath() const {
    auto const& back = j_.legs_.back();
    auto const leg_start_location = Fwd ? back.from_ : back.to_;
    auto const leg_start_time = Fwd ? back.dep_time_ : back.arr_time_;
    if (!q_.start_intermodal_ && is_journey_start(leg_start_location) &&
        is_better_or_eq(j_.start_time_, leg_start_time)) {
      return std::nullopt;
    }
    auto const j_start_time = unix_to_delta(j_.start_time_);
    auto const fp_target_time = rs_.rt(0U, leg_start_location);
    if (q_.start_intermodal_) {
      for (auto const& o : q_.start_) {
        if (o.target_ == leg_start_location &&
            is_better_or_eq(j_.start_time_, leg_start_time - dir(o.duration_))) {
          auto leg = make_leg(Fwd, kStart, leg_start_location, leg_start_time - dir(o.duration_),
                              leg_start_time, NIGIRI_AMD_LEG_OFFSET);
          leg.target_ = o.target_;
          leg.duration_ = o.duration_;
          return leg;
        }
      }
    } else {
      auto const& v = tt_.v_;
      auto const* off = Fwd ? v.fp_in_offsets : v.fp_out_offsets;
      auto const* fps = Fwd ? v.fp_in : v.fp_out;
      for (auto e = off[leg_start_location]; e != off[leg_start_location + 1U]; ++e) {
        auto const target = fp_target(fps[e]);
        auto const fp_dur = adjusted<std::int16_t>(q_.tts_, fp_duration(fps[e]));
        if (is_journey_start(target) && fp_target_time != kInvalid &&
            start_matches(j_start_time + dir(fp_dur), fp_target_time)) {
          auto leg = make_leg(Fwd, target, leg_start_location, j_.start_time_,
                              delta_to_unix(fp_target_time), NIGIRI_AMD_LEG_FOOTPATH);
          leg.target_ = target;
          leg.duration_ = fp_dur;
          return leg;
        }
      }
    }
    throw std::runtime_error("no valid journey start found");
  }

  // reconstruct.cc:799-872
  void run() {
    j_.legs_.clear();
    auto l = j_.dest_;
    for (auto i = 0U; i <= j_.transfers_; ++i) {
      auto const k = j_.transfers_ + 1U - i;
      auto [fp_leg, transport_leg] = get_legs(k, l);
      l = Fwd ? transport_leg.from_ : transport_leg.to_;
      if (i != 0U || fp_leg.from_ != fp_leg.to_ || fp_leg.dep_time_ != fp_leg.arr_time_) {
        j_.legs_.push_back(fp_leg);
      }
      j_.legs_.push_back(transport_leg);
    }
    if (auto init = find_start_footpath()) j_.legs_.push_back(*init);
    if (Fwd) {
      std::reverse(j_.legs_.begin(), j_.legs_.end());
    } else {
      // footpaths begin at the arrival of the previous leg
      for (std::size_t x = 1U; x < j_.legs_.size(); ++x) {
        if (j_.legs_[x].kind_ != NIGIRI_AMD_LEG_FOOTPATH) continue;
        auto const diff = j_.legs_[x].dep_time_ - j_.legs_[x - 1U].arr_time_;
        j_.legs_[x].dep_time_ -= diff;
        j_.legs_[x].arr_time_ -= diff;
      }
    }
    auto journey_q = q_;
    if (!Fwd) {  // query::flip_dir (query.cc:52-57)
      std::swap(journey_q.start_, journey_q.destination_);
      std::swap(journey_q.start_intermodal_, journey_q.dest_intermodal_);
    }
    optimize_footpaths(tt_, journey_q, j_);
    j_.is_reconstructed_ = true;
  }
};

// ---------------------------------------------------------------------------
// dial.h + dijkstra.h:31-70 + dijkstra.cc:18-67
void dijkstra_lb(tt_t const& tt, query_t const& q, std::uint32_t const* off,
                 std::uint32_t const* edges, std::vector<std::uint16_t>& dists) {
  auto const L = tt.n_locations();
  dists.assign(L, kUnreachable);
  std::map<std::uint32_t, std::uint16_t> min;
  for (auto const& d : q.destination_) {
    auto const l = d.target_;  // kExact / kIntermodal without children
    auto it = min.find(l);
    if (it == min.end()) it = min.emplace(l, dists[l]).first;
    it->second = std::min(static_cast<std::uint16_t>(d.duration_), it->second);
  }
  constexpr std::size_t n_buckets = kMaxTravelTime + 1;
  std::vector<std::vector<std::pair<std::uint32_t, std::uint16_t>>> buckets(n_buckets);
  std::size_t size = 0U, current = 0U;
  auto const push = [&](std::uint32_t l, std::uint16_t d) {
    buckets.at(d).emplace_back(l, d);
    current = std::min<std::size_t>(current, d);
    ++size;
  };
  for (auto const& [l, d] : min) {
    push(l, d);
    dists[l] = std::min(d, dists[l]);
  }
  while (size != 0U) {
    while (buckets[current].empty()) ++current;
    auto const [l, d] = buckets[current].back();
    buckets[current].pop_back();
    --size;
    if (dists[l] < d) continue;
    for (auto e = off[l]; e != off[l + 1]; ++e) {
      auto const target = fp_target(edges[e]);
      auto const new_dist = static_cast<int>(d) + fp_duration(edges[e]);
      if (new_dist < dists[target] && new_dist < static_cast<int>(n_buckets) &&
          new_dist < 65535) {
        dists[target] = static_cast<std::uint16_t>(new_dist);
        push(target, static_cast<std::uint16_t>(new_dist));
      }
    }
  }
}

// get_fastest_direct.cc:12-100
std::int16_t get_fastest_direct(tt_t const& tt, query_t const& q, bool fwd) {
  if (q.has_fastest_direct_) return q.fastest_direct_;
  constexpr int max_dist = 65535;
  std::map<std::uint32_t, std::uint16_t> dists;
  auto const get_dist = [&](std::uint32_t l) -> std::uint16_t& {
    return dists.try_emplace(l, static_cast<std::uint16_t>(32767)).first->second;
  };
  for (auto const& s : q.start_) {
    auto& d = get_dist(s.target_);
    d = std::min(static_cast<std::uint16_t>(s.duration_), d);
  }
  std::map<std::uint32_t, std::uint16_t> dest_offsets;
  for (auto const& o : q.destination_) {
    auto const d = static_cast<std::uint16_t>(o.duration_);
    auto const it = dest_offsets.find(o.target_);
    if (it != dest_offsets.end() && it->second > d) {
      it->second = d;
    } else {
      dest_offsets.emplace(o.target_, d);
    }
  }
  // bucket queue with max_dist buckets; the pop order inside a bucket does
  // not change the result (every settled label is visited)
  std::vector<std::vector<std::pair<std::uint32_t, std::uint16_t>>> buckets(max_dist);
  std::size_t size = 0U, current = 0U;
  auto const push = [&](std::uint32_t l, std::uint16_t d) {
    buckets.at(d).emplace_back(l, d);
    current = std::min<std::size_t>(current, d);
    ++size;
  };
  for (auto const& [l, d] : dists) push(l, d);
  auto end_dist = static_cast<std::uint16_t>(kUnreachableDirect);
  auto const* off = fwd ? tt.v_.fp_out_offsets : tt.v_.fp_in_offsets;
  auto const* fps = fwd ? tt.v_.fp_out : tt.v_.fp_in;
  while (size != 0U) {
    while (buckets[current].empty()) ++current;
    auto const [l, d] = buckets[current].back();
    buckets[current].pop_back();
    --size;
    if (l == kEnd) break;
    if (get_dist(l) < d) continue;
    for (auto e = off[l]; e != off[l + 1]; ++e) {
      auto const new_dist =
          static_cast<int>(d) +
          adjusted<std::uint16_t>(q.tts_, static_cast<std::uint16_t>(fp_duration(fps[e])));
      if (new_dist > max_dist) continue;
      auto const target = fp_target(fps[e]);
      if (dest_offsets.find(target) == dest_offsets.end()) continue;
      auto& target_dist = get_dist(target);
      if (new_dist < target_dist && new_dist < max_dist && new_dist < max_dist) {
        target_dist = static_cast<std::uint16_t>(new_dist);
        push(target, static_cast<std::uint16_t>(new_dist));
      }
    }
    if (auto const it = dest_offsets.find(l); it != dest_offsets.end()) {
      auto const new_dist = static_cast<int>(d) + it->second;
      if (new_dist < max_dist && new_dist < end_dist) {
        end_dist = static_cast<std::uint16_t>(new_dist);
      }
    }
  }
  return end_dist == static_cast<std::uint16_t>(kUnreachableDirect)
             ? kUnreachableDirect
             : static_cast<std::int16_t>(std::ceil(end_dist * q.fastest_direct_factor_));
}

// ---------------------------------------------------------------------------
// start_times.h / start_times.cc
struct start_t {
  std::int32_t time_at_start_, time_at_stop_;
  std::uint32_t stop_;
  auto key() const { return std::tie(time_at_start_, time_at_stop_, stop_); }
  bool operator<(start_t const& o) const { return key() < o.key(); }
};

struct interval_t {
  std::int32_t from_, to_;
  bool contains(std::int32_t t) const { return t >= from_ && t < to_; }
  std::int32_t size() const { return to_ - from_; }
  bool operator==(interval_t const& o) const { return from_ == o.from_ && to_ == o.to_; }
};

// add_start_times_at_stop (start_times.cc:40-121)
void add_start_times_at_stop(bool fwd, tt_t const& tt, std::uint32_t r,
                             std::uint32_t stop_idx, std::uint32_t location_idx,
                             interval_t const& iv_at_start, interval_t const& iv_at_stop,
                             std::int16_t offset, std::vector<start_t>& starts) {
  auto const is_better_or_eq = [&](std::int32_t a, std::int32_t b) {
    return fwd ? a <= b : a >= b;
  };
  auto const day_idx = [&](std::int32_t t) {
    auto const m = t - tt.internal_from();
    return static_cast<std::uint16_t>(m / 1440);
  };
  auto const first_day_idx = day_idx(iv_at_stop.from_);
  auto const last_day_idx = day_idx(iv_at_stop.to_);
  auto const n_t = tt.n_transports(r);
  for (auto t_off = 0U; t_off != n_t; ++t_off) {
    auto const t = tt.v_.route_transport_from[r] + t_off;
    auto const stop_time = tt.event(r, t_off, stop_idx, !fwd);
    auto const day_offset = static_cast<std::uint16_t>(delta_count(stop_time) / 1440);
    auto const stop_time_mam = delta_count(stop_time) % 1440;
    for (auto day = first_day_idx; day <= last_day_idx; ++day) {
      auto const ev_time = tt.internal_from() + static_cast<std::int32_t>(day) * 1440 + stop_time_mam;
      if (tt.is_transport_active(t, static_cast<std::uint16_t>(day - day_offset)) &&
          iv_at_stop.contains(ev_time)) {
        auto const d = offset;
        if (d == kFpMaxDuration) continue;
        auto const time_at_start = fwd ? ev_time - d : ev_time + d;
        if (!iv_at_start.contains(time_at_start)) continue;
        if (!starts.empty() && starts.back().time_at_start_ == time_at_start &&
            is_better_or_eq(starts.back().time_at_stop_, ev_time)) {
          continue;
        }
        starts.push_back({time_at_start, ev_time, location_idx});
      }
      if (day == 0xFFFFU) break;
    }
  }
}

// add_starts_in_interval (start_times.cc:123-239), no real-time starts
void add_starts_in_interval(bool fwd, tt_t const& tt, interval_t const& iv,
                            std::uint32_t l, std::int16_t offset,
                            std::int16_t max_start_offset,
                            std::vector<start_t>& starts, bool add_ontrip) {
  for (auto j = tt.v_.location_route_offsets[l]; j != tt.v_.location_route_offsets[l + 1]; ++j) {
    auto const r = tt.v_.location_routes[j];
    auto const n = tt.n_stops(r);
    for (auto i = 0U; i != n; ++i) {
      auto const stp = tt.stop(r, i);
      if (stop_loc(stp) != l) continue;
      if ((!fwd && (i == 0U || !stop_out(stp))) || (fwd && (i == n - 1U || !stop_in(stp)))) {
        continue;
      }
      add_start_times_at_stop(
          fwd, tt, r, i, l, iv,
          fwd ? interval_t{iv.from_, iv.to_ + max_start_offset}
              : interval_t{iv.from_ - max_start_offset, iv.to_},
          offset, starts);
    }
  }
  if (add_ontrip) {
    auto const time_at_start = fwd ? iv.to_ : iv.from_ - 1;
    auto const d = offset;
    if (d != kFpMaxDuration) {
      starts.push_back({time_at_start, fwd ? iv.to_ + d : iv.from_ - 1 - d, l});
    }
  }
}

// get_starts (start_times.cc:241-319), no vias, no td offsets.  The start
// map is iterated in ascending location order.
void get_starts(bool fwd, tt_t const& tt, bool is_interval, std::int32_t t,
                interval_t const& iv, query_t const& q, std::vector<start_t>& starts,
                bool add_ontrip) {
  std::map<std::uint32_t, std::int16_t> shortest_start;
  auto const update = [&](std::uint32_t l, std::int16_t d) {
    auto const it = shortest_start.try_emplace(l, d).first;
    it->second = std::min(it->second, d);
  };
  for (auto const& o : q.start_) {
    auto const l = o.target_;
    update(l, o.duration_);
    if (q.use_start_footpaths_) {
      auto const* off = fwd ? tt.v_.fp_out_offsets : tt.v_.fp_in_offsets;
      auto const* fps = fwd ? tt.v_.fp_out : tt.v_.fp_in;
      for (auto e = off[l]; e != off[l + 1]; ++e) {
        update(fp_target(fps[e]),
               static_cast<std::int16_t>(o.duration_ + adjusted<std::int16_t>(q.tts_, fp_duration(fps[e]))));
      }
    }
  }
  for (auto const& [l, o] : shortest_start) {
    if (is_interval) {
      add_starts_in_interval(fwd, tt, iv, l, o, q.max_start_offset_, starts, add_ontrip);
    } else {
      starts.push_back({t, fwd ? t + o : t - o, l});
    }
  }
}

// collect_destinations (start_times.cc:321-353)
void collect_destinations(tt_t const& tt, query_t const& q, std::vector<bool>& is_dest,
                          std::vector<std::uint16_t>& dist_to_dest) {
  is_dest.assign(tt.n_locations(), false);
  if (q.dest_intermodal_) {
    is_dest[kEnd] = true;
    dist_to_dest.assign(tt.n_locations(), kUnreachable);
  } else {
    dist_to_dest.clear();
  }
  for (auto const& d : q.destination_) {
    if (q.dest_intermodal_) {
      dist_to_dest[d.target_] =
          std::min(dist_to_dest[d.target_], static_cast<std::uint16_t>(d.duration_));
    } else {
      is_dest[d.target_] = true;
    }
  }
}

// ---------------------------------------------------------------------------
struct search_result {
  pareto_set results_;
  interval_t interval_;
  std::uint16_t base_;
  std::int16_t fastest_direct_;
};

// round<days> of a minute time point: nearest day, ties to even (std::chrono)
std::int64_t round_days(std::int64_t minutes) {
  auto const floor_d = minutes >= 0 ? minutes / 1440 : -((-minutes + 1439) / 1440);
  auto const lower = floor_d * 1440;
  auto const diff_lo = minutes - lower, diff_hi = lower + 1440 - minutes;
  if (diff_lo < diff_hi) return floor_d;
  if (diff_lo > diff_hi) return floor_d + 1;
  return (floor_d % 2 == 0) ? floor_d : floor_d + 1;
}

// interval_estimator (include/nigiri/routing/interval_estimate.h:12-156):
// the search interval grows until min_connection_count journeys start in it,
// within the timetable's external interval and the int16 delta range.
template <bool Fwd>
struct interval_estimator {
  static constexpr std::int32_t kMaxSearchIntervalSize = (32767 / 1440 - 7) * 1440;  // limits.h:11-13

  interval_estimator(tt_t const& tt, query_t const& q)
      : q_{q}, ext_{tt.v_.date_range_from, tt.v_.date_range_to} {
    auto const start = q.is_interval_ ? interval_t{q.iv_from_, q.iv_to_}
                                      : interval_t{q.start_time_, q.start_time_};
    auto const days = q.has_fastest_direct_ ? 1 : 2;
    auto const size = static_cast<std::int64_t>(start.to_) - start.from_;
    auto const rd = [](std::int64_t minutes) {
      return static_cast<std::int32_t>(round_days(minutes) * 1440);
    };
    max_.from_ = std::max(rd(static_cast<std::int64_t>(start.from_) - days * 1440),
                          rd(start.from_ + (size - kMaxSearchIntervalSize) / 2));
    max_.to_ = std::min(rd(static_cast<std::int64_t>(start.to_) + days * 1440),
                        rd(start.from_ + (size + kMaxSearchIntervalSize) / 2));
  }

  interval_t initial(interval_t const& itv) const {
    if (q_.min_connection_count_ == 0U || (!q_.extend_earlier_ && !q_.extend_later_)) return itv;
    auto n = itv;
    auto const ext = 60 * static_cast<std::int32_t>(q_.min_connection_count_);
    if (Fwd ? can_extend_later(itv) : can_extend_earlier(itv)) {
      if (Fwd) {
        if (can_extend_earlier(itv)) {
          n.from_ -= 60;
        } else {
          n.to_ += 60;
        }
        n.to_ += ext;
      } else {
        if (can_extend_later(itv)) {
          n.to_ += 60;
        } else {
          n.from_ -= 60;
        }
        n.from_ -= ext;
      }
    } else if (Fwd) {
      if (q_.extend_earlier_) n.from_ -= 60 + ext;
    } else {
      if (q_.extend_later_) n.to_ += 60 + ext;
    }
    clamp(n);
    return n;
  }

  interval_t extension(interval_t const& itv, unsigned num_con_req) const {
    if (num_con_req == 0U || (!q_.extend_earlier_ && !q_.extend_later_)) return itv;
    auto n = itv;
    auto const ext = static_cast<std::int32_t>(static_cast<unsigned>(itv.size()) * num_con_req);
    if (can_extend_earlier(itv) && can_extend_later(itv)) {
      n.from_ -= ext / 2;
      n.to_ += ext / 2;
    } else {
      if (q_.extend_earlier_) n.from_ -= ext;
      if (q_.extend_later_) n.to_ += ext;
    }
    clamp(n);
    return n;
  }

 private:
  bool can_extend_earlier(interval_t const& itv) const {
    return q_.extend_earlier_ && itv.from_ != ext_.from_;
  }
  bool can_extend_later(interval_t const& itv) const {
    return q_.extend_later_ && itv.to_ != ext_.to_;
  }
  void clamp(interval_t& itv) const {
    itv.from_ = std::clamp(itv.from_, ext_.from_, ext_.to_);
    itv.to_ = std::clamp(itv.to_, ext_.from_, ext_.to_);
    itv.from_ = std::clamp(itv.from_, max_.from_, max_.to_);
    itv.to_ = std::clamp(itv.to_, max_.from_, max_.to_);
    if (q_.has_max_interval_) {  // interval_estimate.h:142-145
      itv.from_ = std::clamp(itv.from_, q_.max_iv_from_, q_.max_iv_to_);
      itv.to_ = std::clamp(itv.to_, q_.max_iv_from_, q_.max_iv_to_);
    }
  }

  query_t const& q_;
  interval_t ext_, max_{};
};

// search<Dir, raptor<...>>::search(...) + execute() (search.h:84-345) as
// driven by raptor_search_with_dir (raptor_search.cc:51-79).
template <bool Fwd>
void run_search(tt_t const& tt, query_t q, raptor_state& rs, search_result& out,
                counters& c, std::vector<std::uint16_t>* lb_out,
                std::vector<delta_t>* tad_out) {
  sanitize(q);  // raptor_search_with_dir
  auto search_interval = q.is_interval_ ? interval_t{q.iv_from_, q.iv_to_}
                                        : interval_t{q.start_time_, q.start_time_};
  auto const fastest_direct = get_fastest_direct(tt, q, Fwd);  // raw tts

  // init (search.h:84-172)
  auto& tts = q.tts_;
  tts.factor_ = std::max(tts.factor_, 1.0F);
  tts.min_transfer_time_ = std::max<std::int16_t>(tts.min_transfer_time_, 0);
  tts.additional_time_ = std::max<std::int16_t>(tts.additional_time_, 0);
  tts.default_ = tts.factor_ == 1.0F && tts.min_transfer_time_ == 0 && tts.additional_time_ == 0;
  std::vector<bool> is_dest;
  std::vector<std::uint16_t> dist_to_dest, lb;
  collect_destinations(tt, q, is_dest, dist_to_dest);
  dijkstra_lb(tt, q, Fwd ? tt.v_.lb_fwd_offsets : tt.v_.lb_bwd_offsets,
              Fwd ? tt.v_.lb_fwd : tt.v_.lb_bwd, lb);
  auto const mid = static_cast<std::int64_t>(search_interval.from_) +
                   (search_interval.to_ - search_interval.from_) / 2;
  // day_idx_t{duration_cast<days>(round<days>(mid) - internal_from)}
  auto const base_day = static_cast<std::uint16_t>(
      (round_days(mid) * 1440 - tt.internal_from()) / 1440);
  raptor<Fwd> algo{tt, rs, is_dest, dist_to_dest, lb, base_day, tts, c};
  sanitize(q);

  // execute (search.h:204-345)
  out.results_.els_.clear();
  out.base_ = base_day;
  out.fastest_direct_ = fastest_direct;
  if (lb_out != nullptr) *lb_out = lb;
  auto const is_ontrip = !q.is_interval_;
  if (!q.start_intermodal_ && !q.dest_intermodal_) {
    for (auto const& d : q.destination_) {
      for (auto const& s : q.start_) {
        if (s.target_ == d.target_) {
          out.interval_ = search_interval;
          return;  // start_dest_overlap
        }
      }
    }
  }
  interval_estimator<Fwd> const est{tt, q};
  if (!is_ontrip) search_interval = est.initial(search_interval);

  // add_start_labels (search.h:400-411)
  std::vector<start_t> starts;
  auto const add_start_labels = [&](interval_t const& iv, bool add_ontrip) {
    get_starts(Fwd, tt, true, 0, iv, q, starts, add_ontrip);
    std::sort(starts.begin(), starts.end(),
              [](start_t const& a, start_t const& b) { return Fwd ? b < a : a < b; });
  };
  if (search_interval.size() != 0) {
    add_start_labels(search_interval, true);
  } else if (q.is_interval_) {
    add_start_labels(interval_t{q.iv_from_, q.iv_to_}, true);
  } else {
    get_starts(Fwd, tt, false, q.start_time_, search_interval, q, starts, true);
    std::sort(starts.begin(), starts.end(),
              [](start_t const& a, start_t const& b) { return Fwd ? b < a : a < b; });
  }

  auto& results = out.results_;
  auto const n_results_in_interval = [&]() {  // search.h:382-391
    if (!q.is_interval_) return static_cast<unsigned>(results.els_.size());
    return static_cast<unsigned>(std::count_if(
        results.els_.begin(), results.els_.end(),
        [&](journey_t const& j) { return search_interval.contains(j.start_time_); }));
  };
  auto const min_count = static_cast<unsigned>(q.min_connection_count_);
  // is_timeout_reached (search.h:227-235): measured from here, after the
  // first start labels
  auto const processing_start_time = std::chrono::steady_clock::now();
  auto const is_timeout_reached = [&]() {
    return q.has_timeout_ && std::chrono::steady_clock::now() - processing_start_time >=
                                 std::chrono::seconds{q.timeout_s_};
  };

  // execute() (search.h:204-320)
  while (true) {
    // search_interval() (search.h:419-494)
    auto early_termination = false;
    for (auto from = 0U; from != starts.size();) {
      auto to = from;
      while (to != starts.size() && starts[to].time_at_start_ == starts[from].time_at_start_) ++to;
      if (early_termination) {
        from = to;
        continue;
      }
      ++c.s.n_epochs;
      algo.next_start_time();
      auto const start_time = starts[from].time_at_start_;
      for (auto i = from; i != to; ++i) algo.add_start(starts[i].stop_, starts[i].time_at_stop_);
      auto const worst_time_at_dest =
          start_time + (Fwd ? 1 : -1) * (std::min(fastest_direct, q.max_travel_time_) + 1);
      algo.execute(start_time, q.max_transfers_, worst_time_at_dest, results);
      for (auto& j : results.els_) {
        if (!j.is_reconstructed_ && !j.error_ &&
            (is_ontrip || search_interval.contains(j.start_time_)) &&
            j.travel_time() < fastest_direct) {
          try {
            reconstructor<Fwd>{tt, q, rs, base_day, j}.run();
          } catch (std::exception const&) {
            j.error_ = true;
          }
        }
      }
      if (min_count > 0U && n_results_in_interval() >= min_count &&
          ((Fwd && q.extend_earlier_ && !q.extend_later_) ||
           (!Fwd && !q.extend_earlier_ && q.extend_later_))) {
        early_termination = true;
        if (Fwd) {
          search_interval.from_ = start_time;
        } else {
          search_interval.to_ = start_time + 1;
        }
      }
      from = to;
    }

    if (results.els_.empty() || is_ontrip || n_results_in_interval() >= min_count ||
        is_timeout_reached()) {
      break;
    }

    auto const new_interval = est.extension(search_interval, min_count - n_results_in_interval());
    if (new_interval == search_interval) break;
    starts.clear();
    auto const remove_ontrip_results = [&]() {
      auto& e = results.els_;
      e.erase(std::remove_if(e.begin(), e.end(),
                             [&](journey_t const& j) { return !search_interval.contains(j.start_time_); }),
              e.end());
    };
    if (new_interval.from_ != search_interval.from_) {
      add_start_labels(interval_t{new_interval.from_, search_interval.from_}, !Fwd);
      if (!Fwd) {
        algo.reset_arrivals();
        remove_ontrip_results();
      }
    }
    if (new_interval.to_ != search_interval.to_) {
      add_start_labels(interval_t{search_interval.to_, new_interval.to_}, Fwd);
      if (Fwd) {
        algo.reset_arrivals();
        remove_ontrip_results();
      }
    }
    search_interval = new_interval;
  }

  if (!is_ontrip) {
    auto& e = results.els_;
    e.erase(std::remove_if(e.begin(), e.end(),
                           [&](journey_t const& j) {
                             return !search_interval.contains(j.start_time_) ||
                                    j.travel_time() >= fastest_direct ||
                                    j.travel_time() > q.max_travel_time_;
                           }),
            e.end());
    std::sort(e.begin(), e.end(), [](journey_t const& a, journey_t const& b) {
      return std::tie(a.start_time_, a.transfers_) < std::tie(b.start_time_, b.transfers_);
    });
  }
  auto& e = results.els_;
  e.erase(std::remove_if(e.begin(), e.end(),
                         [](journey_t const& j) { return !j.is_reconstructed_; }),
          e.end());
  out.interval_ = search_interval;
  if (tad_out != nullptr) tad_out->assign(algo.time_at_dest_, algo.time_at_dest_ + kRows);
}

// query::flip_dir (query.cc:52-57; no td offsets, no vias)
void flip_dir(query_t& q) {
  std::swap(q.start_, q.destination_);
  std::swap(q.start_intermodal_, q.dest_intermodal_);
}

constexpr std::int32_t kMinLookAhead = 1440;  // raptor/pong.h:11

// pong<SearchDir, Rt = false, Vias = 0> (src/routing/raptor/pong.cc:62-428):
// a ping search from each start time in the search direction, then for each
// of its journeys (most transfers first) a pong search in the other
// direction from the arrival, pruned by the ping's round times
// (fill_bounds / set_bounds), which finds the latest departure; the next
// ping starts one minute after the earliest of those departures.  Results
// are the pong journeys with start and destination swapped.  Left out:
// enrich_with_slow_direct (a no-op: query::slow_direct_ is false,
// direct.cc:614-616) and the earliest-alternative rewrite of the legs
// between three transit legs (pong.cc:430-470), which changes legs only.
// Both raptors share one raptor_state, as in the reference.
template <bool Fwd>
void run_pong(tt_t const& tt, query_t q, raptor_state& rs, search_result& out, counters& c) {
  auto const processing_start_time = std::chrono::steady_clock::now();  // pong.cc:83
  auto const is_timeout_reached = [&]() {  // pong.cc:235-241
    return q.has_timeout_ && std::chrono::steady_clock::now() - processing_start_time >=
                                 std::chrono::seconds{q.timeout_s_};
  };
  sanitize(q);
  auto const fastest_direct = get_fastest_direct(tt, q, Fwd);
  auto const search_interval = q.is_interval_ ? interval_t{q.iv_from_, q.iv_to_}
                                              : interval_t{q.start_time_, q.start_time_};
  auto const mid = static_cast<std::int64_t>(search_interval.from_) +
                   (search_interval.to_ - search_interval.from_) / 2;
  auto const base_day = static_cast<std::uint16_t>(
      (round_days(mid) * 1440 - tt.internal_from()) / 1440);

  // ping (pong.cc:113-151): lower bounds from the destinations
  std::vector<bool> ping_is_dest;
  std::vector<std::uint16_t> ping_dist_to_dest, ping_lb;
  collect_destinations(tt, q, ping_is_dest, ping_dist_to_dest);
  dijkstra_lb(tt, q, Fwd ? tt.v_.lb_fwd_offsets : tt.v_.lb_bwd_offsets,
              Fwd ? tt.v_.lb_fwd : tt.v_.lb_bwd, ping_lb);
  raptor<Fwd> ping{tt, rs, ping_is_dest, ping_dist_to_dest, ping_lb, base_day, q.tts_, c};

  // pong (pong.cc:153-206): the flipped query, lower bounds all zero
  flip_dir(q);
  std::vector<bool> pong_is_dest;
  std::vector<std::uint16_t> pong_dist_to_dest;
  collect_destinations(tt, q, pong_is_dest, pong_dist_to_dest);
  std::vector<std::uint16_t> const pong_lb(tt.n_locations(), 0U);
  raptor<!Fwd> pong{tt, rs, pong_is_dest, pong_dist_to_dest, pong_lb, base_day, q.tts_, c};
  flip_dir(q);

  // pong.cc:211-395
  auto& results = out.results_;
  results.els_.clear();
  std::vector<start_t> starts;
  auto start_time = Fwd ? search_interval.from_ : search_interval.to_ - 1;
  auto const end_time = Fwd ? search_interval.to_ : search_interval.from_ - 1;
  auto const is_better = [](std::int32_t a, std::int32_t b) { return Fwd ? a < b : a > b; };
  auto const is_validated = [&](journey_t const& j) { return is_better(j.dest_time_, start_time); };
  auto const result_count = [&](bool include_too_slow) {
    return static_cast<unsigned>(std::count_if(
        results.els_.begin(), results.els_.end(), [&](journey_t const& j) {
          return is_validated(j) &&
                 (include_too_slow ||
                  (j.travel_time() < fastest_direct && j.travel_time() <= q.max_travel_time_));
        }));
  };
  interval_t const external{tt.v_.date_range_from, tt.v_.date_range_to};
  while ((is_better(start_time, end_time) ||
          result_count(true) + result_count(false) < 2U * q.min_connection_count_) &&
         external.contains(start_time) && !is_timeout_reached()) {
    starts.clear();
    get_starts(Fwd, tt, false, start_time, search_interval, q, starts, false);
    ping.reset_arrivals();
    ping.next_start_time();
    for (auto const& st : starts) ping.add_start(st.stop_, st.time_at_stop_);
    auto const worst_time_at_dest =
        start_time + (Fwd ? 1 : -1) * std::min<std::int32_t>(q.max_travel_time_ + kMinLookAhead,
                                                             kMaxTravelTime);
    pareto_set ping_results;
    ++c.s.n_epochs;
    ping.execute(start_time, q.max_transfers_, worst_time_at_dest, ping_results);
    if (ping_results.els_.empty()) break;
    auto& pr = ping_results.els_;
    pr.erase(std::remove_if(pr.begin(), pr.end(),
                            [&](journey_t const& x) {  // pareto_set.h:17-19
                              return std::any_of(results.els_.begin(), results.els_.end(),
                                                 [&](journey_t const& y) { return y.dominates(x); });
                            }),
             pr.end());
    if (pr.empty()) break;
    // transfers are unique among one ping's journeys (equal start times)
    std::sort(pr.begin(), pr.end(),
              [](journey_t const& a, journey_t const& b) { return a.transfers_ > b.transfers_; });

    ping.fill_bounds(pr.front().transfers_ + std::size_t{1U});
    flip_dir(q);
    pong.reset_arrivals();
    auto q_pong_starts = q;  // start footpaths unless the pong starts intermodal
    q_pong_starts.use_start_footpaths_ = !q.start_intermodal_;
    for (auto& ping_j : pr) {
      pong.set_bounds(ping_j.transfers_ + 1U);
      starts.clear();
      get_starts(!Fwd, tt, false, ping_j.dest_time_, search_interval, q_pong_starts, starts, false);
      pong.next_start_time();
      for (auto const& st : starts) pong.add_start(st.stop_, st.time_at_stop_);
      ++c.s.n_epochs;
      pong.execute(ping_j.dest_time_, ping_j.transfers_, ping_j.start_time_, results);
      auto const match = std::find_if(results.els_.begin(), results.els_.end(),
                                      [&](journey_t const& pong_j) {
                                        return pong_j.transfers_ == ping_j.transfers_ &&
                                               pong_j.start_time_ == ping_j.dest_time_;
                                      });
      if (match == results.els_.end()) throw std::runtime_error("no pong");
      if (!match->is_reconstructed_ && !match->error_) {
        reconstructor<!Fwd>{tt, q, rs, base_day, *match}.run();
      }
      ping_j.start_time_ = match->dest_time_;
    }
    flip_dir(q);

    auto const first = std::min_element(pr.begin(), pr.end(), [&](journey_t const& a, journey_t const& b) {
      return is_better(a.start_time_, b.start_time_);
    });
    auto const next = first->start_time_ + (Fwd ? 1 : -1);
    if (!is_better(start_time, next)) throw std::runtime_error("no pong progress");
    start_time = next;
  }

  // pong.cc:397-428
  auto& e = results.els_;
  e.erase(std::remove_if(e.begin(), e.end(),
                         [&](journey_t const& j) {
                           auto const j_start_time = j.dest_time_;
                           auto const out_of_interval =
                               Fwd ? !q.extend_later_ && j_start_time >= search_interval.to_
                                   : !q.extend_earlier_ && j_start_time < search_interval.from_;
                           return !j.is_reconstructed_ || !is_validated(j) || out_of_interval ||
                                  j.travel_time() >= fastest_direct ||
                                  j.travel_time() > q.max_travel_time_;
                         }),
          e.end());
  for (auto& x : e) std::swap(x.start_time_, x.dest_time_);
  out.interval_ = {Fwd ? search_interval.from_ : start_time + 1,
                   Fwd ? start_time : search_interval.to_};
  out.base_ = base_day;
  out.fastest_direct_ = fastest_direct;
  auto const swap_special = [](std::uint32_t l) {
    return l == kStart ? kEnd : (l == kEnd ? kStart : l);
  };
  for (auto& j : e) {
    j.legs_.front().from_ = swap_special(j.legs_.front().from_);
    j.legs_.back().to_ = swap_special(j.legs_.back().to_);
  }
  // (start_time, transfers) is unique in a Pareto set
  std::sort(e.begin(), e.end(), [](journey_t const& a, journey_t const& b) {
    return std::tie(a.start_time_, a.transfers_) < std::tie(b.start_time_, b.transfers_);
  });
}

int run(nigiri_tt_view const& v, nigiri_query const& cq, raptor_state& rs,
        search_result& out, counters& c, std::vector<std::uint16_t>* lb,
        std::vector<delta_t>* tad) {
  tt_t const tt{v};
  auto const q = from_c(cq);
  if (q.fwd_) {
    run_search<true>(tt, q, rs, out, c, lb, tad);
  } else {
    run_search<false>(tt, q, rs, out, c, lb, tad);
  }
  return 0;
}

}  // namespace oracle

using namespace oracle;

extern "C" {

int oracle_search(const nigiri_tt_view* tt, const nigiri_query* q, nigiri_journey* out,
                  uint32_t cap, uint32_t* n_out, nigiri_leg* legs, uint32_t leg_cap,
                  uint32_t* n_legs, int16_t* round_times, int16_t* best, int16_t* tmp,
                  int16_t* time_at_dest, uint16_t* lb, int32_t* info, oracle_stats* stats) {
  try {
    raptor_state rs;
    search_result res;
    counters c;
    std::vector<std::uint16_t> lbv;
    std::vector<delta_t> tad;
    run(*tt, *q, rs, res, c, &lbv, &tad);
    auto const& e = res.results_.els_;
    if (n_out != nullptr) *n_out = static_cast<uint32_t>(e.size());
    std::uint32_t leg_pos = 0U;
    for (std::size_t i = 0; i < e.size() && i < cap; ++i) {
      out[i] = nigiri_journey{e[i].start_time_, e[i].dest_time_, e[i].dest_, e[i].transfers_, {},
                              leg_pos, 0U};
      if (legs == nullptr) continue;
      for (auto const& l : e[i].legs_) {
        if (leg_pos < leg_cap) {
          legs[leg_pos] = nigiri_leg{l.kind_, l.from_, l.to_, l.dep_time_, l.arr_time_,
                                     l.transport_, l.route_, l.day_, l.stop_from_, l.stop_to_,
                                     l.duration_, l.target_};
        }
        ++leg_pos;
      }
      out[i].n_legs = static_cast<uint32_t>(e[i].legs_.size());
    }
    if (n_legs != nullptr) *n_legs = leg_pos;
    auto const L = tt->n_locations;
    if (round_times != nullptr && !rs.round_times_.empty()) {
      std::memcpy(round_times, rs.round_times_.data(), sizeof(int16_t) * L * kRows);
    }
    if (best != nullptr && !rs.best_.empty()) std::memcpy(best, rs.best_.data(), sizeof(int16_t) * L);
    if (tmp != nullptr && !rs.tmp_.empty()) std::memcpy(tmp, rs.tmp_.data(), sizeof(int16_t) * L);
    if (time_at_dest != nullptr && tad.size() == kRows) {
      std::memcpy(time_at_dest, tad.data(), sizeof(int16_t) * kRows);
    }
    if (lb != nullptr && lbv.size() == L) std::memcpy(lb, lbv.data(), sizeof(uint16_t) * L);
    if (info != nullptr) {
      info[0] = res.base_;
      info[1] = res.fastest_direct_;
      info[2] = res.interval_.from_;
      info[3] = res.interval_.to_;
    }
    if (stats != nullptr) *stats = c.s;
    return 0;
  } catch (std::exception const&) {
    return -1;
  }
}

int oracle_pong(const nigiri_tt_view* v, const nigiri_query* cq, nigiri_journey* out,
                uint32_t cap, uint32_t* n_out, nigiri_leg* legs, uint32_t leg_cap,
                uint32_t* n_legs, int32_t* info, oracle_stats* stats) {
  try {
    tt_t const tt{*v};
    auto const q = from_c(*cq);
    raptor_state rs;
    search_result res;
    counters c;
    if (q.fwd_) {
      run_pong<true>(tt, q, rs, res, c);
    } else {
      run_pong<false>(tt, q, rs, res, c);
    }
    auto const& e = res.results_.els_;
    if (n_out != nullptr) *n_out = static_cast<uint32_t>(e.size());
    std::uint32_t leg_pos = 0U;
    for (std::size_t i = 0; i < e.size() && i < cap; ++i) {
      out[i] = nigiri_journey{e[i].start_time_, e[i].dest_time_, e[i].dest_, e[i].transfers_, {},
                              leg_pos, static_cast<uint32_t>(e[i].legs_.size())};
      for (auto const& l : e[i].legs_) {
        if (legs != nullptr && leg_pos < leg_cap) {
          legs[leg_pos] = nigiri_leg{l.kind_, l.from_, l.to_, l.dep_time_, l.arr_time_,
                                     l.transport_, l.route_, l.day_, l.stop_from_, l.stop_to_,
                                     l.duration_, l.target_};
        }
        ++leg_pos;
      }
    }
    if (n_legs != nullptr) *n_legs = leg_pos;
    if (info != nullptr) {
      info[0] = res.base_;
      info[1] = res.fastest_direct_;
      info[2] = res.interval_.from_;
      info[3] = res.interval_.to_;
    }
    if (stats != nullptr) *stats = c.s;
    return 0;
  } catch (std::exception const&) {
    return -1;
  }
}

int oracle_optimize_footpaths(const nigiri_tt_view* v, const nigiri_query* cq, nigiri_leg* legs,
                              uint32_t n) {
  try {
    tt_t const tt{*v};
    auto q = from_c(*cq);
    sanitize(q);
    journey_t j{0, 0, 0U, 0U};
    for (auto i = 0U; i != n; ++i) {
      auto const& c = legs[i];
      leg_t l;
      l.kind_ = c.kind;
      l.from_ = c.from;
      l.to_ = c.to;
      l.dep_time_ = c.dep_time;
      l.arr_time_ = c.arr_time;
      l.transport_ = c.transport;
      l.route_ = c.route;
      l.day_ = c.day;
      l.stop_from_ = c.stop_from;
      l.stop_to_ = c.stop_to;
      l.target_ = c.target;
      l.duration_ = c.duration;
      j.legs_.push_back(l);
    }
    optimize_footpaths(tt, q, j);
    for (auto i = 0U; i != n; ++i) {
      auto const& l = j.legs_[i];
      legs[i] = nigiri_leg{l.kind_, l.from_, l.to_, l.dep_time_, l.arr_time_, l.transport_,
                           l.route_, l.day_, l.stop_from_, l.stop_to_, l.duration_, l.target_};
    }
    return 0;
  } catch (std::exception const&) {
    return -1;
  }
}

int oracle_lb(const nigiri_tt_view* v, const nigiri_query* cq, uint16_t* lb) {
  try {
    tt_t const tt{*v};
    auto q = from_c(*cq);
    sanitize(q);
    std::vector<std::uint16_t> d;
    dijkstra_lb(tt, q, q.fwd_ ? v->lb_fwd_offsets : v->lb_bwd_offsets,
                q.fwd_ ? v->lb_fwd : v->lb_bwd, d);
    std::memcpy(lb, d.data(), sizeof(uint16_t) * d.size());
    return 0;
  } catch (std::exception const&) {
    return -1;
  }
}

int oracle_fastest_direct(const nigiri_tt_view* v, const nigiri_query* cq, int16_t* out) {
  try {
    tt_t const tt{*v};
    auto q = from_c(*cq);
    sanitize(q);
    *out = get_fastest_direct(tt, q, q.fwd_);
    return 0;
  } catch (std::exception const&) {
    return -1;
  }
}

int oracle_starts(const nigiri_tt_view* v, const nigiri_query* cq, int32_t* time_at_start,
                  int32_t* time_at_stop, uint32_t* stop, uint32_t cap, uint32_t* n_out) {
  try {
    tt_t const tt{*v};
    auto q = from_c(*cq);
    sanitize(q);
    std::vector<start_t> starts;
    auto const iv = q.is_interval_ ? interval_t{q.iv_from_, q.iv_to_}
                                   : interval_t{q.start_time_, q.start_time_};
    get_starts(q.fwd_, tt, q.is_interval_, q.start_time_, iv, q, starts, true);
    auto const fwd = q.fwd_;
    std::sort(starts.begin(), starts.end(),
              [&](start_t const& a, start_t const& b) { return fwd ? b < a : a < b; });
    *n_out = static_cast<uint32_t>(starts.size());
    for (std::size_t i = 0; i < starts.size() && i < cap; ++i) {
      time_at_start[i] = starts[i].time_at_start_;
      time_at_stop[i] = starts[i].time_at_stop_;
      stop[i] = starts[i].stop_;
    }
    return 0;
  } catch (std::exception const&) {
    return -1;
  }
}

int oracle_search_many(const nigiri_tt_view* tt, const nigiri_query* qs, uint32_t n,
                       int threads, double* latency_ms, double* wall_ms,
                       uint32_t* n_journeys, oracle_stats* total) {
  if (threads < 1) threads = 1;
  std::atomic<uint32_t> next{0U};
  std::vector<counters> per_thread(static_cast<std::size_t>(threads));
  std::atomic<int> failed{0};
  auto const t0 = std::chrono::steady_clock::now();
  std::vector<std::thread> pool;
  for (int th = 0; th != threads; ++th) {
    pool.emplace_back([&, th]() {
      raptor_state rs;
      search_result res;
      bool warmed = false;
      while (true) {
        auto const i = next.fetch_add(1U);
        if (i >= n) break;
        try {
          if (!warmed) {
            counters scratch;
            run(*tt, qs[i], rs, res, scratch, nullptr, nullptr);
            warmed = true;
          }
          auto const s = std::chrono::steady_clock::now();
          run(*tt, qs[i], rs, res, per_thread[static_cast<std::size_t>(th)], nullptr, nullptr);
          auto const e = std::chrono::steady_clock::now();
          if (latency_ms != nullptr) {
            latency_ms[i] = std::chrono::duration<double, std::milli>(e - s).count();
          }
          if (n_journeys != nullptr) n_journeys[i] = static_cast<uint32_t>(res.results_.els_.size());
        } catch (std::exception const&) {
          failed.fetch_add(1);
        }
      }
    });
  }
  for (auto& t : pool) t.join();
  auto const t1 = std::chrono::steady_clock::now();
  if (wall_ms != nullptr) *wall_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  if (total != nullptr) {
    oracle_stats s{};
    for (auto const& c : per_thread) {
      s.n_routes_visited += c.s.n_routes_visited;
      s.n_route_stops_visited += c.s.n_route_stops_visited;
      s.n_earliest_trip_calls += c.s.n_earliest_trip_calls;
      s.n_events_scanned += c.s.n_events_scanned;
      s.n_route_day_tests += c.s.n_route_day_tests;
      s.n_transport_day_tests += c.s.n_transport_day_tests;
      s.n_marked_stations += c.s.n_marked_stations;
      s.n_location_routes_read += c.s.n_location_routes_read;
      s.n_transfer_stations += c.s.n_transfer_stations;
      s.n_footpaths_visited += c.s.n_footpaths_visited;
      s.n_label_writes += c.s.n_label_writes;
      s.n_rounds += c.s.n_rounds;
      s.n_epochs += c.s.n_epochs;
      s.alg_bytes += c.s.alg_bytes;
    }
    *total = s;
  }
  return failed.load() == 0 ? 0 : -1;
}

}  // extern "C"
