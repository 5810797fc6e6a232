// Range search by pong on the plain engine (nigiri's routing::pong_search,
// include/nigiri/routing/raptor/pong.h:22-30, src/routing/raptor/pong.cc:
// 62-428), driven through the Algo entry points as pong.cc drives raptor<>.
//
// Per start time (the search direction's end of the interval first): a ping
// search from the query's starts finds the journeys leaving at that time;
// for each of them, most transfers first, a pong search in the other
// direction from its arrival, pruned by the ping's round times (bounds),
// finds the latest departure with the same number of transfers and is
// reconstructed.  The next ping starts one minute after the earliest such
// departure.  The ping runs on slot ping_slot, the pong on ping_slot + 1;
// fill_bounds copies the ping's round times into the pong slot's bounds,
// which replaces the state both share in the reference.
//
// Left out, as in the CPU restatement: enrich_with_slow_direct (a no-op,
// query::slow_direct_ is false, direct.cc:614-616) and the rewrite of the
// legs between three transit legs by earlier alternatives (pong.cc:430-470),
// which changes legs only.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <stdexcept>
#include <tuple>
#include <vector>

#include "capi_util.h"
#include "engine_access.h"
#include "search_host.h"

namespace nigiri_amd {
namespace {

constexpr std::int32_t kMinLookAhead = 1440;      // raptor/pong.h:11
constexpr std::int32_t kMaxTravelTimeMin = 7200;  // limits.h:10
constexpr std::uint32_t kStart = 0U;              // special_station::kStart

template <typename T>
T adjusted(host_tts const& s, T const d) {  // transfer_time_settings.h:21-31
  if (s.default_) return d;
  return static_cast<T>(static_cast<T>(s.additional_time_) +
                        std::max(static_cast<T>(s.min_transfer_time_),
                                 static_cast<T>(static_cast<float>(d) * s.factor_)));
}

void check(int rc) {
  if (rc != NIGIRI_AMD_OK) throw status_error{rc, nigiri_amd_last_error()};
}

// pareto_set::add (pareto_set.h:21-37)
void pareto_add(std::vector<host_journey>& set, host_journey const& el) {
  std::size_t n_removed = 0U;
  for (std::size_t i = 0U; i < set.size(); ++i) {
    if (set[i].dominates(el)) return;
    if (el.dominates(set[i])) {
      ++n_removed;
      continue;
    }
    set[i - n_removed] = set[i];
  }
  set.resize(set.size() - n_removed + 1U);
  set.back() = el;
}

// get_starts for a time point (start_times.cc:241-295, the unixtime_t
// case): the shortest offset per location, with footpaths from the offset
// locations when use_start_footpaths; transfer-time settings as given.
std::vector<std::pair<std::uint32_t, std::int32_t>> ontrip_starts(
    timetable const& tt, bool fwd, std::uint32_t const* loc, std::int16_t const* dur, std::uint32_t n,
    bool use_start_footpaths, host_tts const& tts, std::int32_t t) {
  std::map<std::uint32_t, std::int16_t> shortest;
  auto const update = [&](std::uint32_t l, std::int16_t d) {
    auto const it = shortest.try_emplace(l, d).first;
    it->second = std::min(it->second, d);
  };
  auto const& fo = fwd ? tt.fp_out_offsets : tt.fp_in_offsets;
  auto const& fps = fwd ? tt.fp_out : tt.fp_in;
  for (auto i = 0U; i != n; ++i) {
    update(loc[i], dur[i]);
    if (!use_start_footpaths) continue;
    for (auto e = fo[loc[i]]; e != fo[loc[i] + 1U]; ++e) {
      update(fp_target(fps[e]),
             static_cast<std::int16_t>(dur[i] + adjusted<std::int16_t>(
                                                     tts, static_cast<std::int16_t>(fp_duration(fps[e])))));
    }
  }
  std::vector<std::pair<std::uint32_t, std::int32_t>> out;
  for (auto const& [l, o] : shortest) out.emplace_back(l, fwd ? t + o : t - o);
  return out;
}

struct slot_search {
  nigiri_amd_engine* e;
  std::uint32_t slot;
  bool fwd;
  std::int32_t base_min;
  std::vector<std::uint32_t> dest_bits;  // ascending (raptor.h:393-413)
  std::vector<std::int16_t> rows;
  std::uint32_t n_rounds = 0U, n_executes = 0U;

  // next_start_time + add_start + execute + journey extraction into `into`
  void run(std::vector<std::pair<std::uint32_t, std::int32_t>> const& starts, std::int32_t start_time,
           std::uint8_t max_transfers, std::int32_t worst_time_at_dest, std::vector<host_journey>& into) {
    check(nigiri_amd_algo_next_start_time(e, slot));
    for (auto const& [l, t] : starts) check(nigiri_amd_algo_add_start(e, slot, l, t));
    auto const K = engine_max_k(e);
    rows.assign(std::max<std::size_t>(1U, dest_bits.size()) * K, 0);
    std::uint32_t n = 0U;
    check(nigiri_amd_algo_execute(e, slot, start_time, max_transfers, worst_time_at_dest, rows.data(), &n));
    n_rounds += n;
    ++n_executes;
    auto const end_k = std::min<std::uint32_t>(std::min<std::uint32_t>(max_transfers, 14U) + 2U, K);
    std::int16_t const invalid = fwd ? 32767 : -32768;
    for (std::size_t j = 0U; j != dest_bits.size(); ++j) {
      for (std::uint32_t k = 1U; k < end_k; ++k) {
        auto const v = rows[j * K + k];
        if (v == invalid) continue;
        pareto_add(into, host_journey{start_time, base_min + v, dest_bits[j], static_cast<std::uint8_t>(k - 1U)});
      }
    }
  }
};

nigiri_query flipped(nigiri_query const& q) {  // query::flip_dir (query.cc:52-57)
  auto f = q;
  f.dir = q.dir == NIGIRI_AMD_DIR_FWD ? NIGIRI_AMD_DIR_BWD : NIGIRI_AMD_DIR_FWD;
  std::swap(f.n_start, f.n_dest);
  std::swap(f.start_loc, f.dest_loc);
  std::swap(f.start_dur, f.dest_dur);
  std::swap(f.start_match_mode, f.dest_match_mode);
  f.use_start_footpaths = f.start_match_mode != NIGIRI_AMD_MATCH_INTERMODAL ? 1 : 0;
  return f;
}

nigiri_algo_init init_of(nigiri_query const& q, prepared_query const& p, std::uint16_t base,
                         std::uint16_t const* lb) {
  nigiri_algo_init a{};
  a.dir = q.dir;
  a.base_day = base;
  a.n_dest = static_cast<std::uint32_t>(p.dests_.size());
  a.dest = p.dests_.data();
  a.intermodal = p.intermodal_dest_ ? 1 : 0;
  a.n_dist = static_cast<std::uint32_t>(p.dist_loc_.size());
  a.dist_loc = p.dist_loc_.data();
  a.dist_min = p.dist_min_.data();
  a.lb = lb;
  a.tts_default = q.tts_default;  // as given: pong hands the query's settings to both raptors
  a.tts_min_transfer_time = q.tts_min_transfer_time;
  a.tts_additional_time = q.tts_additional_time;
  a.tts_factor = q.tts_factor;
  return a;
}

}  // namespace
}  // namespace nigiri_amd

using namespace nigiri_amd;

extern "C" {

int nigiri_amd_pong(nigiri_amd_engine* e, const nigiri_query* q, uint32_t ping_slot, nigiri_result* res,
                    nigiri_journey* journeys, uint32_t journey_capacity, nigiri_leg* legs,
                    uint32_t leg_capacity) {
  if (e == nullptr || q == nullptr || res == nullptr || (journeys == nullptr && journey_capacity != 0U) ||
      (legs == nullptr && leg_capacity != 0U)) {
    set_error("pong: bad argument");
    return NIGIRI_AMD_EINVAL;
  }
  *res = nigiri_result{};
  res->slot = -1;
  auto const rc = guarded([&]() {
    auto const t0 = std::chrono::steady_clock::now();
    if (engine_lanes(e) != 1U) {
      throw status_error{NIGIRI_AMD_EUNSUPPORTED, "pong: bounded searches run on the plain engine (lanes == 1)"};
    }
    if (ping_slot + 1U >= engine_slots(e)) throw status_error{NIGIRI_AMD_EINVAL, "pong: needs slots ping_slot and ping_slot + 1"};
    auto const& tt = engine_timetable(e);
    auto const fwd = q->dir == NIGIRI_AMD_DIR_FWD;
    host_tts const raw{q->tts_default != 0, q->tts_min_transfer_time, q->tts_additional_time, q->tts_factor};

    // ping (pong.cc:83-151): destinations, lower bounds, fastest direct, base day
    host_scratch scratch;
    prepared_query p;
    prepare(tt, *q, p, scratch, true);
    auto const fastest_direct = p.fastest_direct_;
    auto const max_travel_time = p.max_travel_time_;
    auto const iv_from = q->is_interval != 0 ? q->interval_from : q->start_time;
    auto const iv_to = q->is_interval != 0 ? q->interval_to : q->start_time;
    auto const base = p.base_;
    auto const ai = init_of(*q, p, base, p.lb_.data());
    check(nigiri_amd_algo_init(e, ping_slot, &ai));

    // pong (pong.cc:153-206): the flipped query, lower bounds all zero
    auto const fq = flipped(*q);
    prepared_query fp;
    prepare(tt, fq, fp, scratch, false);
    std::vector<std::uint16_t> const zero_lb(tt.n_locations, 0U);
    auto const pi = init_of(fq, fp, base, zero_lb.data());
    check(nigiri_amd_algo_init(e, ping_slot + 1U, &pi));

    auto const base_min = tt.internal_from + static_cast<std::int32_t>(base) * 1440;
    slot_search ping{e, ping_slot, fwd, base_min, {}, {}};
    for (std::uint32_t j = 0; j != p.n_dest_bits(); ++j) ping.dest_bits.push_back(p.dest_bit(j));
    slot_search pong{e, ping_slot + 1U, !fwd, base_min, {}, {}};
    for (std::uint32_t j = 0; j != fp.n_dest_bits(); ++j) pong.dest_bits.push_back(fp.dest_bit(j));

    // pong.cc:211-395
    std::vector<host_journey> results;
    auto start_time = fwd ? iv_from : iv_to - 1;
    auto const end_time = fwd ? iv_to : iv_from - 1;
    auto const is_better = [&](std::int32_t a, std::int32_t b) { return fwd ? a < b : a > b; };
    auto const is_validated = [&](host_journey const& j) { return is_better(j.dest_time_, start_time); };
    auto const result_count = [&](bool include_too_slow) {
      return static_cast<std::uint32_t>(std::count_if(results.begin(), results.end(), [&](host_journey const& j) {
        return is_validated(j) &&
               (include_too_slow || (j.travel_time() < fastest_direct && j.travel_time() <= max_travel_time));
      }));
    };
    // is_timeout_reached (pong.cc:235-241), measured from the call's start
    // (pong.cc:83)
    auto const timeout_reached = [&]() {
      return q->has_timeout != 0 &&
             std::chrono::steady_clock::now() - t0 >= std::chrono::seconds{q->timeout_s};
    };
    std::vector<nigiri_leg> leg_buf(256U);
    while ((is_better(start_time, end_time) ||
            result_count(true) + result_count(false) < 2U * q->min_connection_count) &&
           start_time >= tt.date_range_from && start_time < tt.date_range_to && !timeout_reached()) {
      check(nigiri_amd_algo_reset_arrivals(e, ping.slot));
      auto const starts = ontrip_starts(tt, fwd, q->start_loc, q->start_dur, q->n_start,
                                        q->use_start_footpaths != 0, raw, start_time);
      auto const worst = start_time + (fwd ? 1 : -1) * std::min<std::int32_t>(max_travel_time + kMinLookAhead,
                                                                              kMaxTravelTimeMin);
      std::vector<host_journey> pr;
      ping.run(starts, start_time, q->max_transfers, worst, pr);
      if (pr.empty()) break;
      pr.erase(std::remove_if(pr.begin(), pr.end(),
                              [&](host_journey const& x) {  // pareto_set.h:17-19
                                return std::any_of(results.begin(), results.end(),
                                                   [&](host_journey const& y) { return y.dominates(x); });
                              }),
               pr.end());
      if (pr.empty()) break;
      // transfers are unique among one ping's journeys (equal start times)
      std::sort(pr.begin(), pr.end(),
                [](host_journey const& a, host_journey const& b) { return a.transfers_ > b.transfers_; });

      check(nigiri_amd_algo_fill_bounds(e, ping.slot, pong.slot, pr.front().transfers_ + 1U));
      check(nigiri_amd_algo_reset_arrivals(e, pong.slot));
      for (auto& ping_j : pr) {
        check(nigiri_amd_algo_set_bounds(e, pong.slot, ping_j.transfers_ + 1U));
        auto const pong_starts = ontrip_starts(tt, !fwd, fq.start_loc, fq.start_dur, fq.n_start,
                                               fq.use_start_footpaths != 0, raw, ping_j.dest_time_);
        pong.run(pong_starts, ping_j.dest_time_, ping_j.transfers_, ping_j.start_time_, results);
        auto const match = std::find_if(results.begin(), results.end(), [&](host_journey const& pong_j) {
          return pong_j.transfers_ == ping_j.transfers_ && pong_j.start_time_ == ping_j.dest_time_;
        });
        if (match == results.end()) throw status_error{NIGIRI_AMD_EDEVICE, "pong: no pong journey for a ping journey"};
        if (!match->reconstructed_ && !match->error_) {
          nigiri_journey nj{match->start_time_, match->dest_time_, match->dest_, match->transfers_, {}, 0U, 0U};
          check(nigiri_amd_algo_reconstruct(e, pong.slot, &fq, &nj, leg_buf.data(),
                                            static_cast<std::uint32_t>(leg_buf.size())));
          match->legs_.assign(leg_buf.begin(), leg_buf.begin() + nj.n_legs);
          match->reconstructed_ = true;
        }
        ping_j.start_time_ = match->dest_time_;
      }
      auto const first = std::min_element(pr.begin(), pr.end(), [&](host_journey const& a, host_journey const& b) {
        return is_better(a.start_time_, b.start_time_);
      });
      auto const next = first->start_time_ + (fwd ? 1 : -1);
      if (!is_better(start_time, next)) throw status_error{NIGIRI_AMD_EDEVICE, "pong: no progress"};
      start_time = next;
    }
    check(nigiri_amd_algo_set_bounds(e, pong.slot, 0U));

    // pong.cc:397-428
    results.erase(std::remove_if(results.begin(), results.end(),
                                 [&](host_journey const& j) {
                                   auto const j_start = j.dest_time_;
                                   auto const out_of_interval =
                                       fwd ? q->extend_interval_later == 0 && j_start >= iv_to
                                           : q->extend_interval_earlier == 0 && j_start < iv_from;
                                   return !j.reconstructed_ || !is_validated(j) || out_of_interval ||
                                          j.travel_time() >= fastest_direct || j.travel_time() > max_travel_time;
                                 }),
                  results.end());
    auto const swap_special = [](std::uint32_t l) { return l == kStart ? kEnd : (l == kEnd ? kStart : l); };
    for (auto& j : results) {
      std::swap(j.start_time_, j.dest_time_);
      j.legs_.front().from = swap_special(j.legs_.front().from);
      j.legs_.back().to = swap_special(j.legs_.back().to);
    }
    // (start_time, transfers) is unique in a Pareto set
    std::sort(results.begin(), results.end(), [](host_journey const& a, host_journey const& b) {
      return std::tie(a.start_time_, a.transfers_) < std::tie(b.start_time_, b.transfers_);
    });

    res->status = NIGIRI_AMD_OK;
    res->n_journeys = static_cast<std::uint32_t>(results.size());
    res->journey_offset = 0U;
    res->n_epochs = ping.n_executes + pong.n_executes;
    res->n_rounds = ping.n_rounds + pong.n_rounds;
    res->slot = static_cast<std::int32_t>(ping_slot);
    res->interval_from = fwd ? iv_from : start_time + 1;
    res->interval_to = fwd ? start_time : iv_to;
    res->fastest_direct = fastest_direct;
    res->base_day = static_cast<std::int16_t>(base);
    if (results.size() > journey_capacity) throw status_error{NIGIRI_AMD_ECAPACITY, "pong: journey_capacity too small"};
    std::uint32_t leg_pos = 0U;
    for (std::size_t i = 0; i != results.size(); ++i) {
      auto const& j = results[i];
      journeys[i] = nigiri_journey{j.start_time_, j.dest_time_, j.dest_, j.transfers_, {}, leg_pos, 0U};
      if (legs == nullptr) continue;
      if (leg_pos + j.legs_.size() > leg_capacity) throw status_error{NIGIRI_AMD_ECAPACITY, "pong: leg_capacity too small"};
      std::copy(j.legs_.begin(), j.legs_.end(), legs + leg_pos);
      journeys[i].n_legs = static_cast<std::uint32_t>(j.legs_.size());
      leg_pos += static_cast<std::uint32_t>(j.legs_.size());
    }
    res->latency_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return NIGIRI_AMD_OK;
  });
  if (rc != NIGIRI_AMD_OK) res->status = rc;
  return rc;
}

}  // extern "C"
