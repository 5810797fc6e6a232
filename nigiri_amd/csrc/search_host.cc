// Host half of the search driver.  Follows, in order:
//   raptor_search_with_dir (src/routing/raptor_search.cc:51-79): sanitize;
//   search::search (search.h:174-202): fastest direct, init();
//   search::init (search.h:84-172): transfer-time settings, destinations,
//     lower bounds (dijkstra.cc:18-67), base day;
//   search::execute (search.h:204-345): overlap test, start labels
//     (start_times.cc:241-319) sorted and grouped into epochs;
//   journey extraction / Pareto set (raptor.h:393-413, pareto_set.h:21-37).
#include "search_host.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <map>

#include "capi_util.h"

namespace nigiri_amd {

namespace {

constexpr std::uint16_t kUnreach = 0xFFFFU;
constexpr std::int16_t kUnreachableDirect = 32767;

template <typename T>
T adjusted(host_tts const& s, T const d) {  // transfer_time_settings.h:21-31
  if (s.default_) return d;
  return static_cast<T>(static_cast<T>(s.additional_time_) +
                        std::max(static_cast<T>(s.min_transfer_time_),
                                 static_cast<T>(static_cast<float>(d) * s.factor_)));
}

std::uint16_t event(timetable const& tt, std::uint32_t r, std::uint32_t t_off,
                    std::uint32_t stop_idx, bool arr) {
  auto const n_t = tt.route_transport_from[r + 1] - tt.route_transport_from[r];
  return tt.event_times[tt.route_stop_time_from[r] +
                        n_t * (stop_idx * 2U - (arr ? 1U : 0U)) + t_off];
}

bool bit_test(timetable const& tt, std::uint32_t bf, std::uint32_t day) {
  if (day >= 512U) return false;  // out-of-range day index: not active
  return ((tt.bitfields[bf * kBitfieldWords + day / 64U] >> (day % 64U)) & 1U) != 0U;
}

// Exact single-source shortest travel-time lower bounds on the lb graph
// with the dial cut-off of 7201 minutes (dijkstra.h:31-70, dial.h:9-78).
void lower_bounds(timetable const& tt, bool fwd,
                  std::vector<std::pair<std::uint32_t, std::uint16_t>> const& dests,
                  std::vector<std::uint16_t>& dist, host_scratch& s) {
  auto const& off = fwd ? tt.lb_fwd_offsets : tt.lb_bwd_offsets;
  auto const& edges = fwd ? tt.lb_fwd : tt.lb_bwd;
  dist.assign(tt.n_locations, kUnreach);
  constexpr std::size_t n_buckets = kMaxTravelTime + 1;
  if (s.buckets_.size() != n_buckets) s.buckets_.assign(n_buckets, {});
  std::size_t size = 0U, cur = n_buckets;
  for (auto const& [l, d] : dests) {
    if (d < dist[l]) dist[l] = d;
  }
  for (auto const& [l, d] : dests) {
    if (dist[l] == d && d < n_buckets) {
      s.buckets_[d].push_back(l);
      ++size;
      cur = std::min<std::size_t>(cur, d);
    }
  }
  while (size != 0U) {
    while (s.buckets_[cur].empty()) ++cur;
    auto const l = s.buckets_[cur].back();
    s.buckets_[cur].pop_back();
    --size;
    if (dist[l] < cur) continue;
    for (auto e = off[l]; e != off[l + 1U]; ++e) {
      auto const t = fp_target(edges[e]);
      auto const nd = cur + fp_duration(edges[e]);
      if (nd < dist[t] && nd < n_buckets) {
        dist[t] = static_cast<std::uint16_t>(nd);
        s.buckets_[nd].push_back(t);
        ++size;
      }
    }
  }
}

// get_fastest_direct.cc:12-100: direct walking from start offsets to
// destination offsets; footpaths are only followed into destinations.
std::int16_t fastest_direct(timetable const& tt, nigiri_query const& q, bool fwd,
                            host_tts const& raw_tts) {
  if (q.has_fastest_direct != 0) return q.fastest_direct;
  std::map<std::uint32_t, std::uint32_t> dist;  // location -> minutes
  for (auto i = 0U; i != q.n_start; ++i) {
    auto const it = dist.try_emplace(q.start_loc[i], 32767U).first;
    it->second = std::min<std::uint32_t>(it->second, static_cast<std::uint16_t>(q.start_dur[i]));
  }
  std::map<std::uint32_t, std::uint32_t> dest_off;
  for (auto i = 0U; i != q.n_dest; ++i) {
    auto const d = static_cast<std::uint16_t>(q.dest_dur[i]);
    auto const it = dest_off.try_emplace(q.dest_loc[i], d).first;
    it->second = std::min<std::uint32_t>(it->second, d);
  }
  // Dijkstra restricted to footpaths that end in a destination
  std::multimap<std::uint32_t, std::uint32_t> pq;
  for (auto const& [l, d] : dist) pq.emplace(d, l);
  auto const& off = fwd ? tt.fp_out_offsets : tt.fp_in_offsets;
  auto const& fps = fwd ? tt.fp_out : tt.fp_in;
  std::uint32_t end_dist = kUnreachableDirect;
  while (!pq.empty()) {
    auto const [d, l] = *pq.begin();
    pq.erase(pq.begin());
    if (l == kEnd) break;
    if (dist[l] < d) continue;
    for (auto e = off[l]; e != off[l + 1U]; ++e) {
      auto const nd = d + adjusted<std::uint16_t>(raw_tts, static_cast<std::uint16_t>(fp_duration(fps[e])));
      if (nd > 65535U) continue;
      auto const t = fp_target(fps[e]);
      if (dest_off.find(t) == dest_off.end()) continue;
      auto const it = dist.try_emplace(t, 32767U).first;
      if (nd < it->second && nd < 65535U) {
        it->second = nd;
        pq.emplace(nd, t);
      }
    }
    if (auto const it = dest_off.find(l); it != dest_off.end()) {
      auto const nd = d + it->second;
      if (nd < 65535U && nd < end_dist) end_dist = nd;
    }
  }
  return end_dist == static_cast<std::uint32_t>(kUnreachableDirect)
             ? kUnreachableDirect
             : static_cast<std::int16_t>(std::ceil(end_dist * q.fastest_direct_factor));
}

std::int64_t round_to_days(std::int64_t minutes) {  // std::chrono::round<days>
  auto const fl = minutes >= 0 ? minutes / 1440 : -((-minutes + 1439) / 1440);
  auto const lo = minutes - fl * 1440, hi = (fl + 1) * 1440 - minutes;
  if (lo < hi) return fl;
  if (lo > hi) return fl + 1;
  return (fl % 2 == 0) ? fl : fl + 1;
}

struct start_label {
  std::int32_t time_at_start, time_at_stop;
  std::uint32_t stop;
  bool operator<(start_label const& o) const {
    return std::tie(time_at_start, time_at_stop, stop) <
           std::tie(o.time_at_start, o.time_at_stop, o.stop);
  }
};

// start_times.cc:40-121
void starts_at_stop(timetable const& tt, bool fwd, std::uint32_t r, std::uint32_t stop_idx,
                    std::uint32_t l, std::int32_t iv_from, std::int32_t iv_to,
                    std::int32_t stop_from, std::int32_t stop_to, std::int16_t offset,
                    std::vector<start_label>& starts) {
  auto const day_of = [&](std::int32_t t) {
    return static_cast<std::uint16_t>((t - tt.internal_from) / 1440);
  };
  auto const first_day = day_of(stop_from), last_day = day_of(stop_to);
  auto const t0 = tt.route_transport_from[r], t1 = tt.route_transport_from[r + 1U];
  for (auto t = t0; t != t1; ++t) {
    auto const ev = delta_count(event(tt, r, t - t0, stop_idx, !fwd));
    auto const day_offset = static_cast<std::uint16_t>(ev / 1440);
    auto const mam = ev % 1440;
    for (std::uint32_t day = first_day; day <= last_day; ++day) {
      auto const ev_time = tt.internal_from + static_cast<std::int32_t>(day) * 1440 + mam;
      if (!bit_test(tt, tt.transport_traffic_days[t],
                    static_cast<std::uint16_t>(day - day_offset)) ||
          !(ev_time >= stop_from && ev_time < stop_to)) {
        continue;
      }
      if (offset == static_cast<std::int16_t>(kFpMaxDuration)) continue;
      auto const at_start = fwd ? ev_time - offset : ev_time + offset;
      if (!(at_start >= iv_from && at_start < iv_to)) continue;
      if (!starts.empty() && starts.back().time_at_start == at_start &&
          (fwd ? starts.back().time_at_stop <= ev_time : starts.back().time_at_stop >= ev_time)) {
        continue;
      }
      starts.push_back({at_start, ev_time, l});
    }
  }
}

// start labels of the interval [from, to) (get_starts, start_times.cc:241-319;
// the start map is taken in ascending location order)
void interval_labels(timetable const& tt, prepared_query const& p, std::int32_t from,
                     std::int32_t to, bool add_ontrip, std::vector<start_label>& starts) {
  auto const fwd = p.fwd_;
  for (auto const& [l, o] : p.start_map_) {
    auto const stop_from = fwd ? from : from - p.max_start_offset_;
    auto const stop_to = fwd ? to + p.max_start_offset_ : to;
    for (auto j = tt.location_route_offsets[l]; j != tt.location_route_offsets[l + 1U]; ++j) {
      auto const r = tt.location_routes[j];
      auto const s0 = tt.route_stop_offsets[r];
      auto const n = tt.route_stop_offsets[r + 1U] - s0;
      for (auto i = 0U; i != n; ++i) {
        auto const stp = tt.route_stops[s0 + i];
        if ((stp & kStopLocMask) != l) continue;
        auto const in = ((stp >> 28) & 1U) != 0U, out = ((stp >> 29) & 1U) != 0U;
        if ((!fwd && (i == 0U || !out)) || (fwd && (i == n - 1U || !in))) continue;
        starts_at_stop(tt, fwd, r, i, l, from, to, stop_from, stop_to, o, starts);
      }
    }
    // ontrip label at the interval end (start_times.cc:225-238)
    if (add_ontrip && o != static_cast<std::int16_t>(kFpMaxDuration)) {
      auto const at_start = fwd ? to : from - 1;
      starts.push_back({at_start, fwd ? to + o : from - 1 - o, l});
    }
  }
}

// sorted labels (search.h:408-410) grouped by start time into epochs
void append_epochs(prepared_query& p, std::vector<start_label>& starts, bool reset) {
  auto const fwd = p.fwd_;
  std::sort(starts.begin(), starts.end(), [&](start_label const& a, start_label const& b) {
    return fwd ? b < a : a < b;
  });
  for (std::size_t from = 0U; from != starts.size();) {
    auto to = from;
    while (to != starts.size() && starts[to].time_at_start == starts[from].time_at_start) ++to;
    epoch ep;
    ep.start_time_ = starts[from].time_at_start;
    ep.worst_time_at_dest_ =
        ep.start_time_ + (fwd ? 1 : -1) * (std::min(p.fastest_direct_, p.max_travel_time_) + 1);
    ep.reset_ = reset && from == 0U;
    for (auto i = from; i != to; ++i) {
      ep.loc_.push_back(starts[i].stop);
      ep.time_.push_back(starts[i].time_at_stop);
    }
    p.epochs_.push_back(std::move(ep));
    from = to;
  }
}

// interval_estimator (interval_estimate.h:12-156)
bool can_extend_earlier(prepared_query const& p, std::int32_t from) {
  return p.extend_earlier_ && from != p.ext_from_;
}
bool can_extend_later(prepared_query const& p, std::int32_t to) {
  return p.extend_later_ && to != p.ext_to_;
}
void clamp_interval(prepared_query const& p, std::int32_t& from, std::int32_t& to) {
  from = std::clamp(std::clamp(from, p.ext_from_, p.ext_to_), p.max_from_, p.max_to_);
  to = std::clamp(std::clamp(to, p.ext_from_, p.ext_to_), p.max_from_, p.max_to_);
  if (p.has_max_interval_) {  // interval_estimate.h:142-145
    from = std::clamp(from, p.max_iv_from_, p.max_iv_to_);
    to = std::clamp(to, p.max_iv_from_, p.max_iv_to_);
  }
}

unsigned results_in_interval(prepared_query const& p) {  // search.h:382-391
  if (p.ontrip_) return static_cast<unsigned>(p.results_.size());
  return static_cast<unsigned>(std::count_if(
      p.results_.begin(), p.results_.end(),
      [&](host_journey const& j) { return j.start_time_ >= p.iv_from_ && j.start_time_ < p.iv_to_; }));
}

}  // namespace

std::size_t dest_relaxations(timetable const& tt, bool fwd, bool intermodal,
                             std::vector<std::uint32_t> const& dests) {
  if (intermodal) return dests.size();
  auto const& off = fwd ? tt.fp_in_offsets : tt.fp_out_offsets;
  std::size_t n = 0U;
  for (auto const d : dests) n += 1U + (off[d + 1U] - off[d]);
  return n;
}


void early_termination(prepared_query& p, std::size_t epoch_idx) {
  if (p.min_connection_count_ == 0U || results_in_interval(p) < p.min_connection_count_) return;
  if (!((p.fwd_ && p.extend_earlier_ && !p.extend_later_) ||
        (!p.fwd_ && !p.extend_earlier_ && p.extend_later_))) {
    return;
  }
  auto const start_time = p.epochs_[epoch_idx].start_time_;
  if (p.fwd_) {
    p.iv_from_ = start_time;
  } else {
    p.iv_to_ = start_time + 1;
  }
  p.epochs_.resize(epoch_idx + 1U);
}

void mark_execute_start(prepared_query& p) {
  if (p.exec_started_) return;
  p.exec_started_ = true;
  p.exec_start_ = std::chrono::steady_clock::now();
}

bool timeout_reached(prepared_query const& p) {
  return p.has_timeout_ &&
         (!p.exec_started_ ||
          std::chrono::steady_clock::now() - p.exec_start_ >= std::chrono::seconds{p.timeout_s_});
}

bool wants_extension(prepared_query const& p) {
  return !p.results_.empty() && !p.ontrip_ && results_in_interval(p) < p.min_connection_count_ &&
         !timeout_reached(p);
}

bool extend_interval(timetable const& tt, prepared_query& p) {
  bool reset = false;
  while (wants_extension(p)) {
    // interval_estimator::extension (interval_estimate.h:83-107)
    auto const need = p.min_connection_count_ - results_in_interval(p);
    if (!p.extend_earlier_ && !p.extend_later_) return false;
    auto from = p.iv_from_, to = p.iv_to_;
    auto const ext = static_cast<std::int32_t>(static_cast<unsigned>(to - from) * need);
    if (can_extend_earlier(p, from) && can_extend_later(p, to)) {
      from -= ext / 2;
      to += ext / 2;
    } else {
      if (p.extend_earlier_) from -= ext;
      if (p.extend_later_) to += ext;
    }
    clamp_interval(p, from, to);
    if (from == p.iv_from_ && to == p.iv_to_) return false;
    std::vector<start_label> starts;
    auto const remove_outside = [&]() {  // remove_ontrip_results (search.h:413-417)
      auto& e = p.results_;
      e.erase(std::remove_if(e.begin(), e.end(),
                             [&](host_journey const& j) {
                               return !(j.start_time_ >= p.iv_from_ && j.start_time_ < p.iv_to_);
                             }),
              e.end());
    };
    if (from != p.iv_from_) {
      interval_labels(tt, p, from, p.iv_from_, !p.fwd_, starts);
      if (!p.fwd_) {
        reset = true;
        remove_outside();
      }
    }
    if (to != p.iv_to_) {
      interval_labels(tt, p, p.iv_to_, to, p.fwd_, starts);
      if (p.fwd_) {
        reset = true;
        remove_outside();
      }
    }
    p.iv_from_ = from;
    p.iv_to_ = to;
    if (!starts.empty()) {
      append_epochs(p, starts, reset);
      return true;
    }
  }
  return false;
}

void fastest_direct_nodes(timetable const& tt, nigiri_query const& q, std::vector<fd_node_host>& out) {
  std::map<std::uint32_t, fd_node_host> m;
  auto const at = [&](std::uint32_t l) -> fd_node_host& {
    if (l >= tt.n_locations) throw status_error{NIGIRI_AMD_EINVAL, "location out of range"};
    return m.try_emplace(l, fd_node_host{l, 32767U, 0xFFFFU, 0U, 0U}).first->second;
  };
  for (auto i = 0U; i != q.n_start; ++i) {
    auto& x = at(q.start_loc[i]);
    x.is_start = 1U;
    x.start = std::min(x.start, static_cast<std::uint16_t>(q.start_dur[i]));
  }
  for (auto i = 0U; i != q.n_dest; ++i) {
    auto& x = at(q.dest_loc[i]);
    auto const d = static_cast<std::uint16_t>(q.dest_dur[i]);
    x.off = x.is_dest != 0U ? std::min(x.off, d) : d;
    x.is_dest = 1U;
  }
  out.clear();
  for (auto const& [l, x] : m) out.push_back(x);
}

void prepare(timetable const& tt, nigiri_query const& q, prepared_query& p,
             host_scratch& s, bool host_lb, std::int16_t const* fastest_direct_given) {
  if (q.prf_idx != 0) {
    throw status_error{NIGIRI_AMD_EUNSUPPORTED, "profiles other than 0 are not supported"};
  }
  if (q.dir != NIGIRI_AMD_DIR_FWD && q.dir != NIGIRI_AMD_DIR_BWD) {
    throw status_error{NIGIRI_AMD_EINVAL, "bad direction"};
  }
  auto const start_im = q.start_match_mode == NIGIRI_AMD_MATCH_INTERMODAL;
  auto const dest_im = q.dest_match_mode == NIGIRI_AMD_MATCH_INTERMODAL;
  if (start_im && q.use_start_footpaths != 0) {  // query.cc:59-62
    throw status_error{NIGIRI_AMD_EINVAL,
                       "intermodal start is incompatible with use_start_footpaths"};
  }
  for (auto i = 0U; i != q.n_start; ++i) {
    if (q.start_loc[i] >= tt.n_locations || q.start_dur[i] < 0) {
      throw status_error{NIGIRI_AMD_EINVAL, "start offset out of range"};
    }
  }
  for (auto i = 0U; i != q.n_dest; ++i) {
    if (q.dest_loc[i] >= tt.n_locations || q.dest_dur[i] < 0 ||
        q.dest_dur[i] > kMaxTravelTime) {
      throw status_error{NIGIRI_AMD_EINVAL, "destination offset out of range"};
    }
  }
  auto const fwd = q.dir == NIGIRI_AMD_DIR_FWD;
  p = prepared_query{};
  p.fwd_ = fwd;
  p.ontrip_ = q.is_interval == 0;
  p.intermodal_dest_ = dest_im;
  p.max_transfers_ = q.max_transfers;
  p.end_k_ = static_cast<std::uint8_t>(std::min<int>(q.max_transfers, kMaxTransfers) + 2);
  p.max_travel_time_ = (q.max_travel_time < 0 || q.max_travel_time > kMaxTravelTime)
                           ? static_cast<std::int16_t>(kMaxTravelTime)
                           : q.max_travel_time;
  p.iv_from_ = p.ontrip_ ? q.start_time : q.interval_from;
  p.iv_to_ = p.ontrip_ ? q.start_time : q.interval_to;

  p.start_intermodal_ = start_im;
  for (auto i = 0U; i != q.n_start; ++i) p.start_offsets_.emplace_back(q.start_loc[i], q.start_dur[i]);
  for (auto i = 0U; i != q.n_dest; ++i) p.dest_offsets_.emplace_back(q.dest_loc[i], q.dest_dur[i]);
  auto const by_dur = [](auto const& a, auto const& b) { return a.second < b.second; };
  std::stable_sort(p.start_offsets_.begin(), p.start_offsets_.end(), by_dur);
  std::stable_sort(p.dest_offsets_.begin(), p.dest_offsets_.end(), by_dur);

  host_tts raw{q.tts_default != 0, q.tts_min_transfer_time, q.tts_additional_time, q.tts_factor};
  p.fastest_direct_ = fastest_direct_given != nullptr ? *fastest_direct_given : fastest_direct(tt, q, fwd, raw);

  // search::init: normalised transfer-time settings
  p.tts_ = raw;
  p.tts_.factor_ = std::max(p.tts_.factor_, 1.0F);
  p.tts_.min_transfer_time_ = std::max<std::int16_t>(p.tts_.min_transfer_time_, 0);
  p.tts_.additional_time_ = std::max<std::int16_t>(p.tts_.additional_time_, 0);
  p.tts_.default_ = p.tts_.factor_ == 1.0F && p.tts_.min_transfer_time_ == 0 &&
                    p.tts_.additional_time_ == 0;

  // collect_destinations (start_times.cc:321-353)
  std::vector<std::pair<std::uint32_t, std::uint16_t>> dest_min;
  {
    std::map<std::uint32_t, std::uint16_t> m;
    for (auto i = 0U; i != q.n_dest; ++i) {
      auto const it = m.try_emplace(q.dest_loc[i], static_cast<std::uint16_t>(q.dest_dur[i])).first;
      it->second = std::min(it->second, static_cast<std::uint16_t>(q.dest_dur[i]));
    }
    dest_min.assign(m.begin(), m.end());
  }
  for (auto const& [l, d] : dest_min) {
    if (dest_im) {
      p.dist_loc_.push_back(l);
      p.dist_min_.push_back(d);
    } else {
      p.dests_.push_back(l);
    }
  }
  if ((dest_im && p.dist_loc_.size() > kMaxQueryEgress) ||
      (!dest_im && p.dests_.size() > kMaxQueryDests)) {
    throw status_error{NIGIRI_AMD_EUNSUPPORTED, "too many destination offsets"};
  }
  if (dest_relaxations(tt, fwd, dest_im, dest_im ? p.dist_loc_ : p.dests_) > kMaxQueryDestEvents) {
    throw status_error{NIGIRI_AMD_EUNSUPPORTED, "destinations reachable by more than " +
                                                    std::to_string(kMaxQueryDestEvents) +
                                                    " transfers and footpaths in one round"};
  }
  p.lb_seeds_ = dest_min;
  if (host_lb) {
    lower_bounds(tt, fwd, dest_min, p.lb_, s);
  } else {
    p.lb_.clear();
    ch_sweep_seeds(tt.ch, fwd, dest_min, s.ch_, p.lb_up_);
  }

  // scheduling hint: queries that start close to each other mark the same
  // routes in their first rounds, so the batch scheduler groups them
  if (q.n_start != 0U) {
    auto const l = q.start_loc[0];
    std::uint32_t cx = l, cy = 0U;
    if (!tt.coord_x.empty()) {
      cx = static_cast<std::uint32_t>(std::max(0.0F, tt.coord_x[l]) / 256.0F) & 0xFFFFU;
      cy = static_cast<std::uint32_t>(std::max(0.0F, tt.coord_y[l]) / 256.0F) & 0xFFFFU;
    }
    std::uint64_t key = 0U;
    for (int b = 0; b != 16; ++b) {
      key |= static_cast<std::uint64_t>((cx >> b) & 1U) << (2 * b);
      key |= static_cast<std::uint64_t>((cy >> b) & 1U) << (2 * b + 1);
    }
    p.locality_key_ = key;
  }

  // base day: round<days>(interval midpoint) - internal interval start
  auto const mid = static_cast<std::int64_t>(p.iv_from_) + (p.iv_to_ - p.iv_from_) / 2;
  p.base_ = static_cast<std::uint16_t>((round_to_days(mid) * 1440 - tt.internal_from) / 1440);

  // start_dest_overlap (search.h:363-387)
  if (!start_im && !dest_im) {
    for (auto i = 0U; i != q.n_dest && !p.overlap_; ++i) {
      for (auto j = 0U; j != q.n_start; ++j) {
        if (q.start_loc[j] == q.dest_loc[i]) p.overlap_ = true;
      }
    }
  }
  if (p.overlap_) return;

  // get_starts (start_times.cc:241-319); start map in ascending location order
  std::map<std::uint32_t, std::int16_t> shortest;
  auto const update = [&](std::uint32_t l, std::int16_t d) {
    auto const it = shortest.try_emplace(l, d).first;
    it->second = std::min(it->second, d);
  };
  auto const& fo = fwd ? tt.fp_out_offsets : tt.fp_in_offsets;
  auto const& fps = fwd ? tt.fp_out : tt.fp_in;
  for (auto i = 0U; i != q.n_start; ++i) {
    auto const l = q.start_loc[i];
    update(l, q.start_dur[i]);
    if (q.use_start_footpaths != 0) {
      for (auto e = fo[l]; e != fo[l + 1U]; ++e) {
        update(fp_target(fps[e]),
               static_cast<std::int16_t>(q.start_dur[i] +
                                         adjusted<std::int16_t>(p.tts_, static_cast<std::int16_t>(fp_duration(fps[e])))));
      }
    }
  }
  p.start_map_.assign(shortest.begin(), shortest.end());
  p.max_start_offset_ = q.max_start_offset;

  // interval_estimator (interval_estimate.h:12-156) and its initial
  // widening of a pretrip interval (search.h:216-218)
  p.min_connection_count_ = q.min_connection_count;
  p.extend_earlier_ = q.extend_interval_earlier != 0;
  p.extend_later_ = q.extend_interval_later != 0;
  p.ext_from_ = tt.date_range_from;
  p.ext_to_ = tt.date_range_to;
  p.has_max_interval_ = q.has_max_interval != 0;
  p.max_iv_from_ = q.max_interval_from;
  p.max_iv_to_ = q.max_interval_to;
  p.has_timeout_ = q.has_timeout != 0;
  p.timeout_s_ = q.timeout_s;
  {
    constexpr std::int64_t kMaxSearchIntervalSize = (32767 / 1440 - 7) * 1440;  // limits.h:11-13
    auto const from = static_cast<std::int64_t>(p.iv_from_), to = static_cast<std::int64_t>(p.iv_to_);
    auto const days = q.has_fastest_direct != 0 ? 1 : 2;
    p.max_from_ = static_cast<std::int32_t>(
        std::max(round_to_days(from - days * 1440), round_to_days(from + (to - from - kMaxSearchIntervalSize) / 2)) * 1440);
    p.max_to_ = static_cast<std::int32_t>(
        std::min(round_to_days(to + days * 1440), round_to_days(from + (to - from + kMaxSearchIntervalSize) / 2)) * 1440);
  }
  if (!p.ontrip_ && p.min_connection_count_ != 0U && (p.extend_earlier_ || p.extend_later_)) {
    auto from = p.iv_from_, to = p.iv_to_;
    auto const ext = 60 * static_cast<std::int32_t>(p.min_connection_count_);
    if (fwd ? can_extend_later(p, to) : can_extend_earlier(p, from)) {
      if (fwd) {
        if (can_extend_earlier(p, from)) {
          from -= 60;
        } else {
          to += 60;
        }
        to += ext;
      } else {
        if (can_extend_later(p, to)) {
          to += 60;
        } else {
          from -= 60;
        }
        from -= ext;
      }
    } else if (fwd) {
      if (p.extend_earlier_) from -= 60 + ext;
    } else {
      if (p.extend_later_) to += 60 + ext;
    }
    clamp_interval(p, from, to);
    p.iv_from_ = from;
    p.iv_to_ = to;
  }

  std::vector<start_label> starts;
  if (p.ontrip_) {
    for (auto const& [l, o] : p.start_map_) {
      starts.push_back({q.start_time, fwd ? q.start_time + o : q.start_time - o, l});
    }
  } else {
    interval_labels(tt, p, p.iv_from_, p.iv_to_, true, starts);
  }
  append_epochs(p, starts, false);
}

void add_epoch_results(timetable const& tt, prepared_query& p, std::size_t epoch_idx,
                       std::int16_t const* dest_rows, std::uint32_t max_k) {
  auto const invalid = p.fwd_ ? static_cast<std::int16_t>(SHRT_MAX)
                              : static_cast<std::int16_t>(SHRT_MIN);
  auto const start_time = p.epochs_[epoch_idx].start_time_;
  auto const base = p.base_minutes(tt);
  auto& res = p.results_;
  for (auto j = 0U; j != p.n_dest_bits(); ++j) {
    for (auto k = 1U; k < p.end_k_ && k < max_k; ++k) {
      auto const v = dest_rows[j * max_k + k];
      if (v == invalid) continue;
      host_journey const el{start_time, base + v, p.dest_bit(j),
                            static_cast<std::uint8_t>(k - 1U), false};
      // pareto_set::add
      std::size_t n_removed = 0U;
      bool dominated = false;
      for (std::size_t i = 0U; i < res.size(); ++i) {
        if (res[i].dominates(el)) {
          dominated = true;
          break;
        }
        if (el.dominates(res[i])) {
          ++n_removed;
          continue;
        }
        res[i - n_removed] = res[i];
      }
      if (dominated) {
        // elements were only moved down for indices below the dominator,
        // which equals the identity when nothing was removed before it
        continue;
      }
      res.resize(res.size() - n_removed + 1U);
      res.back() = el;
    }
  }
}

void reconstruction_candidates(prepared_query const& p, std::vector<std::size_t>& out) {
  out.clear();
  for (std::size_t i = 0; i != p.results_.size(); ++i) {
    auto const& j = p.results_[i];
    if (!j.reconstructed_ && !j.error_ &&
        (p.ontrip_ || (j.start_time_ >= p.iv_from_ && j.start_time_ < p.iv_to_)) &&
        j.travel_time() < p.fastest_direct_) {
      out.push_back(i);
    }
  }
}

void finish(prepared_query& p) {
  auto& e = p.results_;
  if (!p.ontrip_) {
    e.erase(std::remove_if(e.begin(), e.end(),
                           [&](host_journey const& j) {
                             return !(j.start_time_ >= p.iv_from_ && j.start_time_ < p.iv_to_) ||
                                    j.travel_time() >= p.fastest_direct_ ||
                                    j.travel_time() > p.max_travel_time_;
                           }),
            e.end());
    std::sort(e.begin(), e.end(), [](host_journey const& a, host_journey const& b) {
      return std::tie(a.start_time_, a.transfers_) < std::tie(b.start_time_, b.transfers_);
    });
  }
  e.erase(std::remove_if(e.begin(), e.end(),
                         [](host_journey const& j) { return !j.reconstructed_; }),
          e.end());
}

}  // namespace nigiri_amd
