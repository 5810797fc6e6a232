// Derived timetable arrays and invariant checks.
#include "timetable.h"

#include <algorithm>
#include <map>
#include <numeric>
#include <sstream>

namespace nigiri_amd {

nigiri_tt_view timetable::view() const {
  nigiri_tt_view v{};
  v.n_locations = n_locations;
  v.n_routes = n_routes();
  v.n_transports = n_transports();
  v.n_bitfields = n_bitfields();
  v.n_route_stops = static_cast<std::uint32_t>(route_stops.size());
  v.n_event_times = static_cast<std::uint32_t>(event_times.size());
  v.n_location_routes = static_cast<std::uint32_t>(location_routes.size());
  v.n_footpaths = static_cast<std::uint32_t>(fp_out.size());
  v.n_lb_fwd = static_cast<std::uint32_t>(lb_fwd.size());
  v.n_lb_bwd = static_cast<std::uint32_t>(lb_bwd.size());
  v.date_range_from = date_range_from;
  v.date_range_to = date_range_to;
  v.internal_from = internal_from;
  v.n_days = n_days;
  v.route_stop_offsets = route_stop_offsets.data();
  v.route_stops = route_stops.data();
  v.route_transport_from = route_transport_from.data();
  v.route_stop_time_from = route_stop_time_from.data();
  v.event_times = event_times.data();
  v.transport_traffic_days = transport_traffic_days.data();
  v.route_traffic_days = route_traffic_days.data();
  v.bitfields = bitfields.data();
  v.location_route_offsets = location_route_offsets.data();
  v.location_routes = location_routes.data();
  v.transfer_time = transfer_time.data();
  v.fp_out_offsets = fp_out_offsets.data();
  v.fp_out = fp_out.data();
  v.fp_in_offsets = fp_in_offsets.data();
  v.fp_in = fp_in.data();
  v.lb_fwd_offsets = lb_fwd_offsets.data();
  v.lb_fwd = lb_fwd.data();
  v.lb_bwd_offsets = lb_bwd_offsets.data();
  v.lb_bwd = lb_bwd.data();
  v.coord_x = coord_x.empty() ? nullptr : coord_x.data();
  v.coord_y = coord_y.empty() ? nullptr : coord_y.data();
  return v;
}

namespace {

std::uint16_t event_at(timetable const& tt, std::uint32_t r, std::uint32_t t,
                       std::uint32_t stop_idx, bool arr) {
  auto const from = tt.route_transport_from[r];
  auto const n_t = tt.route_transport_from[r + 1U] - from;
  auto const idx = tt.route_stop_time_from[r] +
                   n_t * (stop_idx * 2U - (arr ? 1U : 0U)) + (t - from);
  return tt.event_times[idx];
}

// location_routes_: every route once per location, ascending route index.
void build_location_routes(timetable& tt) {
  auto const L = tt.n_locations;
  std::vector<std::vector<std::uint32_t>> lr(L);
  for (auto r = 0U; r != tt.n_routes(); ++r) {
    for (auto i = tt.route_stop_offsets[r]; i != tt.route_stop_offsets[r + 1U];
         ++i) {
      auto const l = tt.route_stops[i] & kStopLocMask;
      if (lr[l].empty() || lr[l].back() != r) {
        lr[l].push_back(r);
      }
    }
  }
  tt.location_route_offsets.assign(L + 1U, 0U);
  tt.location_routes.clear();
  for (auto l = 0U; l != L; ++l) {
    std::sort(lr[l].begin(), lr[l].end());
    lr[l].erase(std::unique(lr[l].begin(), lr[l].end()), lr[l].end());
    tt.location_routes.insert(tt.location_routes.end(), lr[l].begin(),
                              lr[l].end());
    tt.location_route_offsets[l + 1U] =
        static_cast<std::uint32_t>(tt.location_routes.size());
  }
}

// route_traffic_days_ = OR over the route's transports of the transport
// bitfield shifted by every day offset its events reach
// (init_finish.cc:168-198).  One new bitfield is appended per route.
void build_route_traffic_days(timetable& tt) {
  auto const n_base = tt.n_bitfields();
  // drop previously appended route bitfields when called twice
  tt.route_traffic_days.assign(tt.n_routes(), 0U);
  std::vector<std::uint64_t> out;
  out.reserve(static_cast<std::size_t>(tt.n_routes()) * kBitfieldWords);
  for (auto r = 0U; r != tt.n_routes(); ++r) {
    std::uint64_t combined[kBitfieldWords] = {};
    auto const n_stops = tt.route_stop_offsets[r + 1U] - tt.route_stop_offsets[r];
    for (auto t = tt.route_transport_from[r]; t != tt.route_transport_from[r + 1U];
         ++t) {
      int max_delta = 0;
      for (auto s = 0U; s != n_stops; ++s) {
        if (s != 0U) {
          max_delta = std::max(max_delta, delta_days(event_at(tt, r, t, s, true)));
        }
        if (s + 1U != n_stops) {
          max_delta =
              std::max(max_delta, delta_days(event_at(tt, r, t, s, false)));
        }
      }
      auto const* bf = &tt.bitfields[static_cast<std::size_t>(
                                         tt.transport_traffic_days[t]) *
                                     kBitfieldWords];
      for (int d = 0; d <= max_delta; ++d) {
        // combined |= bf << d   (bit i moves to i + d)
        for (int w = static_cast<int>(kBitfieldWords) - 1; w >= 0; --w) {
          std::uint64_t v = bf[w] << d;
          if (d != 0 && w > 0) {
            v |= bf[w - 1] >> (64 - d);
          }
          combined[w] |= v;
        }
      }
    }
    out.insert(out.end(), combined, combined + kBitfieldWords);
    tt.route_traffic_days[r] = n_base + r;
  }
  tt.bitfields.insert(tt.bitfields.end(), out.begin(), out.end());
}

void build_csr(std::uint32_t const L,
               std::vector<std::vector<std::uint32_t>>& adj,
               std::vector<std::uint32_t>& off, std::vector<std::uint32_t>& data) {
  off.assign(L + 1U, 0U);
  data.clear();
  for (auto l = 0U; l != L; ++l) {
    data.insert(data.end(), adj[l].begin(), adj[l].end());
    off[l + 1U] = static_cast<std::uint32_t>(data.size());
  }
}

auto const fp_less = [](std::uint32_t a, std::uint32_t b) {
  return std::make_pair(fp_duration(a), fp_target(a)) <
         std::make_pair(fp_duration(b), fp_target(b));
};

void build_footpaths_in(timetable& tt) {
  auto const L = tt.n_locations;
  std::vector<std::vector<std::uint32_t>> in(L), out(L);
  for (auto l = 0U; l != L; ++l) {
    for (auto i = tt.fp_out_offsets[l]; i != tt.fp_out_offsets[l + 1U]; ++i) {
      auto const fp = tt.fp_out[i];
      out[l].push_back(fp);
      in[fp_target(fp)].push_back(fp_pack(l, fp_duration(fp)));
    }
  }
  for (auto l = 0U; l != L; ++l) {
    std::sort(out[l].begin(), out[l].end(), fp_less);
    std::sort(in[l].begin(), in[l].end(), fp_less);
  }
  build_csr(L, out, tt.fp_out_offsets, tt.fp_out);
  build_csr(L, in, tt.fp_in_offsets, tt.fp_in);
}

// loader/build_lb_graph.h:11-99 for one direction (all locations are roots
// in these timetables: no parents, children or equivalences).
void build_lb_graph(timetable& tt, bool fwd) {
  auto const L = tt.n_locations;
  std::vector<std::vector<std::uint32_t>> adj(L);
  std::map<std::uint32_t, std::uint32_t> weights;
  auto const update = [&](std::uint32_t target, std::uint32_t d) {
    auto const it = weights.find(target);
    if (it != weights.end()) {
      it->second = std::min(it->second, d);
    } else {
      weights.emplace(target, d);
    }
  };
  for (auto l = 0U; l != L; ++l) {
    weights.clear();
    auto const& off = fwd ? tt.fp_in_offsets : tt.fp_out_offsets;
    auto const& fps = fwd ? tt.fp_in : tt.fp_out;
    for (auto i = off[l]; i != off[l + 1U]; ++i) {
      auto const target = fp_target(fps[i]);
      if (target != l) {
        update(target, fp_duration(fps[i]));
      }
    }
    for (auto j = tt.location_route_offsets[l];
         j != tt.location_route_offsets[l + 1U]; ++j) {
      auto const r = tt.location_routes[j];
      auto const s0 = tt.route_stop_offsets[r];
      auto const n = tt.route_stop_offsets[r + 1U] - s0;
      for (auto from = 0U; from + 1U < n; ++from) {
        auto const to = from + 1U;
        auto const from_l = tt.route_stops[s0 + from] & kStopLocMask;
        auto const to_l = tt.route_stops[s0 + to] & kStopLocMask;
        if ((fwd ? to_l : from_l) != l) {
          continue;
        }
        auto const target = fwd ? from_l : to_l;
        if (target == l) {
          continue;
        }
        auto min = std::uint32_t{32767U};
        for (auto t = tt.route_transport_from[r];
             t != tt.route_transport_from[r + 1U]; ++t) {
          auto const from_time = event_at(tt, r, t, from, false);
          auto const to_time = event_at(tt, r, t, to, true);
          // delta::operator- then as_duration() (types.h:601-610)
          auto const diff = static_cast<std::uint16_t>(
              (delta_days(to_time) - delta_days(from_time)) * 1440 +
              (delta_mam(to_time) - delta_mam(from_time)));
          auto const dur = static_cast<std::uint32_t>((diff / 1440U) * 1440U +
                                                      diff % 1440U);
          min = std::min(min, dur);
        }
        // footpath{target, min} saturates the duration at 511 minutes
        update(target, std::min(min, kFpMaxDuration));
      }
    }
    for (auto const& [target, d] : weights) {
      adj[l].push_back(fp_pack(target, d));
    }
  }
  build_csr(L, adj, fwd ? tt.lb_fwd_offsets : tt.lb_bwd_offsets,
            fwd ? tt.lb_fwd : tt.lb_bwd);
}

}  // namespace

namespace {

void build_importance(timetable& tt) {
  std::vector<unsigned> route_days(tt.n_routes(), 0U);
  for (auto r = 0U; r != tt.n_routes(); ++r) {
    for (auto t = tt.route_transport_from[r]; t != tt.route_transport_from[r + 1U]; ++t) {
      auto const* bf = &tt.bitfields[static_cast<std::size_t>(tt.transport_traffic_days[t]) * kBitfieldWords];
      for (auto w = 0U; w != kBitfieldWords; ++w) {
        route_days[r] += static_cast<unsigned>(__builtin_popcountll(bf[w]));
      }
    }
  }
  tt.location_importance.assign(tt.n_locations, 0U);
  tt.max_importance = 0U;
  for (auto l = 0U; l != tt.n_locations; ++l) {
    unsigned count = 0U;
    for (auto j = tt.location_route_offsets[l]; j != tt.location_route_offsets[l + 1U]; ++j) {
      count += route_days[tt.location_routes[j]];
    }
    auto& x = tt.location_importance[l];
    x += 1.0F * static_cast<float>(count);
    tt.max_importance = std::max(tt.max_importance, x);
  }
}

}  // namespace

void finalize(timetable& tt) {
  build_location_routes(tt);
  build_importance(tt);
  build_route_traffic_days(tt);
  build_footpaths_in(tt);
  build_lb_graph(tt, true);
  build_lb_graph(tt, false);
  tt.ch = build_ch(tt.n_locations, tt.lb_fwd_offsets, tt.lb_fwd);
}

std::string validate(timetable const& tt) {
  std::ostringstream err;
  auto const L = tt.n_locations;
  auto const R = tt.n_routes();
  if (tt.transfer_time.size() != L) err << "transfer_time size; ";
  if (tt.route_transport_from.size() != R + 1U) err << "route_transport_from; ";
  if (tt.route_stop_time_from.size() != R + 1U) err << "route_stop_time_from; ";
  if (tt.route_traffic_days.size() != R) err << "route_traffic_days; ";
  if (tt.transport_traffic_days.size() != tt.n_transports())
    err << "transport_traffic_days; ";
  if (tt.n_days + static_cast<std::uint32_t>(kTimetableOffsetDays) > 512U)
    err << "n_days > 512; ";
  for (auto r = 0U; r != R && err.tellp() == 0; ++r) {
    auto const n = tt.route_stop_offsets[r + 1U] - tt.route_stop_offsets[r];
    auto const n_t = tt.route_transport_from[r + 1U] - tt.route_transport_from[r];
    if (n < 2U || n_t == 0U) {
      err << "route " << r << " too small; ";
      break;
    }
    if (tt.route_stop_time_from[r + 1U] - tt.route_stop_time_from[r] !=
        2U * (n - 1U) * n_t) {
      err << "route " << r << " event block size; ";
      break;
    }
    for (auto i = 0U; i != n; ++i) {
      if ((tt.route_stops[tt.route_stop_offsets[r] + i] & kStopLocMask) >= L) {
        err << "route " << r << " stop out of range; ";
      }
    }
    // FIFO modulo 1440 (loader/get_index.h:8-32): transports sorted by first
    // departure mam, no overtaking (mod 1440) at any stop event.
    auto const base = tt.route_stop_time_from[r];
    for (auto j = 0U; j != 2U * (n - 1U); ++j) {
      for (auto t = 1U; t < n_t; ++t) {
        auto const a = delta_count(tt.event_times[base + j * n_t + t - 1U]);
        auto const b = delta_count(tt.event_times[base + j * n_t + t]);
        if (a % 1440 > b % 1440) {
          err << "route " << r << " not FIFO mod 1440; ";
          j = 2U * (n - 1U);
          break;
        }
      }
    }
  }
  for (auto l = 0U; l != L && err.tellp() == 0; ++l) {
    for (auto i = tt.fp_out_offsets[l]; i + 1U < tt.fp_out_offsets[l + 1U]; ++i) {
      if (fp_duration(tt.fp_out[i]) > fp_duration(tt.fp_out[i + 1U])) {
        err << "footpaths_out not sorted at " << l << "; ";
        break;
      }
    }
  }
  return err.str();
}

}  // namespace nigiri_amd
