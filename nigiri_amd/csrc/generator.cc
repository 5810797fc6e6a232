// Seeded synthetic timetables and query workloads.
//
// Workloads (sizes at scale 1.0):
//   1: 2,000 stops on a 45x45 grid, 200 lines of 10-40 stops, 20 trips per
//      line at 10-30 min headways from 05:00, 7 service days, all active,
//      grid-neighbour footpaths of 2-10 min.
//   2: 50k stops (random geometric), 5k lines of 10-60 stops, 100 trips per
//      line at 5-60 min headways, 365 days (60 % Mon-Fri, 25 % daily, 15 %
//      random 50 %), footpaths <= 400 m, 1-15 min, mean out-degree 4.
//   3/4: 200k stops, 20k lines, 800k trips, out-degree 6.
//   5: 500k stops, 60k lines, 2.5M trips, out-degree 6.
// Hop times 2-8 min, dwell 0-1 min, transfer time 2 min everywhere.
//
// Lines are expanded into nigiri routes the way the GTFS loader does: trips
// of one stop sequence go into the first route where they keep FIFO order
// modulo 1440 (loader/get_index.h:8-32, used at gtfs/load_timetable.cc:
// 291-301); first departures >= 24:00 are normalised to a time of day with
// the traffic-day bitfield shifted by whole days.
#include <algorithm>
#include <cmath>
#include <map>
#include <numeric>
#include <stdexcept>
#include <unordered_map>

#include "generator.h"
#include "rng.h"
#include "timetable.h"

namespace nigiri_amd {

namespace {

struct gen_cfg {
  std::uint32_t n_stops;
  bool grid;
  std::uint32_t grid_w;
  std::uint32_t n_lines, len_min, len_max;
  std::uint64_t n_trips;
  std::uint32_t hw_min, hw_max;
  std::int32_t first_dep_min, first_dep_max;
  std::uint32_t ext_days;
  bool mixed_bitfields;
  double fp_radius, fp_degree;
  std::uint32_t fp_min, fp_max;  // grid footpath durations
};

gen_cfg make_cfg(int cfg, double scale) {
  auto const s = [&](double x) {
    return static_cast<std::uint32_t>(std::max(1.0, std::round(x * scale)));
  };
  switch (cfg) {
    case 1:
      return {s(2000), true, 45, s(200), 10, 40, s(200) * 20ULL, 10, 30,
              300, 300, 7, false, 0.0, 0.0, 2, 10};
    case 2:
      return {s(50000), false, 0, s(5000), 10, 60, s(5000) * 100ULL, 5, 60,
              0, 1439, 365, true, 400.0, 4.0, 0, 0};
    case 3:
    case 4:
      return {s(200000), false, 0, s(20000), 10, 60, s(20000) * 40ULL, 5, 60,
              0, 1439, 365, true, 400.0, 6.0, 0, 0};
    case 5: {
      auto const lines = s(60000);
      return {s(500000), false, 0, lines, 10, 60,
              static_cast<std::uint64_t>(std::llround(lines * (2.5e6 / 6e4))),
              5, 60, 0, 1439, 365, true, 400.0, 6.0, 0, 0};
    }
    default: throw std::invalid_argument("workload must be 1..5");
  }
}

// 2026-01-05 (a Monday), in days since the unix epoch.
constexpr std::int32_t kFirstDay = 20458;

struct spatial_grid {
  spatial_grid(std::vector<float> const& x, std::vector<float> const& y,
               double cell, double side)
      : x_{x}, y_{y}, cell_{cell} {
    n_ = std::max<std::int64_t>(1, static_cast<std::int64_t>(side / cell) + 1);
    cells_.resize(static_cast<std::size_t>(n_ * n_));
    for (auto i = 0U; i != x.size(); ++i) {
      cells_[idx(x[i], y[i])].push_back(i);
    }
  }
  std::size_t idx(double px, double py) const {
    auto cx = std::clamp<std::int64_t>(static_cast<std::int64_t>(px / cell_), 0,
                                       n_ - 1);
    auto cy = std::clamp<std::int64_t>(static_cast<std::int64_t>(py / cell_), 0,
                                       n_ - 1);
    return static_cast<std::size_t>(cy * n_ + cx);
  }
  template <typename Fn>
  void within(double px, double py, double radius, Fn&& fn) const {
    auto const r = static_cast<std::int64_t>(std::ceil(radius / cell_));
    auto const cx = static_cast<std::int64_t>(px / cell_);
    auto const cy = static_cast<std::int64_t>(py / cell_);
    for (auto yy = std::max<std::int64_t>(0, cy - r);
         yy <= std::min(n_ - 1, cy + r); ++yy) {
      for (auto xx = std::max<std::int64_t>(0, cx - r);
           xx <= std::min(n_ - 1, cx + r); ++xx) {
        for (auto const i : cells_[static_cast<std::size_t>(yy * n_ + xx)]) {
          auto const dx = x_[i] - px, dy = y_[i] - py;
          auto const d = std::sqrt(dx * dx + dy * dy);
          if (d <= radius) {
            fn(i, d);
          }
        }
      }
    }
  }
  std::vector<float> const& x_;
  std::vector<float> const& y_;
  double cell_;
  std::int64_t n_;
  std::vector<std::vector<std::uint32_t>> cells_;
};

struct line {
  std::vector<std::uint32_t> stops;   // stop ids (0-based, without specials)
  std::vector<std::uint32_t> offset;  // per stop event (dep0, arr1, dep1, ...)
};

}  // namespace

double side_length(gen_cfg const& c) {
  if (c.grid) {
    return 300.0 * c.grid_w;
  }
  return std::sqrt(c.n_stops * M_PI * c.fp_radius * c.fp_radius / c.fp_degree);
}

timetable generate(int cfg_id, std::uint64_t seed, double scale) {
  auto const c = make_cfg(cfg_id, scale);
  rng r{seed};
  timetable tt;
  auto const N = c.n_stops;
  auto const L = N + kNSpecialStations;
  tt.n_locations = L;
  tt.date_range_from = kFirstDay * 1440;
  tt.date_range_to = (kFirstDay + static_cast<std::int32_t>(c.ext_days)) * 1440;
  tt.internal_from = tt.date_range_from - kTimetableOffsetDays * 1440;
  tt.n_days = c.ext_days + kTimetableOffsetDays + 1U;  // [from-5d, to+1d)
  tt.transfer_time.assign(L, 2U);
  for (auto i = 0U; i != kNSpecialStations; ++i) {
    tt.transfer_time[i] = 0U;
  }

  // ---- stops and footpaths ----------------------------------------------
  std::vector<float> x(N), y(N);
  std::vector<std::vector<std::uint32_t>> walk_nb(N);
  std::vector<std::vector<std::uint32_t>> fp_out(L);
  auto const side = side_length(c);
  if (c.grid) {
    for (auto i = 0U; i != N; ++i) {
      x[i] = static_cast<float>((i % c.grid_w) * 300.0);
      y[i] = static_cast<float>((i / c.grid_w) * 300.0);
    }
    auto const link = [&](std::uint32_t a, std::uint32_t b) {
      walk_nb[a].push_back(b);
      walk_nb[b].push_back(a);
      auto const d = static_cast<std::uint32_t>(r.uniform(c.fp_min, c.fp_max));
      fp_out[a + kNSpecialStations].push_back(fp_pack(b + kNSpecialStations, d));
      fp_out[b + kNSpecialStations].push_back(fp_pack(a + kNSpecialStations, d));
    };
    for (auto i = 0U; i != N; ++i) {
      if ((i % c.grid_w) + 1U < c.grid_w && i + 1U < N) link(i, i + 1U);
      if (i + c.grid_w < N) link(i, i + c.grid_w);
    }
  } else {
    for (auto i = 0U; i != N; ++i) {
      x[i] = static_cast<float>(r.unit() * side);
      y[i] = static_cast<float>(r.unit() * side);
    }
    spatial_grid g{x, y, c.fp_radius, side};
    auto const walk_radius = 1.5 * c.fp_radius;
    for (auto i = 0U; i != N; ++i) {
      g.within(x[i], y[i], walk_radius, [&](std::uint32_t j, double d) {
        if (j == i) return;
        walk_nb[i].push_back(j);
        if (d <= c.fp_radius) {
          auto const dur = static_cast<std::uint32_t>(
              std::min(15.0, 1.0 + std::floor(d * 14.0 / c.fp_radius)));
          fp_out[i + kNSpecialStations].push_back(
              fp_pack(j + kNSpecialStations, dur));
        }
      });
      std::sort(walk_nb[i].begin(), walk_nb[i].end());
    }
  }
  tt.coord_x.assign(L, 0.0F);
  tt.coord_y.assign(L, 0.0F);
  for (auto i = 0U; i != N; ++i) {
    tt.coord_x[i + kNSpecialStations] = x[i];
    tt.coord_y[i + kNSpecialStations] = y[i];
  }
  tt.fp_out_offsets.assign(L + 1U, 0U);
  for (auto l = 0U; l != L; ++l) {
    tt.fp_out.insert(tt.fp_out.end(), fp_out[l].begin(), fp_out[l].end());
    tt.fp_out_offsets[l + 1U] = static_cast<std::uint32_t>(tt.fp_out.size());
  }

  // ---- lines: self-avoiding walks with a persistent heading --------------
  std::vector<line> lines;
  lines.reserve(c.n_lines);
  std::vector<std::uint32_t> visited(N, 0U);
  std::uint32_t stamp = 0U;
  for (auto li = 0U; li != c.n_lines; ++li) {
    auto const want = static_cast<std::uint32_t>(r.uniform(c.len_min, c.len_max));
    std::vector<std::uint32_t> best_walk;
    for (auto attempt = 0; attempt != 20 && best_walk.size() < want; ++attempt) {
      ++stamp;
      std::vector<std::uint32_t> w{static_cast<std::uint32_t>(r.uniform(0, N - 1))};
      visited[w[0]] = stamp;
      auto heading = r.unit() * 2.0 * M_PI;
      while (w.size() < want) {
        auto const cur = w.back();
        auto best = std::uint32_t{~0U};
        auto best_score = -1e30;
        for (auto const nb : walk_nb[cur]) {
          if (visited[nb] == stamp) continue;
          auto const dx = x[nb] - x[cur], dy = y[nb] - y[cur];
          auto const a = std::atan2(dy, dx);
          auto const score = std::cos(a - heading) + 0.8 * r.unit();
          if (score > best_score) {
            best_score = score;
            best = nb;
          }
        }
        if (best == ~0U) break;
        heading = std::atan2(y[best] - y[cur], x[best] - x[cur]);
        visited[best] = stamp;
        w.push_back(best);
      }
      if (w.size() > best_walk.size()) best_walk = std::move(w);
    }
    if (best_walk.size() < 2U) continue;
    line ln;
    ln.stops = std::move(best_walk);
    auto const n = static_cast<std::uint32_t>(ln.stops.size());
    std::uint32_t t = 0U;
    ln.offset.push_back(0U);  // dep at stop 0
    for (auto i = 1U; i != n; ++i) {
      t += static_cast<std::uint32_t>(r.uniform(2, 8));
      ln.offset.push_back(t);  // arr at stop i
      if (i + 1U != n) {
        t += static_cast<std::uint32_t>(r.uniform(0, 1));
        ln.offset.push_back(t);  // dep at stop i
      }
    }
    lines.push_back(std::move(ln));
  }

  // ---- trips, bitfields, FIFO route grouping ------------------------------
  std::map<std::array<std::uint64_t, kBitfieldWords>, std::uint32_t> bf_idx;
  auto const get_bf = [&](std::array<std::uint64_t, kBitfieldWords> const& b) {
    auto const it = bf_idx.find(b);
    if (it != bf_idx.end()) return it->second;
    auto const idx = static_cast<std::uint32_t>(bf_idx.size());
    bf_idx.emplace(b, idx);
    tt.bitfields.insert(tt.bitfields.end(), b.begin(), b.end());
    return idx;
  };

  auto const n_lines = static_cast<std::uint32_t>(lines.size());
  std::uint64_t trips_left = c.n_trips;
  for (auto li = 0U; li != n_lines; ++li) {
    auto const& ln = lines[li];
    auto const n_trips = static_cast<std::uint32_t>(
        trips_left / (n_lines - li) +
        ((trips_left % (n_lines - li)) != 0U ? 1U : 0U));
    trips_left -= n_trips;

    // service days of the line (external day e is internal day e + 5)
    std::vector<bool> service(c.ext_days, true);
    if (c.mixed_bitfields) {
      auto const cls = r.unit();
      for (auto e = 0U; e != c.ext_days; ++e) {
        if (cls < 0.60) {
          service[e] = (e % 7U) < 5U;  // day 0 is a Monday
        } else if (cls < 0.85) {
          service[e] = true;
        } else {
          service[e] = r.coin(0.5);
        }
      }
    }

    struct trip {
      std::uint32_t mam0;
      std::uint32_t bf;
    };
    std::vector<trip> trips;
    auto dep = static_cast<std::uint32_t>(r.uniform(c.first_dep_min, c.first_dep_max));
    for (auto j = 0U; j != n_trips; ++j) {
      if (j != 0U) dep += static_cast<std::uint32_t>(r.uniform(c.hw_min, c.hw_max));
      auto const shift = dep / 1440U;
      std::array<std::uint64_t, kBitfieldWords> b{};
      for (auto e = 0U; e != c.ext_days; ++e) {
        if (!service[e]) continue;
        auto const bit = e + kTimetableOffsetDays + shift;
        if (bit < 512U) b[bit / 64U] |= 1ULL << (bit % 64U);
      }
      trips.push_back({dep % 1440U, get_bf(b)});
    }

    // get_index (loader/get_index.h:8-32) over the line's route groups
    std::vector<std::vector<std::uint32_t>> groups;
    auto const time_at = [&](std::uint32_t tr, std::size_t ev) {
      return trips[tr].mam0 + ln.offset[ev];
    };
    for (auto tr = 0U; tr != trips.size(); ++tr) {
      bool placed = false;
      for (auto& grp : groups) {
        auto const it = std::lower_bound(
            grp.begin(), grp.end(), tr, [&](std::uint32_t a, std::uint32_t b) {
              return time_at(a, 0) % 1440U < time_at(b, 0) % 1440U;
            });
        auto const index = static_cast<std::size_t>(it - grp.begin());
        bool ok = true;
        for (auto ev = 0U; ev != ln.offset.size() && ok; ++ev) {
          auto const v = time_at(tr, ev) % 1440U;
          if (index > 0 && v < time_at(grp[index - 1], ev) % 1440U) ok = false;
          if (index < grp.size() && v > time_at(grp[index], ev) % 1440U) ok = false;
        }
        if (ok) {
          grp.insert(grp.begin() + static_cast<std::ptrdiff_t>(index), tr);
          placed = true;
          break;
        }
      }
      if (!placed) groups.push_back({tr});
    }

    for (auto const& grp : groups) {
      for (auto const s : ln.stops) {
        tt.route_stops.push_back(
            stop_pack(s + kNSpecialStations, true, true, true, true));
      }
      tt.route_stop_offsets.push_back(static_cast<std::uint32_t>(tt.route_stops.size()));
      for (auto const tr : grp) {
        tt.transport_traffic_days.push_back(trips[tr].bf);
      }
      tt.route_transport_from.push_back(
          static_cast<std::uint32_t>(tt.transport_traffic_days.size()));
      for (auto ev = 0U; ev != ln.offset.size(); ++ev) {
        for (auto const tr : grp) {
          tt.event_times.push_back(delta_pack(time_at(tr, ev)));
        }
      }
      tt.route_stop_time_from.push_back(static_cast<std::uint32_t>(tt.event_times.size()));
    }
  }

  finalize(tt);
  return tt;
}

// ---- query workloads -------------------------------------------------------

namespace {

nigiri_query default_query() {
  nigiri_query q{};
  q.dir = NIGIRI_AMD_DIR_FWD;
  q.start_match_mode = NIGIRI_AMD_MATCH_EXACT;
  q.dest_match_mode = NIGIRI_AMD_MATCH_EXACT;
  q.use_start_footpaths = 1;
  q.max_start_offset = static_cast<std::int16_t>(kMaxTravelTime);
  q.max_travel_time = static_cast<std::int16_t>(kMaxTravelTime);
  q.max_transfers = kMaxTransfers;
  q.tts_default = 1;
  q.tts_factor = 1.0F;
  q.fastest_direct_factor = 1.0;
  return q;
}

void draw_offsets(timetable const& tt, spatial_grid const& g, rng& r,
                  double side, std::vector<std::uint32_t>& loc,
                  std::vector<std::int16_t>& dur) {
  auto const n = static_cast<std::uint32_t>(r.uniform(1, 50));
  auto const px = r.unit() * side, py = r.unit() * side;
  std::vector<std::pair<double, std::uint32_t>> cand;
  for (auto radius = 1000.0; cand.empty() && radius < 4.0 * side + 2000.0;
       radius *= 2.0) {
    g.within(px, py, radius, [&](std::uint32_t l, double d) {
      if (l >= kNSpecialStations) cand.emplace_back(d, l);
    });
    if (!cand.empty() && radius > 1000.0) {
      // no stop within 1 km: use the nearest stop only
      auto const best = *std::min_element(cand.begin(), cand.end());
      cand = {best};
    }
  }
  std::sort(cand.begin(), cand.end(),
            [](auto const& a, auto const& b) { return a.second < b.second; });
  // partial Fisher-Yates over the candidates
  auto const take = std::min<std::size_t>(n, cand.size());
  for (auto i = 0U; i != take; ++i) {
    auto const j = static_cast<std::size_t>(
        r.uniform(i, static_cast<std::int64_t>(cand.size()) - 1));
    std::swap(cand[i], cand[j]);
    loc.push_back(cand[i].second);
    dur.push_back(static_cast<std::int16_t>(r.uniform(0, 30)));
  }
}

}  // namespace

std::vector<gen_query> generate_queries(timetable const& tt, int cfg_id,
                                        std::uint64_t seed, std::uint32_t n) {
  rng r{seed};
  std::vector<gen_query> out;
  out.reserve(n);
  auto const N = tt.n_locations - kNSpecialStations;
  auto const ext_days = static_cast<std::uint32_t>(
      (tt.date_range_to - tt.date_range_from) / 1440);
  double side = 0.0;
  for (auto l = kNSpecialStations; l != tt.n_locations; ++l) {
    side = std::max(side, static_cast<double>(std::max(tt.coord_x[l], tt.coord_y[l])));
  }
  spatial_grid const grid{tt.coord_x, tt.coord_y, 1000.0, side + 1.0};
  for (auto i = 0U; i != n; ++i) {
    gen_query g;
    g.q = default_query();
    // kind: 0 point, 1 range, 2 intermodal
    int kind = 0;
    auto const u = r.unit();
    switch (cfg_id) {
      case 1: g.q.max_transfers = 7; break;
      case 2: break;
      case 3: g.q.max_transfers = 8; kind = u < 0.70 ? 0 : 1; break;
      case 4: g.q.max_transfers = 8; kind = 2; break;
      case 5:
        g.q.max_transfers = 8;
        kind = u < 0.60 ? 0 : (u < 0.85 ? 1 : 2);
        break;
      default: throw std::invalid_argument("workload must be 1..5");
    }
    auto const day = static_cast<std::int32_t>(r.uniform(0, ext_days - 1));
    auto const t = tt.date_range_from + day * 1440 +
                   static_cast<std::int32_t>(r.uniform(360, 1200));
    if (kind == 2) {
      g.q.start_match_mode = NIGIRI_AMD_MATCH_INTERMODAL;
      g.q.dest_match_mode = NIGIRI_AMD_MATCH_INTERMODAL;
      g.q.use_start_footpaths = 0;
      draw_offsets(tt, grid, r, side, g.start_loc, g.start_dur);
      draw_offsets(tt, grid, r, side, g.dest_loc, g.dest_dur);
    } else {
      auto const src = static_cast<std::uint32_t>(r.uniform(0, N - 1));
      auto dst = static_cast<std::uint32_t>(r.uniform(0, N - 2));
      if (dst >= src) ++dst;
      g.start_loc = {src + kNSpecialStations};
      g.dest_loc = {dst + kNSpecialStations};
      g.start_dur = {0};
      g.dest_dur = {0};
    }
    if (kind == 1) {
      g.q.is_interval = 1;
      g.q.interval_from = t;
      g.q.interval_to = t + static_cast<std::int32_t>(r.uniform(60, 240));
    } else {
      g.q.start_time = t;
    }
    out.push_back(std::move(g));
  }
  return out;
}

}  // namespace nigiri_amd
