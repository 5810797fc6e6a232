// Legs of reconstructed journeys (see journey_legs.h).
#include "journey_legs.h"

#include <algorithm>
#include <climits>

namespace nigiri_amd {

namespace {

using offsets_t = std::vector<std::pair<std::uint32_t, std::int16_t>>;

template <typename T>
T adjusted(host_tts const& s, T const d) {  // transfer_time_settings.h:21-31
  if (s.default_) return d;
  return static_cast<T>(static_cast<T>(s.additional_time_) +
                        std::max(static_cast<T>(s.min_transfer_time_),
                                 static_cast<T>(static_cast<float>(d) * s.factor_)));
}

std::int16_t clamp16(long long v) {
  return static_cast<std::int16_t>(std::clamp<long long>(v, SHRT_MIN, SHRT_MAX));
}

// One trip of the timetable on a traffic day (rt::frun over a whole route).
struct trip {
  timetable const& tt;
  std::uint32_t r, t;
  std::uint16_t day;

  std::uint32_t size() const { return tt.route_stop_offsets[r + 1U] - tt.route_stop_offsets[r]; }
  std::uint32_t word(std::uint32_t i) const { return tt.route_stops[tt.route_stop_offsets[r] + i]; }
  std::uint32_t loc(std::uint32_t i) const { return word(i) & kStopLocMask; }
  bool in(std::uint32_t i) const { return ((word(i) >> 28) & 1U) != 0U; }
  bool out(std::uint32_t i) const { return ((word(i) >> 29) & 1U) != 0U; }
  // timetable::event_time (timetable.h:185-190)
  std::int32_t time(std::uint32_t i, bool arr) const {
    auto const n_t = tt.route_transport_from[r + 1U] - tt.route_transport_from[r];
    auto const d = tt.event_times[tt.route_stop_time_from[r] + n_t * (2U * i - (arr ? 1U : 0U)) +
                                  (t - tt.route_transport_from[r])];
    return tt.internal_from + static_cast<std::int32_t>(day) * 1440 + delta_count(d);
  }
};

trip trip_of(timetable const& tt, nigiri_leg const& l) {
  return trip{tt, l.route, l.transport, l.day};
}

// journey::leg's constructor (journey.h:47-57): a and b in search direction
nigiri_leg make_leg(bool fwd, std::uint32_t a, std::uint32_t b, std::int32_t time_a,
                    std::int32_t time_b, std::int32_t kind) {
  nigiri_leg l{};
  l.kind = kind;
  l.from = fwd ? a : b;
  l.to = fwd ? b : a;
  l.dep_time = fwd ? time_a : time_b;
  l.arr_time = fwd ? time_b : time_a;
  return l;
}

std::uint32_t fp_list_begin(timetable const& tt, bool in, std::uint32_t l) {
  return (in ? tt.fp_in_offsets : tt.fp_out_offsets)[l];
}

// get_penalty (optimize_footpaths.cc:262-285)
double penalty(timetable const& tt, int duration, int buffer, std::uint32_t from, std::uint32_t to) {
  auto w = static_cast<double>(duration);
  if (from == to) {  // the same location is the only equivalent one here
    w -= (static_cast<double>(tt.location_importance[from]) / tt.max_importance) * 3U;
  }
  w -= 0.2 * static_cast<double>(std::clamp(buffer, 0, 15));
  return w;
}

// first footpath of `list` at location `a` that ends at `b`, or -1
long first_fp_to(timetable const& tt, bool in, std::uint32_t a, std::uint32_t b) {
  auto const& off = in ? tt.fp_in_offsets : tt.fp_out_offsets;
  auto const& fps = in ? tt.fp_in : tt.fp_out;
  for (auto e = off[a]; e != off[a + 1U]; ++e) {
    if (fp_target(fps[e]) == b) return static_cast<long>(e);
  }
  return -1;
}

}  // namespace

void optimize_footpaths(timetable const& tt, host_tts const& tts, offsets_t const& start,
                        offsets_t const& dest, std::vector<nigiri_leg>& legs) {
  auto const n_legs = legs.size();
  auto const is = [](nigiri_leg const& l, std::int32_t kind) { return l.kind == kind; };

  // optimize_initial_departure (:15-62): a shorter start offset to an
  // earlier-or-equal departure stop of the first trip
  if (n_legs > 1U && is(legs[0], NIGIRI_AMD_LEG_OFFSET) && is(legs[1], NIGIRI_AMD_LEG_RUN)) {
    auto& o_leg = legs[0];
    auto& t_leg = legs[1];
    auto const tr = trip_of(tt, t_leg);
    auto best = o_leg.duration;
    for (std::uint32_t i = 0U; i + 1U < t_leg.stop_to; ++i) {
      if (!tr.in(i)) continue;
      for (auto const& [loc, dur] : start) {
        if (best <= dur || loc != tr.loc(i)) continue;
        auto const dep = tr.time(i, false);
        if (o_leg.dep_time <= dep - dur) {
          o_leg.to = loc;
          o_leg.dep_time = dep - dur;
          o_leg.arr_time = dep;
          o_leg.target = loc;
          o_leg.duration = dur;
          t_leg.from = loc;
          t_leg.dep_time = dep;
          t_leg.stop_from = static_cast<std::uint16_t>(i);
          best = dur;
        }
      }
    }
  }

  // optimize_initial_start_footpath (:64-127)
  if (n_legs > 1U && is(legs[0], NIGIRI_AMD_LEG_FOOTPATH) && is(legs[1], NIGIRI_AMD_LEG_RUN)) {
    auto& fp_leg = legs[0];
    auto& t_leg = legs[1];
    auto const tr = trip_of(tt, t_leg);
    auto const from = fp_leg.from;
    auto best = adjusted<std::int16_t>(tts, fp_leg.duration);
    for (std::uint32_t i = 0U; i + 1U < t_leg.stop_to; ++i) {
      if (!tr.in(i)) continue;
      auto const e = first_fp_to(tt, false, from, tr.loc(i));
      if (e < 0) continue;
      auto const dur = adjusted<std::int16_t>(tts, static_cast<std::int16_t>(fp_duration(tt.fp_out[e])));
      if (dur >= best) continue;
      auto const dep = tr.time(i, false);
      if (fp_leg.dep_time <= dep - dur) {
        fp_leg.to = tr.loc(i);
        fp_leg.dep_time = dep - dur;
        fp_leg.arr_time = dep;
        fp_leg.target = tr.loc(i);
        fp_leg.duration = dur;
        t_leg.from = tr.loc(i);
        t_leg.dep_time = dep;
        t_leg.stop_from = static_cast<std::uint16_t>(i);
        best = dur;
      }
    }
  }

  // optimize_last_arrival (:129-184): a shorter egress offset from a later
  // stop of the last trip that still arrives in time
  if (n_legs > 1U && is(legs[n_legs - 1U], NIGIRI_AMD_LEG_OFFSET) &&
      is(legs[n_legs - 2U], NIGIRI_AMD_LEG_RUN)) {
    auto& o_leg = legs[n_legs - 1U];
    auto& t_leg = legs[n_legs - 2U];
    auto const tr = trip_of(tt, t_leg);
    auto best = o_leg.duration;
    for (std::uint32_t i = t_leg.stop_from + 1U; i < tr.size(); ++i) {
      if (!tr.out(i)) continue;
      for (auto const& [loc, dur] : dest) {
        if (best <= dur || loc != tr.loc(i)) continue;
        auto const arr = tr.time(i, true);
        if (arr + dur <= o_leg.arr_time) {
          o_leg.from = loc;
          o_leg.dep_time = arr;
          o_leg.arr_time = arr + dur;
          o_leg.target = loc;
          o_leg.duration = dur;
          t_leg.to = loc;
          t_leg.arr_time = arr;
          t_leg.stop_to = static_cast<std::uint16_t>(i + 1U);
          best = dur;
        }
      }
    }
  }

  // optimize_final_egress_footpath (:186-260)
  if (n_legs > 1U && is(legs[n_legs - 1U], NIGIRI_AMD_LEG_FOOTPATH) &&
      is(legs[n_legs - 2U], NIGIRI_AMD_LEG_RUN)) {
    auto& fp_leg = legs[n_legs - 1U];
    auto& t_leg = legs[n_legs - 2U];
    auto const tr = trip_of(tt, t_leg);
    auto const to = fp_leg.to;
    auto best = adjusted<std::int16_t>(tts, fp_leg.duration);
    for (std::uint32_t i = t_leg.stop_from + 1U; i < tr.size(); ++i) {
      if (!tr.out(i)) continue;
      auto const e = first_fp_to(tt, true, to, tr.loc(i));
      if (e < 0) continue;
      auto const dur = adjusted<std::int16_t>(tts, static_cast<std::int16_t>(fp_duration(tt.fp_in[e])));
      if (dur >= best) continue;
      auto const arr = tr.time(i, true);
      if (arr + dur <= fp_leg.arr_time) {
        fp_leg.from = tr.loc(i);
        fp_leg.dep_time = arr;
        fp_leg.arr_time = arr + dur;
        fp_leg.target = to;
        fp_leg.duration = dur;
        t_leg.to = tr.loc(i);
        t_leg.arr_time = arr;
        t_leg.stop_to = static_cast<std::uint16_t>(i + 1U);
        best = dur;
      }
    }
  }

  // optimize_transfers (:287-442): the transfer between two trips with the
  // lowest penalty over all stop pairs joined by a footpath.  The reference
  // tries the pairs (x, y) in order and keeps strict improvements, i.e. the
  // first pair of lowest penalty; here each stop x's footpaths are read once
  // and matched against the boarding stops of the second trip by location,
  // and ties are broken by (x, y).
  std::vector<std::pair<std::uint32_t, std::uint32_t>> board;  // (location, y)
  std::vector<std::uint32_t> seen;
  // locations of the current board list: stamp[l] == epoch
  thread_local std::vector<std::uint32_t> stamp;
  thread_local std::uint32_t epoch = 0U;
  if (stamp.size() != tt.n_locations) {
    stamp.assign(tt.n_locations, 0U);
    epoch = 0U;
  }
  for (std::size_t i = 0U; i + 2U < legs.size(); ++i) {
    auto& a_leg = legs[i];
    auto& fp_leg = legs[i + 1U];
    auto& b_leg = legs[i + 2U];
    if (!is(a_leg, NIGIRI_AMD_LEG_RUN) || !is(fp_leg, NIGIRI_AMD_LEG_FOOTPATH) ||
        !is(b_leg, NIGIRI_AMD_LEG_RUN)) {
      continue;
    }
    auto const ta = trip_of(tt, a_leg), tb = trip_of(tt, b_leg);
    auto const initial =
        penalty(tt, fp_leg.duration, b_leg.dep_time - fp_leg.arr_time, a_leg.from, b_leg.to);
    board.clear();
    if (++epoch == 0U) {
      std::fill(stamp.begin(), stamp.end(), 0U);
      epoch = 1U;
    }
    for (std::uint32_t y = 0U; y + 1U < b_leg.stop_to; ++y) {
      if (tb.in(y)) {
        board.emplace_back(tb.loc(y), y);
        stamp[tb.loc(y)] = epoch;
      }
    }
    std::sort(board.begin(), board.end());
    bool found = false;
    double best = 0.0;
    std::uint32_t bx = 0U, by = 0U;
    long be = -1;
    for (std::uint32_t x = a_leg.stop_from + 1U; x < ta.size(); ++x) {
      if (!ta.out(x)) continue;
      auto const la = ta.loc(x);
      std::int32_t arr = 0;
      bool have_arr = false;
      seen.clear();
      for (auto e = tt.fp_out_offsets[la]; e != tt.fp_out_offsets[la + 1U]; ++e) {
        auto const target = fp_target(tt.fp_out[e]);
        if (stamp[target] != epoch) continue;  // no boarding stop of the second trip
        if (std::find(seen.begin(), seen.end(), target) != seen.end()) continue;  // first footpath per target
        seen.push_back(target);
        auto it = std::lower_bound(board.begin(), board.end(), std::make_pair(target, 0U));
        for (; it != board.end() && it->first == target; ++it) {
          auto const y = it->second;
          auto const dur = adjusted<std::int16_t>(tts, static_cast<std::int16_t>(fp_duration(tt.fp_out[e])));
          if (!have_arr) {
            arr = ta.time(x, true);
            have_arr = true;
          }
          auto const dep = tb.time(y, false);
          if (arr + dur > dep) continue;
          auto const pen = penalty(tt, dur, dep - (arr + dur), la, target);
          if (!(pen < initial)) continue;
          if (!found || pen < best || (pen == best && std::make_pair(x, y) < std::make_pair(bx, by))) {
            found = true;
            best = pen;
            bx = x;
            by = y;
            be = static_cast<long>(e);
          }
        }
      }
    }
    if (!found) continue;
    auto const raw = static_cast<std::int16_t>(fp_duration(tt.fp_out[be]));
    auto const dur = adjusted<std::int16_t>(tts, raw);
    auto const arr = ta.time(bx, true);
    a_leg.to = ta.loc(bx);
    a_leg.arr_time = arr;
    a_leg.stop_to = static_cast<std::uint16_t>(bx + 1U);
    b_leg.from = tb.loc(by);
    b_leg.dep_time = tb.time(by, false);
    b_leg.stop_from = static_cast<std::uint16_t>(by);
    fp_leg.from = ta.loc(bx);
    fp_leg.to = tb.loc(by);
    fp_leg.dep_time = arr;
    fp_leg.arr_time = arr + dur;
    fp_leg.target = tb.loc(by);
    fp_leg.duration = raw;
  }
}

void assemble_legs(timetable const& tt, prepared_query const& p, rc_out const& o,
                   host_journey& j) {
  auto const fwd = p.fwd_;
  auto const dir = [&](int a) { return fwd ? a : -a; };
  auto const better_eq = [&](int a, int b) { return fwd ? a <= b : a >= b; };
  auto const base = p.base_minutes(tt);
  auto const top = static_cast<std::uint32_t>(j.transfers_) + 1U;
  auto& legs = j.legs_;
  legs.clear();
  if (o.status != 0 || o.n_steps != top) {
    j.error_ = true;
    return;
  }

  // the walk and the trip of every round, destination first
  // (reconstruct.cc:681-811)
  auto l = j.dest_;
  for (std::uint32_t i = 0U; i != top; ++i) {
    auto const k = top - i;
    auto const& s = o.steps[i];
    std::uint32_t target = l;
    std::int16_t used = 0, stored = 0;
    std::int32_t kind = NIGIRI_AMD_LEG_FOOTPATH;
    if (p.intermodal_dest_ && k == top) {
      auto const& [loc, dur] = p.dest_offsets_.at(static_cast<std::size_t>(s.walk));
      target = loc;
      used = stored = dur;
      kind = NIGIRI_AMD_LEG_OFFSET;
    } else if (s.walk < 0) {
      used = stored = static_cast<std::int16_t>(
          k == top ? 0 : adjusted<std::uint8_t>(p.tts_, tt.transfer_time[l]));
    } else {
      auto const e = fp_list_begin(tt, fwd, l) + static_cast<std::uint32_t>(s.walk);
      auto const fp = (fwd ? tt.fp_in : tt.fp_out).at(e);
      target = fp_target(fp);
      stored = static_cast<std::int16_t>(fp_duration(fp));
      used = adjusted<std::int16_t>(p.tts_, stored);
    }
    auto const fp_start = static_cast<std::int16_t>(s.curr_time - dir(used));
    trip const tr{tt, s.route, s.transport, s.day};
    auto const ev = clamp16(static_cast<long long>(tr.time(s.stop_idx, !fwd)) - base);
    auto run = make_leg(fwd, tr.loc(s.stop_idx), tr.loc(s.finish_idx), base + ev,
                        tr.time(s.finish_idx, fwd), NIGIRI_AMD_LEG_RUN);
    run.transport = s.transport;
    run.route = s.route;
    run.day = s.day;
    run.stop_from = std::min(s.stop_idx, s.finish_idx);
    run.stop_to = static_cast<std::uint16_t>(std::max(s.stop_idx, s.finish_idx) + 1U);
    auto walk = make_leg(fwd, target, l, base + fp_start,
                         base + static_cast<std::int16_t>(fp_start + dir(used)), kind);
    walk.target = target;
    walk.duration = stored;
    l = tr.loc(s.stop_idx);
    if (i != 0U || walk.from != walk.to || walk.dep_time != walk.arr_time) legs.push_back(walk);
    legs.push_back(run);
  }
  if (l != o.start_loc) {
    j.error_ = true;
    return;
  }

  // find_start_footpath (reconstruct.cc:30-160)
  {
    auto const& back = legs.back();
    auto const at = fwd ? back.from : back.to;
    auto const at_time = fwd ? back.dep_time : back.arr_time;
    auto const is_start = [&](std::uint32_t x) {
      return std::any_of(p.start_offsets_.begin(), p.start_offsets_.end(),
                         [&](auto const& so) { return so.first == x; });
    };
    if (p.start_intermodal_ || !is_start(at) || !better_eq(j.start_time_, at_time)) {
      bool found = false;
      if (p.start_intermodal_) {
        for (auto const& [loc, dur] : p.start_offsets_) {
          if (loc == at && better_eq(j.start_time_, at_time - dir(dur))) {
            auto lg = make_leg(fwd, 0U, at, at_time - dir(dur), at_time, NIGIRI_AMD_LEG_OFFSET);
            lg.target = loc;
            lg.duration = dur;
            legs.push_back(lg);
            found = true;
            break;
          }
        }
      } else {
        auto const invalid = fwd ? SHRT_MAX : SHRT_MIN;
        auto const j_start = clamp16(static_cast<long long>(j.start_time_) - base);
        auto const& off = fwd ? tt.fp_in_offsets : tt.fp_out_offsets;
        auto const& fps = fwd ? tt.fp_in : tt.fp_out;
        for (auto e = off[at]; e != off[at + 1U]; ++e) {
          auto const src = fp_target(fps[e]);
          auto const dur = adjusted<std::int16_t>(p.tts_, static_cast<std::int16_t>(fp_duration(fps[e])));
          auto const a = j_start + dir(dur);
          if (is_start(src) && o.rt0 != invalid && (p.ontrip_ ? better_eq(a, o.rt0) : a == o.rt0)) {
            auto lg = make_leg(fwd, src, at, j.start_time_, base + o.rt0, NIGIRI_AMD_LEG_FOOTPATH);
            lg.target = src;
            lg.duration = dur;
            legs.push_back(lg);
            found = true;
            break;
          }
        }
      }
      if (!found) {
        legs.clear();
        j.error_ = true;
        return;
      }
    }
  }

  // travel order (reconstruct.cc:819-859)
  if (fwd) {
    std::reverse(legs.begin(), legs.end());
  } else {
    for (std::size_t x = 1U; x < legs.size(); ++x) {
      if (legs[x].kind != NIGIRI_AMD_LEG_FOOTPATH) continue;
      auto const diff = legs[x].dep_time - legs[x - 1U].arr_time;
      legs[x].dep_time -= diff;
      legs[x].arr_time -= diff;
    }
  }
  // the journey-direction query: a backward search flips start and destination
  optimize_footpaths(tt, p.tts_, fwd ? p.start_offsets_ : p.dest_offsets_,
                     fwd ? p.dest_offsets_ : p.start_offsets_, legs);
  j.reconstructed_ = true;
}

}  // namespace nigiri_amd
