// GTFS loader (see gtfs.h).  Follows, in order:
//   read_agencies (gtfs/agency.cc): agency_id -> time zone;
//   read_stops / read_transfers (gtfs/stop.cc:129-300): one location per
//     stop, transfer_time_ from same-stop transfers (default 2 min,
//     loader_interface.h:19), footpaths from transfers.txt made symmetric
//     (stop.cc:352-363);
//   read_routes (gtfs/route.cc), read_calendar / read_calendar_date /
//     merge_traffic_days (gtfs/services.cc:12-73): service bitfields over the
//     internal interval [date_range.from - 5 d, date_range.to + 1 d);
//   read_trips / read_stop_times (gtfs/trip.cc, stop_time.cc:155-200):
//     stop sequences with pickup / drop-off flags, sorted by stop_sequence;
//   precompute_noon_offsets (noon_offsets.cc:14-40): the first agency's UTC
//     offset at local noon of every day;
//   expand_local_to_utc (local_to_utc.h:150-250): one UTC trip per
//     (first-departure day offset, UTC offset) with its UTC traffic days;
//   route grouping by (route type, stop sequence) with FIFO placement
//     (load_timetable.cc:277-310, get_index.h:8-32) and the event layout of
//     load_timetable.cc:459-474;
//   finalize (init_finish.cc:200-234) via nigiri_amd::finalize.
// Not restated (refused with an error where a feed uses them): frequencies,
// stay-seated transfers and block_id trips, flex, stop times that need
// interpolation.  Not generated: footpaths between nearby or same-name
// stops (stop.cc:365-398) and their transitive closure (build_footpaths.cc).
#include "gtfs.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <stdexcept>
#include <cstdio>
#include <tuple>
#include <unordered_map>

#include "capi_util.h"
#include "tz.h"

namespace nigiri_amd {

namespace {

constexpr std::int32_t kInterpolate = -1;

// ---- CSV (RFC 4180 quoting) ------------------------------------------------
struct csv {
  std::vector<std::string> header;
  std::vector<std::vector<std::string>> rows;
  int col(std::string const& name) const {
    auto const it = std::find(header.begin(), header.end(), name);
    return it == header.end() ? -1 : static_cast<int>(it - header.begin());
  }
  std::string const& get(std::vector<std::string> const& row, int c) const {
    static std::string const empty;
    return c < 0 || static_cast<std::size_t>(c) >= row.size() ? empty : row[static_cast<std::size_t>(c)];
  }
};

std::string trim(std::string s) {
  auto const b = s.find_first_not_of(" \t\r\n");
  if (b == std::string::npos) return {};
  auto const e = s.find_last_not_of(" \t\r\n");
  return s.substr(b, e - b + 1U);
}

csv parse_csv(std::string const& text) {
  csv out;
  std::size_t i = 0U;
  if (text.compare(0, 3, "\xEF\xBB\xBF") == 0) i = 3U;  // UTF-8 byte order mark
  bool first = true;
  while (i < text.size()) {
    std::vector<std::string> fields;
    std::string f;
    bool quoted = false, any = false;
    for (; i < text.size(); ++i) {
      auto const ch = text[i];
      if (quoted) {
        if (ch == '"') {
          if (i + 1U < text.size() && text[i + 1U] == '"') {
            f += '"';
            ++i;
          } else {
            quoted = false;
          }
        } else {
          f += ch;
        }
      } else if (ch == '"') {
        quoted = true;
        any = true;
      } else if (ch == ',') {
        fields.push_back(trim(f));
        f.clear();
        any = true;
      } else if (ch == '\n') {
        ++i;
        break;
      } else if (ch != '\r') {
        f += ch;
        any = true;
      }
    }
    if (!any && trim(f).empty()) continue;  // blank line
    fields.push_back(trim(f));
    if (first) {
      out.header = std::move(fields);
      first = false;
    } else {
      out.rows.push_back(std::move(fields));
    }
  }
  return out;
}

// HH:MM:SS (hours may exceed 24) -> minutes; empty -> kInterpolate
std::int32_t hhmm_to_min(std::string const& s) {
  if (s.empty()) return kInterpolate;
  int h = 0, m = 0, sec = 0;
  if (std::sscanf(s.c_str(), "%d:%d:%d", &h, &m, &sec) < 2) {
    throw std::invalid_argument("stop_times.txt: bad time \"" + s + "\"");
  }
  return h * 60 + m;
}

// YYYYMMDD -> days since 1970-01-01
std::int32_t parse_date(std::string const& s) {
  if (s.size() != 8U) throw std::invalid_argument("bad GTFS date \"" + s + "\"");
  auto const y = std::stoi(s.substr(0, 4));
  auto const m = static_cast<unsigned>(std::stoi(s.substr(4, 2)));
  auto const d = static_cast<unsigned>(std::stoi(s.substr(6, 2)));
  return days_from_civil(y, m, d);
}

int to_int(std::string const& s, int def) { return s.empty() ? def : std::stoi(s); }

using bitfield = std::array<std::uint64_t, kBitfieldWords>;

void set_bit(bitfield& b, std::int64_t i, bool v = true) {
  if (i < 0 || i >= 512) return;
  auto const w = static_cast<std::size_t>(i / 64);
  auto const s = static_cast<unsigned>(i % 64);
  if (v) {
    b[w] |= 1ULL << s;
  } else {
    b[w] &= ~(1ULL << s);
  }
}
bool test_bit(bitfield const& b, std::int64_t i) {
  return i >= 0 && i < 512 && ((b[static_cast<std::size_t>(i / 64)] >> (i % 64)) & 1ULL) != 0ULL;
}

struct trip {
  std::string id;
  int route_type = 3;
  std::string service;
  std::vector<std::pair<int, std::uint32_t>> seq;  // (stop_sequence, stop word)
  std::vector<std::pair<std::int32_t, std::int32_t>> times;  // (arr, dep) local minutes
};

struct utc_trip {
  std::uint32_t trip;
  std::vector<std::int32_t> utc_times;  // dep(0), arr(1), dep(1), ..., arr(n-1)
  bitfield days{};
};

}  // namespace

timetable load_gtfs(std::vector<std::pair<std::string, std::string>> const& files,
                    std::int32_t date_from_day, std::int32_t date_to_day,
                    std::string const& tz_dir, gtfs_names& names) {
  if (date_to_day <= date_from_day || date_to_day - date_from_day + kTimetableOffsetDays + 1 > 512) {
    throw status_error{NIGIRI_AMD_EINVAL, "gtfs: bad date range"};
  }
  std::map<std::string, std::string> dir;
  for (auto const& [n, c] : files) dir[n] = c;
  auto const file = [&](std::string const& n) {
    auto const it = dir.find(n);
    return it == dir.end() ? csv{} : parse_csv(it->second);
  };
  for (auto const* unsupported : {"frequencies.txt", "location_groups.txt", "booking_rules.txt"}) {
    auto const it = dir.find(unsupported);
    if (it != dir.end() && !parse_csv(it->second).rows.empty()) {
      throw status_error{NIGIRI_AMD_EUNSUPPORTED, std::string{"gtfs: "} + unsupported + " is not supported"};
    }
  }

  timetable tt;
  tt.date_range_from = date_from_day * 1440;
  tt.date_range_to = date_to_day * 1440;
  auto const internal_day0 = date_from_day - kTimetableOffsetDays;
  tt.internal_from = internal_day0 * 1440;
  // internal_interval_days (timetable.h:232-237): [from - 5 d, to + 1 d)
  auto const n_days = static_cast<std::uint32_t>(date_to_day + 1 - internal_day0);
  tt.n_days = n_days;

  // ---- agencies ----------------------------------------------------------
  auto const agencies = file("agency.txt");
  std::map<std::string, std::string> agency_tz;
  std::string first_tz;
  for (auto const& r : agencies.rows) {
    auto const tzname = agencies.get(r, agencies.col("agency_timezone"));
    agency_tz[agencies.get(r, agencies.col("agency_id"))] = tzname;
    if (first_tz.empty()) first_tz = tzname;
  }
  if (first_tz.empty()) throw status_error{NIGIRI_AMD_EINVAL, "gtfs: agency.txt names no time zone"};
  auto const tz = load_time_zone(tz_dir, first_tz);

  // ---- stops: special stations first (special_stations.h:7-22) -----------
  names.location_ids = {"START", "END", "VIA0", "VIA1", "VIA2", "VIA3", "VIA4", "VIA5", "VIA6"};
  std::unordered_map<std::string, std::uint32_t> stop_idx;
  auto const stops = file("stops.txt");
  std::vector<bool> accessible(kNSpecialStations, true);
  for (auto const& r : stops.rows) {
    auto const type = to_int(stops.get(r, stops.col("location_type")), 0);
    if (type == 2 || type == 3) continue;  // entrance / generic node (stop.cc:237-241)
    auto const id = stops.get(r, stops.col("stop_id"));
    if (id.empty() || stop_idx.count(id) != 0U) continue;
    stop_idx[id] = static_cast<std::uint32_t>(names.location_ids.size());
    names.location_ids.push_back(id);
    accessible.push_back(to_int(stops.get(r, stops.col("wheelchair_boarding")), 0) != 2);
    auto const lat = stops.get(r, stops.col("stop_lat")), lon = stops.get(r, stops.col("stop_lon"));
    tt.coord_x.push_back(lon.empty() ? 0.0F : std::stof(lon));
    tt.coord_y.push_back(lat.empty() ? 0.0F : std::stof(lat));
  }
  auto const L = static_cast<std::uint32_t>(names.location_ids.size());
  tt.n_locations = L;
  tt.coord_x.insert(tt.coord_x.begin(), kNSpecialStations, 0.0F);
  tt.coord_y.insert(tt.coord_y.begin(), kNSpecialStations, 0.0F);
  tt.transfer_time.assign(L, 2U);  // loader_config::default_transfer_time_
  for (auto i = 0U; i != kNSpecialStations; ++i) tt.transfer_time[i] = 0U;

  // ---- transfers.txt (stop.cc:129-194, 352-363) ----------------------------
  std::vector<std::vector<std::pair<std::uint32_t, std::uint32_t>>> fps(L);
  std::vector<int> own_transfer(L, -1);
  {
    auto const tr = file("transfers.txt");
    for (auto const& r : tr.rows) {
      auto const type = to_int(tr.get(r, tr.col("transfer_type")), 0);
      if (type == 3) continue;  // not possible
      if (type == 4) {
        throw status_error{NIGIRI_AMD_EUNSUPPORTED, "gtfs: stay-seated transfers are not supported"};
      }
      auto const a = stop_idx.find(tr.get(r, tr.col("from_stop_id")));
      auto const b = stop_idx.find(tr.get(r, tr.col("to_stop_id")));
      if (a == stop_idx.end() || b == stop_idx.end()) continue;  // logged and skipped
      auto const minutes = to_int(tr.get(r, tr.col("min_transfer_time")), 0) / 60;
      if (a->second == b->second) {
        auto& t = own_transfer[a->second];
        t = t < 0 ? minutes : std::min(t, minutes);
        continue;
      }
      auto& v = fps[a->second];
      if (std::none_of(v.begin(), v.end(), [&](auto const& x) { return x.first == b->second; })) {
        v.emplace_back(b->second, static_cast<std::uint32_t>(minutes));
      }
    }
    for (auto l = 0U; l != L; ++l) {
      if (own_transfer[l] >= 0) tt.transfer_time[l] = static_cast<std::uint8_t>(std::min(own_transfer[l], 255));
    }
    // symmetric: add the reverse direction where it is missing
    auto const fwd = fps;
    for (auto l = 0U; l != L; ++l) {
      for (auto const& [t, d] : fwd[l]) {
        auto& v = fps[t];
        if (std::none_of(v.begin(), v.end(), [&](auto const& x) { return x.first == l; })) v.emplace_back(l, d);
      }
    }
  }

  // ---- services (services.cc:12-73) ----------------------------------------
  std::map<std::string, bitfield> service;
  {
    auto const cal = file("calendar.txt");
    static char const* const kWeekdays[] = {"sunday", "monday", "tuesday", "wednesday", "thursday", "friday",
                                            "saturday"};
    for (auto const& r : cal.rows) {
      bitfield b{};
      auto const from = parse_date(cal.get(r, cal.col("start_date")));
      auto const to = parse_date(cal.get(r, cal.col("end_date"))) + 1;
      for (auto d = std::max(from, internal_day0); d < std::min(to, internal_day0 + static_cast<std::int32_t>(n_days));
           ++d) {
        auto const wd = static_cast<int>(((d % 7) + 11) % 7);  // 0 = Sunday
        if (cal.get(r, cal.col(kWeekdays[wd])) == "1") set_bit(b, d - internal_day0);
      }
      service[cal.get(r, cal.col("service_id"))] = b;
    }
    auto const dates = file("calendar_dates.txt");
    for (auto const& r : dates.rows) {
      auto& b = service[dates.get(r, dates.col("service_id"))];
      auto const day = parse_date(dates.get(r, dates.col("date"))) - internal_day0;
      set_bit(b, day, to_int(dates.get(r, dates.col("exception_type")), 1) == 1);
    }
  }

  // ---- routes, trips, stop times -------------------------------------------
  std::map<std::string, int> route_type;
  {
    auto const routes = file("routes.txt");
    for (auto const& r : routes.rows) {
      route_type[routes.get(r, routes.col("route_id"))] = to_int(routes.get(r, routes.col("route_type")), 3);
    }
  }
  std::vector<trip> trips;
  std::unordered_map<std::string, std::size_t> trip_idx;
  {
    auto const tf = file("trips.txt");
    for (auto const& r : tf.rows) {
      if (!tf.get(r, tf.col("block_id")).empty()) {
        throw status_error{NIGIRI_AMD_EUNSUPPORTED, "gtfs: block_id (stay-seated trips) is not supported"};
      }
      trip t;
      t.id = tf.get(r, tf.col("trip_id"));
      auto const rt = route_type.find(tf.get(r, tf.col("route_id")));
      if (rt == route_type.end()) continue;  // unknown route: logged and skipped
      t.route_type = rt->second;
      t.service = tf.get(r, tf.col("service_id"));
      trip_idx[t.id] = trips.size();
      trips.push_back(std::move(t));
    }
    auto const st = file("stop_times.txt");
    auto const c_trip = st.col("trip_id"), c_arr = st.col("arrival_time"), c_dep = st.col("departure_time"),
               c_stop = st.col("stop_id"), c_seq = st.col("stop_sequence"), c_pick = st.col("pickup_type"),
               c_drop = st.col("drop_off_type");
    for (auto const& r : st.rows) {
      auto const t = trip_idx.find(st.get(r, c_trip));
      auto const s = stop_idx.find(st.get(r, c_stop));
      if (t == trip_idx.end() || s == stop_idx.end()) continue;  // logged and skipped
      auto const arr = hhmm_to_min(st.get(r, c_arr)), dep = hhmm_to_min(st.get(r, c_dep));
      if (arr == kInterpolate || dep == kInterpolate) {
        throw status_error{NIGIRI_AMD_EUNSUPPORTED, "gtfs: stop times without times (interpolation)"};
      }
      auto const in = to_int(st.get(r, c_pick), 0) != 1, out = to_int(st.get(r, c_drop), 0) != 1;
      auto const l = s->second;
      auto& tr = trips[t->second];
      tr.seq.emplace_back(to_int(st.get(r, c_seq), 0),
                          stop_pack(l, in, out, accessible[l] && in, accessible[l] && out));
      tr.times.emplace_back(arr, dep);
    }
    for (auto& tr : trips) {  // sort by stop_sequence
      std::vector<std::size_t> ord(tr.seq.size());
      for (std::size_t i = 0; i != ord.size(); ++i) ord[i] = i;
      std::stable_sort(ord.begin(), ord.end(), [&](auto a, auto b) { return tr.seq[a].first < tr.seq[b].first; });
      auto seq = tr.seq;
      auto times = tr.times;
      for (std::size_t i = 0; i != ord.size(); ++i) {
        tr.seq[i] = seq[ord[i]];
        tr.times[i] = times[ord[i]];
      }
    }
  }

  // ---- noon offsets of the first agency's zone (noon_offsets.cc:14-40) -----
  std::vector<std::int32_t> noon(n_days);
  for (auto d = 0U; d != n_days; ++d) noon[d] = noon_offset_minutes(tz, internal_day0 + static_cast<std::int32_t>(d));

  // ---- expand_local_to_utc + route grouping --------------------------------
  struct route_key {
    int type;
    std::vector<std::uint32_t> stops;
    bool operator<(route_key const& o) const { return std::tie(type, stops) < std::tie(o.type, o.stops); }
  };
  std::map<route_key, std::size_t> key_idx;
  std::vector<std::pair<route_key, std::vector<std::vector<utc_trip>>>> route_services;  // first-seen order
  auto const sel_from = date_from_day, sel_to = date_to_day;  // selection = date_range_
  for (std::uint32_t ti = 0; ti != trips.size(); ++ti) {
    auto const& t = trips[ti];
    if (t.times.size() <= 1U) continue;  // invalid event times: skipped (local_to_utc.h:159-166)
    auto const sv = service.find(t.service);
    if (sv == service.end()) continue;
    auto const first_dep = t.times.front().second, last_arr = t.times.back().first;
    auto const first_day_offset = first_dep / 1440, last_day_offset = last_arr / 1440;
    std::vector<std::pair<std::pair<std::int32_t, std::int32_t>, bitfield>> keyed;  // insertion order
    for (std::uint32_t d = 0; d != n_days; ++d) {
      auto const day = internal_day0 + static_cast<std::int32_t>(d);
      // service days [day + first offset, day + last offset + 1) overlap the selection
      if (!(day + first_day_offset < sel_to && sel_from < day + last_day_offset + 1)) continue;
      if (!test_bit(sv->second, d)) continue;
      auto const tz_off = noon[d];
      auto const first_dep_utc = first_dep - tz_off;
      auto const day_off = static_cast<std::int32_t>(std::floor(static_cast<double>(first_dep_utc) / 1440.0));
      auto const utc_day = static_cast<std::int64_t>(d) + day_off;
      if (utc_day < 0 || utc_day >= 512) continue;
      auto const key = std::make_pair(day_off, tz_off);
      auto it = std::find_if(keyed.begin(), keyed.end(), [&](auto const& x) { return x.first == key; });
      if (it == keyed.end()) {
        keyed.emplace_back(key, bitfield{});
        it = std::prev(keyed.end());
      }
      set_bit(it->second, utc_day);
    }
    for (auto const& [key, days] : keyed) {
      utc_trip u;
      u.trip = ti;
      u.days = days;
      for (std::size_t i = 0; i + 1U < t.times.size(); ++i) {
        u.utc_times.push_back(t.times[i].second - key.second - key.first * 1440);
        u.utc_times.push_back(t.times[i + 1U].first - key.second - key.first * 1440);
      }
      std::int32_t pred = 0;
      for (auto& x : u.utc_times) {
        x = std::max(pred, x);
        pred = x;
      }
      route_key rk{t.route_type, {}};
      for (auto const& s : t.seq) rk.stops.push_back(s.second);
      auto const [kit, inserted] = key_idx.try_emplace(rk, route_services.size());
      if (inserted) route_services.emplace_back(rk, std::vector<std::vector<utc_trip>>{});
      auto& subs = route_services[kit->second].second;
      // get_index (get_index.h:8-32): first sub-route where the trip keeps
      // FIFO order modulo 1440 at every event
      bool placed = false;
      for (auto& sub : subs) {
        auto const pos = static_cast<std::size_t>(
            std::lower_bound(sub.begin(), sub.end(), u,
                             [](utc_trip const& a, utc_trip const& b) {
                               return a.utc_times.front() % 1440 < b.utc_times.front() % 1440;
                             }) -
            sub.begin());
        bool ok = true;
        for (std::size_t i = 0; i != u.utc_times.size() && ok; ++i) {
          auto const v = u.utc_times[i] % 1440;
          if (pos > 0U && v < sub[pos - 1U].utc_times[i] % 1440) ok = false;
          if (pos < sub.size() && v > sub[pos].utc_times[i] % 1440) ok = false;
        }
        if (ok) {
          sub.insert(sub.begin() + static_cast<std::ptrdiff_t>(pos), std::move(u));
          placed = true;
          break;
        }
      }
      if (!placed) subs.push_back({std::move(u)});
    }
  }

  // ---- write routes, transports and events (load_timetable.cc:404-474) -----
  std::map<bitfield, std::uint32_t> bf_idx;
  for (auto const& [key, subs] : route_services) {
    for (auto const& sub : subs) {
      tt.route_stops.insert(tt.route_stops.end(), key.stops.begin(), key.stops.end());
      tt.route_stop_offsets.push_back(static_cast<std::uint32_t>(tt.route_stops.size()));
      for (auto const& s : sub) {
        auto const [it, inserted] = bf_idx.try_emplace(s.days, static_cast<std::uint32_t>(bf_idx.size()));
        if (inserted) tt.bitfields.insert(tt.bitfields.end(), s.days.begin(), s.days.end());
        tt.transport_traffic_days.push_back(it->second);
        names.transport_trip.push_back(trips[s.trip].id);
      }
      tt.route_transport_from.push_back(static_cast<std::uint32_t>(tt.transport_traffic_days.size()));
      auto const n_ev = 2U * (key.stops.size() - 1U);
      for (std::size_t ev = 0; ev != n_ev; ++ev) {
        for (auto const& s : sub) {
          auto const m = s.utc_times[ev];
          if (m < 0 || m / 1440 > 31) throw status_error{NIGIRI_AMD_EUNSUPPORTED, "gtfs: trip longer than 31 days"};
          tt.event_times.push_back(delta_pack(static_cast<std::uint32_t>(m)));
        }
      }
      tt.route_stop_time_from.push_back(static_cast<std::uint32_t>(tt.event_times.size()));
    }
  }

  // ---- footpaths_out_ (sorted by duration in finalize) ---------------------
  tt.fp_out_offsets.assign(L + 1U, 0U);
  for (auto l = 0U; l != L; ++l) {
    auto v = fps[l];
    std::stable_sort(v.begin(), v.end(), [](auto const& a, auto const& b) { return a.second < b.second; });
    for (auto const& [t, d] : v) tt.fp_out.push_back(fp_pack(t, d));
    tt.fp_out_offsets[l + 1U] = static_cast<std::uint32_t>(tt.fp_out.size());
  }
  finalize(tt);
  return tt;
}

}  // namespace nigiri_amd
