// C ABI: host timetable and query-workload entry points.
#include <algorithm>
#include <cstring>
#include <exception>
#include <string>

#include "capi_util.h"
#include "generator.h"
#include "gtfs.h"
#include "handles.h"
#include "timetable.h"

namespace nigiri_amd {
thread_local std::string g_last_error;
void set_error(std::string msg) { g_last_error = std::move(msg); }
}  // namespace nigiri_amd


using namespace nigiri_amd;

extern "C" {

const char* nigiri_amd_last_error(void) { return g_last_error.c_str(); }

int nigiri_amd_version(void) { return 1; }

int nigiri_amd_tt_generate(int cfg, uint64_t seed, double scale, nigiri_amd_tt** out) {
  if (out == nullptr || !(scale > 0.0 && scale <= 1.0)) {
    set_error("tt_generate: bad argument");
    return NIGIRI_AMD_EINVAL;
  }
  return guarded([&]() {
    auto h = std::make_unique<nigiri_amd_tt>();
    h->tt = generate(cfg, seed, scale);
    auto const err = validate(h->tt);
    if (!err.empty()) {
      set_error("tt_generate: generated timetable invalid: " + err);
      return NIGIRI_AMD_EINVAL;
    }
    *out = h.release();
    return NIGIRI_AMD_OK;
  });
}

int nigiri_amd_tt_from_arrays(const nigiri_tt_view* v, nigiri_amd_tt** out) {
  if (v == nullptr || out == nullptr || v->n_locations < kNSpecialStations) {
    set_error("tt_from_arrays: bad argument");
    return NIGIRI_AMD_EINVAL;
  }
  return guarded([&]() {
    auto h = std::make_unique<nigiri_amd_tt>();
    auto& tt = h->tt;
    auto const L = v->n_locations, R = v->n_routes;
    tt.n_locations = L;
    tt.date_range_from = v->date_range_from;
    tt.date_range_to = v->date_range_to;
    tt.internal_from = v->internal_from;
    tt.n_days = v->n_days;
    tt.route_stop_offsets.assign(v->route_stop_offsets, v->route_stop_offsets + R + 1U);
    tt.route_stops.assign(v->route_stops, v->route_stops + tt.route_stop_offsets.back());
    tt.route_transport_from.assign(v->route_transport_from, v->route_transport_from + R + 1U);
    tt.route_stop_time_from.assign(v->route_stop_time_from, v->route_stop_time_from + R + 1U);
    tt.event_times.assign(v->event_times, v->event_times + tt.route_stop_time_from.back());
    tt.transport_traffic_days.assign(v->transport_traffic_days,
                                     v->transport_traffic_days + tt.route_transport_from.back());
    tt.bitfields.assign(v->bitfields, v->bitfields + static_cast<std::size_t>(v->n_bitfields) * 8U);
    tt.transfer_time.assign(v->transfer_time, v->transfer_time + L);
    tt.fp_out_offsets.assign(v->fp_out_offsets, v->fp_out_offsets + L + 1U);
    tt.fp_out.assign(v->fp_out, v->fp_out + tt.fp_out_offsets.back());
    if (v->coord_x != nullptr && v->coord_y != nullptr) {
      tt.coord_x.assign(v->coord_x, v->coord_x + L);
      tt.coord_y.assign(v->coord_y, v->coord_y + L);
    }
    for (auto const bf : tt.transport_traffic_days) {
      if (bf >= v->n_bitfields) {
        set_error("tt_from_arrays: bitfield index out of range");
        return NIGIRI_AMD_EINVAL;
      }
    }
    for (auto const fp : tt.fp_out) {
      if (fp_target(fp) >= L) {
        set_error("tt_from_arrays: footpath target out of range");
        return NIGIRI_AMD_EINVAL;
      }
    }
    finalize(tt);
    auto const err = validate(tt);
    if (!err.empty()) {
      set_error("tt_from_arrays: " + err);
      return NIGIRI_AMD_EINVAL;
    }
    *out = h.release();
    return NIGIRI_AMD_OK;
  });
}

int nigiri_amd_tt_load_gtfs(const char* const* file_names, const char* const* file_contents, uint32_t n_files,
                            int32_t date_from_day, int32_t date_to_day, const char* tz_dir, nigiri_amd_tt** out) {
  if (out == nullptr || (n_files != 0U && (file_names == nullptr || file_contents == nullptr))) {
    set_error("tt_load_gtfs: null argument");
    return NIGIRI_AMD_EINVAL;
  }
  return guarded([&]() {
    std::vector<std::pair<std::string, std::string>> files;
    for (uint32_t i = 0; i != n_files; ++i) {
      if (file_names[i] == nullptr || file_contents[i] == nullptr) {
        throw status_error{NIGIRI_AMD_EINVAL, "tt_load_gtfs: null file"};
      }
      files.emplace_back(file_names[i], file_contents[i]);
    }
    auto h = std::make_unique<nigiri_amd_tt>();
    gtfs_names names;
    h->tt = load_gtfs(files, date_from_day, date_to_day, tz_dir == nullptr ? "" : tz_dir, names);
    auto const err = validate(h->tt);
    if (!err.empty()) throw status_error{NIGIRI_AMD_EINVAL, "tt_load_gtfs: timetable invalid: " + err};
    h->location_ids = std::move(names.location_ids);
    h->transport_trip_ids = std::move(names.transport_trip);
    *out = h.release();
    return NIGIRI_AMD_OK;
  });
}

namespace {
int copy_id(std::vector<std::string> const& ids, uint32_t i, char* buf, uint32_t cap) {
  if (i >= ids.size() || buf == nullptr || cap <= ids[i].size()) {
    set_error("id lookup: index out of range or buffer too small");
    return NIGIRI_AMD_EINVAL;
  }
  std::memcpy(buf, ids[i].c_str(), ids[i].size() + 1U);
  return NIGIRI_AMD_OK;
}
}  // namespace

int nigiri_amd_tt_location_id(const nigiri_amd_tt* tt, uint32_t l, char* buf, uint32_t cap) {
  if (tt == nullptr) return NIGIRI_AMD_EINVAL;
  return copy_id(tt->location_ids, l, buf, cap);
}

int nigiri_amd_tt_find_location(const nigiri_amd_tt* tt, const char* stop_id, uint32_t* l) {
  if (tt == nullptr || stop_id == nullptr || l == nullptr) return NIGIRI_AMD_EINVAL;
  for (std::size_t i = kNSpecialStations; i < tt->location_ids.size(); ++i) {
    if (tt->location_ids[i] == stop_id) {
      *l = static_cast<uint32_t>(i);
      return NIGIRI_AMD_OK;
    }
  }
  set_error(std::string{"find_location: unknown stop_id "} + stop_id);
  return NIGIRI_AMD_EINVAL;
}

int nigiri_amd_tt_transport_trip_id(const nigiri_amd_tt* tt, uint32_t t, char* buf, uint32_t cap) {
  if (tt == nullptr) return NIGIRI_AMD_EINVAL;
  return copy_id(tt->transport_trip_ids, t, buf, cap);
}

int nigiri_amd_tt_lower_bounds(const nigiri_amd_tt* tt, int32_t dir, uint32_t n_seeds, const uint32_t* seed_loc,
                               const uint16_t* seed_min, uint16_t* lb) {
  if (tt == nullptr || lb == nullptr || (n_seeds != 0U && (seed_loc == nullptr || seed_min == nullptr)) ||
      (dir != NIGIRI_AMD_DIR_FWD && dir != NIGIRI_AMD_DIR_BWD)) {
    set_error("tt_lower_bounds: bad argument");
    return NIGIRI_AMD_EINVAL;
  }
  return guarded([&]() {
    std::vector<std::pair<std::uint32_t, std::uint16_t>> seeds;
    for (uint32_t i = 0; i != n_seeds; ++i) {
      if (seed_loc[i] >= tt->tt.n_locations) throw status_error{NIGIRI_AMD_EINVAL, "tt_lower_bounds: seed out of range"};
      seeds.emplace_back(seed_loc[i], seed_min[i]);
    }
    ch_search_scratch s;
    std::vector<std::uint16_t> out;
    ch_lower_bounds(tt->tt.ch, dir == NIGIRI_AMD_DIR_FWD, seeds, s, out);
    std::copy(out.begin(), out.end(), lb);
    return NIGIRI_AMD_OK;
  });
}

int nigiri_amd_tt_view(const nigiri_amd_tt* tt, nigiri_tt_view* out) {
  if (tt == nullptr || out == nullptr) {
    set_error("tt_view: null argument");
    return NIGIRI_AMD_EINVAL;
  }
  *out = tt->tt.view();
  return NIGIRI_AMD_OK;
}

void nigiri_amd_tt_free(nigiri_amd_tt* tt) { delete tt; }

int nigiri_amd_queries_generate(const nigiri_amd_tt* tt, int cfg, uint64_t seed,
                                uint32_t n_queries, nigiri_query* out,
                                uint32_t* offset_loc, int16_t* offset_dur,
                                uint32_t* n_offsets) {
  if (tt == nullptr || n_offsets == nullptr) {
    set_error("queries_generate: null argument");
    return NIGIRI_AMD_EINVAL;
  }
  return guarded([&]() {
    auto const qs = generate_queries(tt->tt, cfg, seed, n_queries);
    std::size_t need = 0U;
    for (auto const& g : qs) need += g.start_loc.size() + g.dest_loc.size();
    if (out == nullptr) {
      *n_offsets = static_cast<uint32_t>(need);
      return NIGIRI_AMD_OK;
    }
    if (need > *n_offsets || offset_loc == nullptr || offset_dur == nullptr) {
      set_error("queries_generate: offset arrays too small");
      return NIGIRI_AMD_ECAPACITY;
    }
    std::size_t pos = 0U;
    for (auto i = 0U; i != qs.size(); ++i) {
      auto const& g = qs[i];
      out[i] = g.q;
      out[i].n_start = static_cast<uint32_t>(g.start_loc.size());
      out[i].start_loc = offset_loc + pos;
      out[i].start_dur = offset_dur + pos;
      std::memcpy(offset_loc + pos, g.start_loc.data(), g.start_loc.size() * sizeof(uint32_t));
      std::memcpy(offset_dur + pos, g.start_dur.data(), g.start_dur.size() * sizeof(int16_t));
      pos += g.start_loc.size();
      out[i].n_dest = static_cast<uint32_t>(g.dest_loc.size());
      out[i].dest_loc = offset_loc + pos;
      out[i].dest_dur = offset_dur + pos;
      std::memcpy(offset_loc + pos, g.dest_loc.data(), g.dest_loc.size() * sizeof(uint32_t));
      std::memcpy(offset_dur + pos, g.dest_dur.data(), g.dest_dur.size() * sizeof(int16_t));
      pos += g.dest_loc.size();
    }
    *n_offsets = static_cast<uint32_t>(pos);
    return NIGIRI_AMD_OK;
  });
}

}  // extern "C"
