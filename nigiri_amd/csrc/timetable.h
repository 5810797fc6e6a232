// Host-side timetable: the arrays of nigiri::timetable that the RAPTOR path
// reads (include/nigiri/timetable.h:342-398 and 89-98, 421-422), stored as
// flat CSR vectors so that they upload to the device image unchanged.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/nigiri_amd.h"
#include "ch.h"

namespace nigiri_amd {

constexpr std::uint32_t kNSpecialStations = 9U;  // special_stations.h:7-22
constexpr std::uint32_t kEnd = 1U;               // special_station::kEnd
constexpr int kMaxTransfers = 14;                // routing/limits.h:9
constexpr int kMaxK = kMaxTransfers + 2;         // rows of round_times
constexpr int kMaxTravelTime = 7200;             // 5 days, limits.h:10
constexpr int kTimetableOffsetDays = 5;          // types.h:53-54
constexpr std::uint32_t kBitfieldWords = 8U;     // bitset<512>
constexpr std::uint32_t kFpTargetBits = 23U;     // footpath.h:14-20
constexpr std::uint32_t kFpTargetMask = (1U << kFpTargetBits) - 1U;
constexpr std::uint32_t kFpMaxDuration = 511U;
constexpr std::uint32_t kStopLocMask = (1U << 28) - 1U;  // stop.h:85

inline std::uint32_t fp_pack(std::uint32_t target, std::uint32_t dur) {
  if (dur > kFpMaxDuration) dur = kFpMaxDuration;
  return (target & kFpTargetMask) | (dur << kFpTargetBits);
}
inline std::uint32_t fp_target(std::uint32_t fp) { return fp & kFpTargetMask; }
inline std::uint32_t fp_duration(std::uint32_t fp) {
  return fp >> kFpTargetBits;
}

// delta {days:5, mam:11} (types.h:569-632), packed as nigiri's bitfield
// layout: days in the low 5 bits, minutes after midnight above.
inline std::uint16_t delta_pack(std::uint32_t minutes) {
  return static_cast<std::uint16_t>(((minutes % 1440U) << 5) |
                                    (minutes / 1440U));
}
inline int delta_days(std::uint16_t d) { return d & 31; }
inline int delta_mam(std::uint16_t d) { return d >> 5; }
inline int delta_count(std::uint16_t d) {
  return delta_days(d) * 1440 + delta_mam(d);
}

inline std::uint32_t stop_pack(std::uint32_t loc, bool in, bool out,
                               bool in_wc, bool out_wc) {
  return (loc & kStopLocMask) | (in ? 1U << 28 : 0U) | (out ? 1U << 29 : 0U) |
         (in_wc ? 1U << 30 : 0U) | (out_wc ? 1U << 31 : 0U);
}

struct timetable {
  std::uint32_t n_locations = kNSpecialStations;
  std::int32_t date_range_from = 0, date_range_to = 0, internal_from = 0;
  std::uint32_t n_days = 0;

  std::vector<std::uint32_t> route_stop_offsets{0U}, route_stops;
  std::vector<std::uint32_t> route_transport_from{0U};
  std::vector<std::uint32_t> route_stop_time_from{0U};
  std::vector<std::uint16_t> event_times;
  std::vector<std::uint32_t> transport_traffic_days, route_traffic_days;
  std::vector<std::uint64_t> bitfields;
  std::vector<std::uint32_t> location_route_offsets, location_routes;
  std::vector<std::uint8_t> transfer_time;
  std::vector<std::uint32_t> fp_out_offsets, fp_out, fp_in_offsets, fp_in;
  std::vector<std::uint32_t> lb_fwd_offsets, lb_fwd, lb_bwd_offsets, lb_bwd;
  std::vector<float> coord_x, coord_y;
  // location_importance_ / max_importance_ as assign_importance computes
  // them (src/loader/init_finish.cc:64-101).  Routes carry no transport
  // class here: every trip counts with the priority of class kOther (1).
  std::vector<std::uint32_t> location_importance;
  std::uint32_t max_importance = 0U;
  // contraction hierarchy of lb_fwd (exact lower bounds by PHAST)
  contraction_hierarchy ch;

  std::uint32_t n_routes() const {
    return static_cast<std::uint32_t>(route_stop_offsets.size() - 1U);
  }
  std::uint32_t n_transports() const { return route_transport_from.back(); }
  std::uint32_t n_bitfields() const {
    return static_cast<std::uint32_t>(bitfields.size() / kBitfieldWords);
  }
  nigiri_tt_view view() const;
};

// Rebuilds the derived arrays the way nigiri's finalize() does
// (src/loader/init_finish.cc:200-234): location_routes_, route_traffic_days_
// (init_finish.cc:168-198, appending one bitfield per route), footpaths_in_
// as the duration-sorted transpose of footpaths_out_ (build_footpaths.cc:
// 404-436), and both lower-bound graphs (loader/build_lb_graph.h:11-99);
// then the contraction hierarchy of the lower-bound graph.
void finalize(timetable& tt);

// Checks structural invariants (CSR bounds, FIFO order modulo 1440 inside
// every route, sorted footpaths).  Returns an empty string when valid.
std::string validate(timetable const& tt);

}  // namespace nigiri_amd
