// Time zones for the GTFS loader: TZif files (RFC 8536, version 2 and
// later, as shipped by tzdata) with their POSIX TZ footer rule for instants
// after the last listed transition.  Replaces the date library's
// time_zone lookups that nigiri's loader makes (noon_offsets.cc:7-12).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace nigiri_amd {

struct posix_rule {
  bool valid = false;
  std::int32_t std_off = 0;  // UTC offset of standard time, seconds
  bool has_dst = false;
  std::int32_t dst_off = 0;
  // transition dates: kind 'M' (month, week 1-5, weekday 0-6), 'J' (day
  // 1-365, no leap day) or 'D' (day 0-365); time of day in seconds
  struct date_rule {
    char kind = 'M';
    int m = 0, w = 0, d = 0, n = 0;
    std::int32_t time = 7200;
  } start, end;
};

struct time_zone {
  std::string name;
  std::vector<std::int64_t> transitions;  // UTC seconds, ascending
  std::vector<std::uint8_t> trans_type;
  std::vector<std::int32_t> type_off;     // UTC offset per local time type, seconds
  posix_rule footer;
  // UTC offset in seconds in force at the UTC instant t (seconds)
  std::int32_t offset_at(std::int64_t t) const;
};

// Parses TZif data; throws std::invalid_argument on malformed input.
time_zone parse_tzif(std::string const& name, std::string const& bytes);

// Reads <dir>/<name> (e.g. "Europe/Berlin").  dir empty: $TZDIR, else
// /usr/share/zoneinfo.  Throws std::invalid_argument if it is missing.
time_zone load_time_zone(std::string const& dir, std::string const& name);

// Parses a POSIX TZ string such as "CET-1CEST,M3.5.0,M10.5.0/3".
posix_rule parse_posix_tz(std::string const& s);

// UTC offset in minutes at local noon of the local calendar day `day`
// (days since 1970-01-01): get_noon_offset (noon_offsets.cc:7-12).
std::int32_t noon_offset_minutes(time_zone const& tz, std::int32_t day);

// Civil date helpers (proleptic Gregorian).
std::int32_t days_from_civil(int y, unsigned m, unsigned d);
void civil_from_days(std::int32_t z, int& y, unsigned& m, unsigned& d);

}  // namespace nigiri_amd
