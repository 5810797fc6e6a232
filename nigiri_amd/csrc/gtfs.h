// GTFS feed -> host timetable, restating the parts of nigiri's GTFS loader
// (src/loader/gtfs/load_timetable.cc:84-545) that the RAPTOR path reads.
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "timetable.h"

namespace nigiri_amd {

struct gtfs_names {
  std::vector<std::string> location_ids;   // per location (specials: nigiri's names)
  std::vector<std::string> transport_trip; // per transport: GTFS trip_id
};

// files: (file name, content), as nigiri's loader::mem_dir holds them.
// date_from_day / date_to_day: timetable::date_range_ in days since
// 1970-01-01 ([from, to)).  tz_dir: directory of TZif files (empty: $TZDIR
// or /usr/share/zoneinfo).  Throws status_error / std::invalid_argument.
timetable load_gtfs(std::vector<std::pair<std::string, std::string>> const& files,
                    std::int32_t date_from_day, std::int32_t date_to_day,
                    std::string const& tz_dir, gtfs_names& names);

}  // namespace nigiri_amd
