// Opaque handle types of the C ABI that wrap C++ objects.
#pragma once

#include <string>
#include <vector>

#include "timetable.h"

struct nigiri_amd_tt {
  nigiri_amd::timetable tt;
  // feed ids of a loaded feed (empty for generated timetables)
  std::vector<std::string> location_ids;
  std::vector<std::string> transport_trip_ids;
};
