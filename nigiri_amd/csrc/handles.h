// Opaque handle types of the C ABI that wrap C++ objects.
#pragma once

#include "timetable.h"

struct nigiri_amd_tt {
  nigiri_amd::timetable tt;
};
