// Read access to an engine handle for host drivers outside engine.hip.
#pragma once

#include <cstdint>

#include "../../include/nigiri_amd.h"
#include "timetable.h"

namespace nigiri_amd {

timetable const& engine_timetable(nigiri_amd_engine const* e);
std::uint32_t engine_lanes(nigiri_amd_engine const* e);  // 1 or 64
std::uint32_t engine_slots(nigiri_amd_engine const* e);
std::uint32_t engine_max_k(nigiri_amd_engine const* e);

}  // namespace nigiri_amd
