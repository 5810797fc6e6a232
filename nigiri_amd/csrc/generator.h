// Seeded synthetic timetables and query workloads (generator.cc).
#pragma once

#include <cstdint>
#include <vector>

#include "timetable.h"

namespace nigiri_amd {

struct gen_query {
  nigiri_query q;
  std::vector<std::uint32_t> start_loc, dest_loc;
  std::vector<std::int16_t> start_dur, dest_dur;
};

timetable generate(int cfg_id, std::uint64_t seed, double scale);
std::vector<gen_query> generate_queries(timetable const& tt, int cfg_id,
                                        std::uint64_t seed, std::uint32_t n);

}  // namespace nigiri_amd
