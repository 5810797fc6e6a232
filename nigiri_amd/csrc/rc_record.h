// Records of the device journey reconstruction (see launch_reconstruct in
// kernels.h), shared by the kernels and the host code that writes the legs.
#pragma once

#include <cstdint>

namespace nigiri_amd {

constexpr int kRcMaxSteps = 15;  // rounds k = 1 .. 15

struct rc_job {
  std::uint32_t slot;
  std::uint32_t dest;       // journey destination (END for intermodal)
  std::uint32_t off_begin;  // egress offsets off_loc/off_dur[off_begin ..
  std::uint16_t n_off;      //   off_begin + n_off), sanitized order
  std::uint8_t transfers;
  std::uint8_t intermodal;
};

struct rc_step {
  std::uint32_t transport;
  std::uint32_t route;
  std::uint16_t day;         // traffic day of the trip
  std::uint16_t stop_idx;    // trip stop where the round's ride begins (search direction)
  std::uint16_t finish_idx;  // trip stop where the walk to the round's location begins
  std::int16_t curr_time;    // round_times[k][location of the round]
  std::int32_t walk;         // -1: transfer at the location; >= 0: index in the
                             // location's footpath list (footpaths_in forward,
                             // footpaths_out backward) or egress offset index
};

struct rc_out {
  std::int32_t status;       // 0: reconstructed, 1: some round has no leg
  std::uint32_t n_steps;
  std::uint32_t start_loc;   // where the ride of the last step begins
  std::int16_t rt0;          // round_times[0][start_loc]
  std::int16_t pad;
  rc_step steps[kRcMaxSteps];
};

}  // namespace nigiri_amd
