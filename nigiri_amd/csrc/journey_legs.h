// Host half of the journey reconstruction: the device records, for every
// round of a journey, which walk and which trip reach the round's location
// (rc_out, filled by launch_reconstruct); this file writes the legs nigiri's
// reconstruct_journey produces from them (src/routing/raptor/reconstruct.cc:
// 30-160 and 799-877) and applies optimize_footpaths
// (src/routing/raptor/optimize_footpaths.cc:15-453).
#pragma once

#include "rc_record.h"
#include "search_host.h"

namespace nigiri_amd {

// Fills j.legs_ (travel order) and sets j.reconstructed_, or sets j.error_
// when the device found no legs or no first walk exists (the exceptions of
// reconstruct.cc:159, 738-741, 794-796).
void assemble_legs(timetable const& tt, prepared_query const& p, rc_out const& o,
                   host_journey& j);

// optimize_footpaths without via stops on legs in travel order; start and
// dest are the journey-direction query's offsets in sanitized order.
void optimize_footpaths(timetable const& tt, host_tts const& tts,
                        std::vector<std::pair<std::uint32_t, std::int16_t>> const& start,
                        std::vector<std::pair<std::uint32_t, std::int16_t>> const& dest,
                        std::vector<nigiri_leg>& legs);

}  // namespace nigiri_amd
