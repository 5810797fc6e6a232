/*
 * oracle.h — C entry points of the CPU oracle (liboracle.so).
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a line-faithful CPU restatement
 * of nigiri's RAPTOR search path (Rt = false, Vias = 0, kOneToOne, forward
 * and backward).  Only tests/, the smoke check and the CPU-baseline leg of
 * the benchmark may load it; the product library never links or calls it.
 *
 * Parity status per function (how each one is pinned) is listed in the
 * header comment of oracle.cc.
 */
#ifndef NIGIRI_ORACLE_H
#define NIGIRI_ORACLE_H

#include <stdint.h>

#include "../include/nigiri_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Work counters of one search (raptor_stats.h:9-52 plus the byte model). */
typedef struct oracle_stats {
  uint64_t n_routes_visited;
  uint64_t n_route_stops_visited;
  uint64_t n_earliest_trip_calls;
  uint64_t n_events_scanned;
  uint64_t n_route_day_tests;
  uint64_t n_transport_day_tests;
  uint64_t n_marked_stations;
  uint64_t n_location_routes_read;
  uint64_t n_transfer_stations;
  uint64_t n_footpaths_visited;
  uint64_t n_label_writes;
  uint64_t n_rounds;
  uint64_t n_epochs;
  uint64_t alg_bytes; /* algorithmic bytes (model in oracle.cc)          */
} oracle_stats;

/* Full search of one query (search<Dir, raptor<...>>::execute), including
 * journey reconstruction (raptor/reconstruct.cc, optimize_footpaths.cc).
 * Journeys go to out[0 .. *n_out), their legs to the optional legs array
 * (*n_legs = legs needed).  State after the last epoch goes to the
 * optional arrays: round_times[16][n_locations], best/tmp[n_locations],
 * time_at_dest[16], lb[n_locations].  info = {base_day, fastest_direct,
 * interval_from, interval_to}. */
int oracle_search(const nigiri_tt_view* tt, const nigiri_query* q,
                  nigiri_journey* out, uint32_t cap, uint32_t* n_out,
                  nigiri_leg* legs, uint32_t leg_cap, uint32_t* n_legs,
                  int16_t* round_times, int16_t* best, int16_t* tmp,
                  int16_t* time_at_dest, uint16_t* lb, int32_t* info,
                  oracle_stats* stats);

/* Range search by pong (routing::pong_search, raptor/pong.cc:62-428):
 * ping/pong pairs of searches until the interval is covered.  Outputs as
 * in oracle_search; info = {base_day, fastest_direct, interval_from,
 * interval_to} of the result.  Returns -1 where the reference throws. */
int oracle_pong(const nigiri_tt_view* tt, const nigiri_query* q,
                nigiri_journey* out, uint32_t cap, uint32_t* n_out,
                nigiri_leg* legs, uint32_t leg_cap, uint32_t* n_legs,
                int32_t* info, oracle_stats* stats);

/* optimize_footpaths (raptor/optimize_footpaths.cc:444-453) on n legs in
 * travel order, in place; q is the journey-direction query. */
int oracle_optimize_footpaths(const nigiri_tt_view* tt, const nigiri_query* q,
                              nigiri_leg* legs, uint32_t n);

/* Lower bounds only (routing/dijkstra.cc:18-67). */
int oracle_lb(const nigiri_tt_view* tt, const nigiri_query* q, uint16_t* lb);

/* get_fastest_direct (routing/get_fastest_direct.cc:12-100). */
int oracle_fastest_direct(const nigiri_tt_view* tt, const nigiri_query* q,
                          int16_t* out);

/* get_starts + the search driver's sort (start_times.cc:241-319,
 * search.h:400-411). */
int oracle_starts(const nigiri_tt_view* tt, const nigiri_query* q,
                  int32_t* time_at_start, int32_t* time_at_stop,
                  uint32_t* stop, uint32_t cap, uint32_t* n_out);

/* Multi-threaded throughput run in the style of exe/benchmark.cc:210-270:
 * one state per thread, an atomic query counter, one warm-up query per
 * thread.  latency_ms[i] receives the time of query i. */
int oracle_search_many(const nigiri_tt_view* tt, const nigiri_query* qs,
                       uint32_t n, int threads, double* latency_ms,
                       double* wall_ms, uint32_t* n_journeys,
                       oracle_stats* total_stats);

#ifdef __cplusplus
}
#endif

#endif
