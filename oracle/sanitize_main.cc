// Host-code sanitizer check (test infrastructure): the oracle and the
// library's host sources (timetable, generator, GTFS loader, time zones,
// contraction hierarchy, query preparation, journey legs, host C ABI) built
// with -fsanitize=address,undefined, run on generated workloads and on a
// GTFS feed.  No GPU code is built or run.  Besides running clean under
// the sanitizers, it checks that the host's contraction-hierarchy lower
// bounds equal the oracle's Dijkstra, that query preparation with host
// and device-style lower bounds agrees, and that the host's fastest direct
// walk equals the oracle's.
//
// usage: sanitize_check [gtfs_dir date_from_day date_to_day stop_from stop_to]
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "handles.h"
#include "nigiri_amd.h"
#include "oracle.h"
#include "search_host.h"

namespace {

int g_failures = 0;

void fail(char const* what, int cfg, unsigned i) {
  std::fprintf(stderr, "FAIL %s (workload %d, query %u): %s\n", what, cfg, i, nigiri_amd_last_error());
  ++g_failures;
}

void run_queries(nigiri_amd_tt* tt, int cfg, std::vector<nigiri_query> const& qs) {
  nigiri_tt_view v{};
  if (nigiri_amd_tt_view(tt, &v) != NIGIRI_AMD_OK) {
    fail("tt_view", cfg, 0U);
    return;
  }
  auto const L = v.n_locations;
  std::vector<nigiri_journey> jr(4096);
  std::vector<nigiri_leg> legs(1U << 15);
  std::vector<std::int16_t> rt(16U * L), best(L), tmp(L), tad(16);
  std::vector<std::uint16_t> lb(L), lb_ch(L);
  nigiri_amd::host_scratch scratch;
  for (unsigned i = 0; i != qs.size(); ++i) {
    auto const& q = qs[i];
    std::uint32_t n = 0, n_legs = 0;
    std::int32_t info[4] = {};
    oracle_stats st{};
    if (oracle_search(&v, &q, jr.data(), static_cast<std::uint32_t>(jr.size()), &n, legs.data(),
                      static_cast<std::uint32_t>(legs.size()), &n_legs, rt.data(), best.data(), tmp.data(),
                      tad.data(), lb.data(), info, &st) != 0) {
      fail("oracle_search", cfg, i);
      continue;
    }
    if (q.is_interval != 0 && (i % 2U) == 0U) {
      std::uint32_t np = 0, nl = 0;
      std::int32_t pinfo[4] = {};
      oracle_stats ps{};
      if (oracle_pong(&v, &q, jr.data(), static_cast<std::uint32_t>(jr.size()), &np, legs.data(),
                      static_cast<std::uint32_t>(legs.size()), &nl, pinfo, &ps) != 0) {
        fail("oracle_pong", cfg, i);
      }
    }
    // lower bounds: contraction hierarchy (host C ABI) against the oracle
    if (oracle_lb(&v, &q, lb.data()) != 0) fail("oracle_lb", cfg, i);
    std::vector<std::uint32_t> seed_loc;
    std::vector<std::uint16_t> seed_min;
    if (q.dest_match_mode == NIGIRI_AMD_MATCH_INTERMODAL) {
      for (std::uint32_t d = 0; d != q.n_dest; ++d) {
        seed_loc.push_back(q.dest_loc[d]);
        seed_min.push_back(static_cast<std::uint16_t>(q.dest_dur[d]));
      }
    }
    // exact destinations: the oracle seeds the destination's matching
    // locations; compare through query preparation instead (below)
    nigiri_amd::prepared_query p_host, p_dev;
    try {
      nigiri_amd::prepare(tt->tt, q, p_host, scratch, true);
      nigiri_amd::prepare(tt->tt, q, p_dev, scratch, false);
    } catch (std::exception const& e) {
      std::fprintf(stderr, "FAIL prepare (workload %d, query %u): %s\n", cfg, i, e.what());
      ++g_failures;
      continue;
    }
    if (p_host.lb_ != std::vector<std::uint16_t>(lb.begin(), lb.end())) fail("host lb != oracle lb", cfg, i);
    // fastest direct walk: the host function against the oracle, and the
    // node sets the device kernel gets
    std::int16_t fd_ref = 0;
    if (oracle_fastest_direct(&v, &q, &fd_ref) != 0 || fd_ref != p_host.fastest_direct_) {
      fail("fastest_direct != oracle", cfg, i);
    }
    std::vector<nigiri_amd::fd_node_host> fd_nodes;
    nigiri_amd::fastest_direct_nodes(tt->tt, q, fd_nodes);
    for (std::size_t j = 1; j < fd_nodes.size(); ++j) {
      if (fd_nodes[j - 1U].loc >= fd_nodes[j].loc) fail("fastest_direct_nodes order", cfg, i);
    }
    if (p_host.epochs_.size() != p_dev.epochs_.size() || p_host.lb_seeds_ != p_dev.lb_seeds_) {
      fail("prepare host / device-style", cfg, i);
    }
    if (!seed_loc.empty()) {
      if (nigiri_amd_tt_lower_bounds(tt, q.dir, static_cast<std::uint32_t>(seed_loc.size()), seed_loc.data(),
                                     seed_min.data(), lb_ch.data()) != NIGIRI_AMD_OK) {
        fail("tt_lower_bounds", cfg, i);
      } else if (lb_ch != lb) {
        fail("tt_lower_bounds != oracle lb", cfg, i);
      }
    }
  }
}

void generated(int cfg, double scale, unsigned n_queries) {
  nigiri_amd_tt* tt = nullptr;
  if (nigiri_amd_tt_generate(cfg, 0x51DEULL + static_cast<unsigned>(cfg), scale, &tt) != NIGIRI_AMD_OK) {
    fail("tt_generate", cfg, 0U);
    return;
  }
  std::uint32_t n_off = 0;
  nigiri_amd_queries_generate(tt, cfg, 0x7172ULL + static_cast<unsigned>(cfg), n_queries, nullptr, nullptr,
                              nullptr, &n_off);
  std::vector<nigiri_query> qs(n_queries);
  std::vector<std::uint32_t> off_loc(n_off + 1U);
  std::vector<std::int16_t> off_dur(n_off + 1U);
  if (nigiri_amd_queries_generate(tt, cfg, 0x7172ULL + static_cast<unsigned>(cfg), n_queries, qs.data(),
                                  off_loc.data(), off_dur.data(), &n_off) != NIGIRI_AMD_OK) {
    fail("queries_generate", cfg, 0U);
  } else {
    run_queries(tt, cfg, qs);
  }
  nigiri_amd_tt_free(tt);
  std::printf("workload %d (scale %.3f): %u queries checked\n", cfg, scale, n_queries);
}

void gtfs(char const* dir, int from_day, int to_day, char const* a, char const* c) {
  std::vector<std::string> names, contents;
  for (auto const& e : std::filesystem::directory_iterator(dir)) {
    std::ifstream in{e.path()};
    std::stringstream ss;
    ss << in.rdbuf();
    names.push_back(e.path().filename().string());
    contents.push_back(ss.str());
  }
  std::vector<char const*> np, cp;
  for (std::size_t i = 0; i != names.size(); ++i) {
    np.push_back(names[i].c_str());
    cp.push_back(contents[i].c_str());
  }
  nigiri_amd_tt* tt = nullptr;
  if (nigiri_amd_tt_load_gtfs(np.data(), cp.data(), static_cast<std::uint32_t>(np.size()), from_day, to_day,
                              nullptr, &tt) != NIGIRI_AMD_OK) {
    fail("tt_load_gtfs", 0, 0U);
    return;
  }
  std::uint32_t la = 0, lc = 0;
  if (nigiri_amd_tt_find_location(tt, a, &la) != NIGIRI_AMD_OK ||
      nigiri_amd_tt_find_location(tt, c, &lc) != NIGIRI_AMD_OK) {
    fail("tt_find_location", 0, 0U);
    nigiri_amd_tt_free(tt);
    return;
  }
  char buf[64];
  if (nigiri_amd_tt_location_id(tt, la, buf, sizeof(buf)) != NIGIRI_AMD_OK) fail("tt_location_id", 0, 0U);
  if (nigiri_amd_tt_transport_trip_id(tt, 0U, buf, sizeof(buf)) != NIGIRI_AMD_OK) fail("trip_id", 0, 0U);
  nigiri_tt_view v{};
  nigiri_amd_tt_view(tt, &v);
  std::vector<nigiri_query> qs;
  std::int16_t zero = 0;
  for (int dir = 0; dir != 2; ++dir) {
    nigiri_query q{};
    q.dir = dir;
    q.is_interval = 1;
    q.interval_from = v.date_range_from;
    q.interval_to = v.date_range_to;
    q.start_time = v.date_range_from;
    q.n_start = 1;
    q.start_loc = dir == 0 ? &la : &lc;
    q.start_dur = &zero;
    q.n_dest = 1;
    q.dest_loc = dir == 0 ? &lc : &la;
    q.dest_dur = &zero;
    q.max_start_offset = 7200;
    q.max_travel_time = 7200;
    q.max_transfers = 7;
    q.tts_default = dir == 0 ? 1 : 0;
    q.tts_factor = 1.5F;
    q.tts_min_transfer_time = 3;
    q.tts_additional_time = 1;
    q.fastest_direct_factor = 1.0;
    q.prf_idx = 0;
    qs.push_back(q);
  }
  run_queries(tt, 0, qs);
  nigiri_amd_tt_free(tt);
  std::printf("gtfs feed: %zu files, 2 queries checked\n", names.size());
}

}  // namespace

int main(int argc, char** argv) {
  generated(1, 1.0, 40);
  generated(2, 0.02, 24);
  generated(3, 0.01, 24);
  generated(4, 0.01, 16);
  if (argc >= 6) gtfs(argv[1], std::atoi(argv[2]), std::atoi(argv[3]), argv[4], argv[5]);
  if (g_failures != 0) {
    std::fprintf(stderr, "%d failures\n", g_failures);
    return 1;
  }
  std::printf("sanitize_check: ok\n");
  return 0;
}
