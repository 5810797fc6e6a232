// Contraction hierarchy of the lower-bound graph, for exact one-to-all
// shortest travel times by PHAST: a small upward search from the query's
// destinations on the host, then one downward sweep over every location on
// the device (64 queries share each node's 128 B row).  Replaces the
// per-query dial Dijkstra of routing/dijkstra.cc:18-67: the lower bounds are
// shortest-path distances, unique whatever algorithm computes them, cut off
// at 7201 minutes (kMaxTravelTime + 1 dial buckets, dijkstra.h:52).
//
// The hierarchy is built on G = fwd_search_lb_graph_ (edges u -> t of
// lb_fwd, the graph the forward search's lower bounds run on).  The
// backward graph is its transpose, so the same node order and shortcuts
// serve both directions:
//   G   : upward search over up_out, sweep over down_in;
//   G^T : upward search over down_in, sweep over up_out.
#pragma once

#include <cstdint>
#include <utility>
#include <vector>

namespace nigiri_amd {

constexpr std::uint32_t kLbCutoff = 7201U;  // dial buckets (dijkstra.h:52)

struct ch_csr {
  std::vector<std::uint32_t> off;  // [L + 1]
  std::vector<std::uint32_t> node;
  std::vector<std::uint16_t> w;    // minutes, clamped at kLbCutoff
};

// Sweep order of one direction: the locations with edges from higher
// ranks, grouped into levels (a location comes after every higher-ranked
// neighbour its sweep edges read), and those edges in that order.
// The leading `host_levels` levels (the top of the hierarchy: a long chain
// of small levels) are swept on the host right after the upward search,
// so that the device sweep starts at level host_levels.
struct ch_sweep {
  std::vector<std::uint32_t> level_off;  // [n_levels + 1] into order
  std::vector<std::uint32_t> order;      // locations
  std::vector<std::uint32_t> edge_off;   // [order.size() + 1]
  std::vector<std::uint32_t> src;        // higher-ranked neighbour
  std::vector<std::uint16_t> w;
  std::uint32_t host_levels = 0U;
  std::vector<std::uint8_t> on_host;     // [n]: 1 for the locations of those levels
};

struct contraction_hierarchy {
  std::uint32_t n = 0U;
  std::vector<std::uint32_t> rank;  // contraction position of each location
  ch_csr up_out;   // v -> u of the overlay graph with rank[u] > rank[v]
  ch_csr down_in;  // u -> v of the overlay graph with rank[u] > rank[v], by v
  ch_sweep sweep_fwd, sweep_bwd;
  std::uint64_t n_shortcuts = 0U;
  bool empty() const { return n == 0U; }
};

// Builds the hierarchy of the graph given as CSR (edges u -> target,
// footpath-packed: bits 0..22 target, 23..31 minutes).
contraction_hierarchy build_ch(std::uint32_t n, std::vector<std::uint32_t> const& off,
                               std::vector<std::uint32_t> const& edges);

// Reused by one thread across upward searches.
struct ch_search_scratch {
  std::vector<std::uint32_t> dist, stamp;
  std::uint32_t cur = 0U;
  std::vector<std::pair<std::uint32_t, std::uint32_t>> heap;  // (dist, node)
};

// Upward search of PHAST from the seeds (location, minutes) for the
// forward (fwd = true: graph G) or backward lower bounds: every location it
// settles with its tentative distance (< kLbCutoff), in settle order.
void ch_upward(contraction_hierarchy const& ch, bool fwd,
               std::vector<std::pair<std::uint32_t, std::uint16_t>> const& seeds,
               ch_search_scratch& s,
               std::vector<std::pair<std::uint32_t, std::uint16_t>>& out);

// The labels a sweep from level host_levels starts from: the upward
// search's labels of the locations below the host levels, and the final
// bounds (< kLbCutoff) of the host levels' locations, each location once.
// Every location without a label starts unreachable.
void ch_sweep_seeds(contraction_hierarchy const& ch, bool fwd,
                    std::vector<std::pair<std::uint32_t, std::uint16_t>> const& seeds,
                    ch_search_scratch& s,
                    std::vector<std::pair<std::uint32_t, std::uint16_t>>& out);

// Complete lower bounds on the host (sweep seeds, then the sweep from level
// host_levels, as the device runs it): the reference for the device sweep,
// and the plain host path.
void ch_lower_bounds(contraction_hierarchy const& ch, bool fwd,
                     std::vector<std::pair<std::uint32_t, std::uint16_t>> const& seeds,
                     ch_search_scratch& s, std::vector<std::uint16_t>& lb);

}  // namespace nigiri_amd
