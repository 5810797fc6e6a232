// Contraction hierarchy builder, upward search and host sweep (see ch.h).
//
// Node order: priority 2 x (shortcuts - live edges) + contracted
// neighbours + level, witness searches bounded by a settle limit (a missed
// witness only adds a shortcut that is a real path, so the distances stay
// exact).  Shortcuts of kLbCutoff minutes or more are left
// out: no lower bound below the cut-off can use them.
#include "ch.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <functional>
#include <queue>
#include <stdexcept>
#include <thread>

namespace nigiri_amd {

namespace {

struct arc {
  std::uint32_t node;
  std::uint32_t w;
};

// Per-thread witness search state.
struct witness_state {
  std::vector<std::uint32_t> dist, stamp, target;
  std::uint32_t cur = 0U, n_targets = 0U;
  std::vector<std::pair<std::uint32_t, std::uint32_t>> heap;
  std::vector<arc> ins, outs;  // live neighbours of the node being contracted
  explicit witness_state(std::uint32_t n) : dist(n, 0U), stamp(n, 0U), target(n, 0U) {}
  std::uint32_t dist_of(std::uint32_t v) const { return stamp[v] == cur ? dist[v] : ~0U; }
};

// Contraction in parallel rounds: every round computes the priority of the
// nodes whose neighbourhood changed, takes the nodes whose priority is a
// strict local minimum (an independent set) and contracts them together.
// Witness searches treat every node of the set as already removed, so each
// contraction is valid whatever rank order the set gets; shortcuts the
// others would have provided as witnesses only make more shortcuts.
struct builder {
  std::uint32_t n;
  std::vector<std::vector<arc>> out, in;
  std::vector<std::uint8_t> removed;  // contracted or in the current set
  std::vector<std::uint32_t> deleted_nb, level;
  std::uint64_t n_shortcuts = 0U;

  explicit builder(std::uint32_t nodes)
      : n{nodes}, out(nodes), in(nodes), removed(nodes, 0U), deleted_nb(nodes, 0U), level(nodes, 0U) {}

  // Dijkstra from s over nodes not removed, other than `via`, until every
  // target (nodes with target[x] == cur, set by the caller after begin())
  // is settled, or up to distance max_d or `limit` settled nodes
  static void begin(witness_state& w) {
    if (++w.cur == 0U) {
      std::fill(w.stamp.begin(), w.stamp.end(), 0U);
      std::fill(w.target.begin(), w.target.end(), 0U);
      w.cur = 1U;
    }
    w.n_targets = 0U;
  }
  void witness(witness_state& w, std::uint32_t s, std::uint32_t via, std::uint32_t max_d,
               std::uint32_t limit) const {
    w.heap.clear();
    w.stamp[s] = w.cur;
    w.dist[s] = 0U;
    w.heap.emplace_back(0U, s);
    std::uint32_t settled = 0U;
    auto const cmp = std::greater<>{};
    while (!w.heap.empty()) {
      std::pop_heap(w.heap.begin(), w.heap.end(), cmp);
      auto const [d, u] = w.heap.back();
      w.heap.pop_back();
      if (d != w.dist_of(u)) continue;
      if (d > max_d || ++settled > limit) break;
      if (w.target[u] == w.cur && --w.n_targets == 0U) break;
      for (auto const& a : out[u]) {
        if (a.node == via || removed[a.node] != 0U) continue;
        auto const nd = d + a.w;
        if (nd < w.dist_of(a.node)) {
          w.stamp[a.node] = w.cur;
          w.dist[a.node] = nd;
          w.heap.emplace_back(nd, a.node);
          std::push_heap(w.heap.begin(), w.heap.end(), cmp);
        }
      }
    }
  }

  // removes the arcs to removed nodes from the list of live node x
  void prune(std::vector<std::vector<arc>>& adj, std::uint32_t x) {
    if (removed[x] != 0U) return;
    auto& v = adj[x];
    v.erase(std::remove_if(v.begin(), v.end(), [&](arc const& a) { return removed[a.node] != 0U; }), v.end());
  }

  void add_edge(std::uint32_t u, std::uint32_t x, std::uint32_t c) {
    for (auto& a : out[u]) {
      if (a.node == x) {
        if (c < a.w) {
          a.w = c;
          for (auto& b : in[x]) {
            if (b.node == u) b.w = c;
          }
        }
        return;
      }
    }
    out[u].push_back({x, c});
    in[x].push_back({u, c});
    ++n_shortcuts;
  }

  // shortcuts needed to contract v (appended to `sc` as (u, x, c) when
  // given), and the number of live edges at v
  std::pair<int, int> contract(witness_state& w, std::uint32_t v, std::uint32_t limit,
                               std::vector<std::array<std::uint32_t, 3>>* sc) const {
    int shortcuts = 0, live = 0;
    std::uint32_t max_out = 0U;
    auto& ins = w.ins;
    auto& outs = w.outs;
    ins.clear();
    outs.clear();
    for (auto const& o : out[v]) {
      if (removed[o.node] == 0U) {
        outs.push_back(o);
        max_out = std::max(max_out, o.w);
      }
    }
    for (auto const& i : in[v]) {
      if (removed[i.node] == 0U) ins.push_back(i);
    }
    live = static_cast<int>(ins.size() + outs.size());
    for (auto const& i : ins) {
      if (std::none_of(outs.begin(), outs.end(),
                       [&](arc const& o) { return o.node != i.node && i.w + o.w < kLbCutoff; })) {
        continue;
      }
      begin(w);
      for (auto const& o : outs) {
        if (o.node != i.node && w.target[o.node] != w.cur) {
          w.target[o.node] = w.cur;
          ++w.n_targets;
        }
      }
      witness(w, i.node, v, std::min<std::uint32_t>(i.w + max_out, kLbCutoff), limit);
      for (auto const& o : outs) {
        if (o.node == i.node) continue;
        auto const c = i.w + o.w;
        if (c >= kLbCutoff || w.dist_of(o.node) <= c) continue;
        ++shortcuts;
        if (sc != nullptr) sc->push_back({i.node, o.node, c});
      }
    }
    return {shortcuts, live};
  }

  std::int64_t priority(witness_state& w, std::uint32_t v) const {
    auto const [sc, live] = contract(w, v, 40U, nullptr);
    return 2 * (static_cast<std::int64_t>(sc) - live) + deleted_nb[v] + level[v];
  }
};

// f(i, thread) for i in [0, n) on up to `threads` threads
template <typename F>
void parallel_for(std::size_t n, unsigned threads, F&& f) {
  if (threads <= 1U || n < 256U) {
    for (std::size_t i = 0; i != n; ++i) f(i, 0U);
    return;
  }
  std::atomic<std::size_t> next{0U};
  std::vector<std::thread> pool;
  for (unsigned t = 0; t != threads; ++t) {
    pool.emplace_back([&, t]() {
      for (auto i = next.fetch_add(64U); i < n; i = next.fetch_add(64U)) {
        for (auto j = i; j != std::min(n, i + 64U); ++j) f(j, t);
      }
    });
  }
  for (auto& th : pool) th.join();
}

void make_csr(std::uint32_t n, std::vector<std::vector<arc>> const& adj,
              std::function<bool(std::uint32_t, std::uint32_t)> const& keep, ch_csr& out) {
  out.off.assign(n + 1U, 0U);
  out.node.clear();
  out.w.clear();
  for (std::uint32_t v = 0; v != n; ++v) {
    for (auto const& a : adj[v]) {
      if (!keep(v, a.node)) continue;
      out.node.push_back(a.node);
      out.w.push_back(static_cast<std::uint16_t>(std::min(a.w, kLbCutoff)));
    }
    out.off[v + 1U] = static_cast<std::uint32_t>(out.node.size());
  }
}

// levels of a sweep over `edges` (per v: its higher-ranked sources)
void make_sweep(std::uint32_t n, std::vector<std::uint32_t> const& by_rank_desc, ch_csr const& edges,
                ch_sweep& s) {
  std::vector<std::uint32_t> lvl(n, 0U);
  std::uint32_t max_lvl = 0U;
  for (auto const v : by_rank_desc) {
    std::uint32_t l = 0U;
    for (auto e = edges.off[v]; e != edges.off[v + 1U]; ++e) l = std::max(l, lvl[edges.node[e]] + 1U);
    lvl[v] = l;
    max_lvl = std::max(max_lvl, l);
  }
  // nodes without sweep edges (level 0) need no work
  std::vector<std::uint32_t> count(max_lvl + 2U, 0U);
  for (std::uint32_t v = 0; v != n; ++v) {
    if (lvl[v] != 0U) ++count[lvl[v]];
  }
  s.level_off.assign(1U, 0U);
  for (std::uint32_t l = 1U; l <= max_lvl; ++l) s.level_off.push_back(s.level_off.back() + count[l]);
  std::vector<std::uint32_t> pos(s.level_off.begin(), s.level_off.end());
  s.order.assign(s.level_off.back(), 0U);
  for (std::uint32_t v = 0; v != n; ++v) {
    if (lvl[v] != 0U) s.order[pos[lvl[v] - 1U]++] = v;
  }
  s.edge_off.assign(1U, 0U);
  s.src.clear();
  s.w.clear();
  for (auto const v : s.order) {
    for (auto e = edges.off[v]; e != edges.off[v + 1U]; ++e) {
      s.src.push_back(edges.node[e]);
      s.w.push_back(edges.w[e]);
    }
    s.edge_off.push_back(static_cast<std::uint32_t>(s.src.size()));
  }
  // host levels: the leading levels of fewer than kHostLevelNodes locations,
  // while their edges stay within 1/32 of all sweep edges (each is one
  // dependent device launch for a handful of locations)
  constexpr std::uint32_t kHostLevelNodes = 128U;
  auto const edge_budget = s.src.size() / 32U;
  s.host_levels = 0U;
  while (s.host_levels + 1U < s.level_off.size() &&
         s.level_off[s.host_levels + 1U] - s.level_off[s.host_levels] < kHostLevelNodes &&
         s.edge_off[s.level_off[s.host_levels + 1U]] <= edge_budget) {
    ++s.host_levels;
  }
  s.on_host.assign(n, 0U);
  for (std::uint32_t p = 0U; p != s.level_off[s.host_levels]; ++p) s.on_host[s.order[p]] = 1U;
}

}  // namespace

contraction_hierarchy build_ch(std::uint32_t n, std::vector<std::uint32_t> const& off,
                               std::vector<std::uint32_t> const& edges) {
  builder b{n};
  for (std::uint32_t u = 0; u != n; ++u) {
    for (auto e = off[u]; e != off[u + 1U]; ++e) {
      auto const t = edges[e] & ((1U << 23) - 1U), w = edges[e] >> 23;
      if (t == u || t >= n) continue;
      b.out[u].push_back({t, w});
      b.in[t].push_back({u, w});
    }
  }
  auto const threads = std::max(1U, std::min(16U, std::thread::hardware_concurrency()));
  std::vector<witness_state> ws;
  for (unsigned t = 0; t != threads; ++t) ws.emplace_back(n);
  contraction_hierarchy ch;
  ch.n = n;
  ch.rank.assign(n, 0U);
  std::vector<std::int64_t> prio(n, 0);
  std::vector<std::uint8_t> dirty(n, 1U);
  std::vector<std::uint32_t> remaining(n);
  for (std::uint32_t v = 0; v != n; ++v) remaining[v] = v;
  // tie break between equal priorities: a fixed pseudo-random order
  auto const tie = [](std::uint32_t v) { return (v * 2654435761U) ^ (v >> 16); };
  std::uint32_t next_rank = 0U;
  std::vector<std::vector<std::array<std::uint32_t, 3>>> sc(threads);
  while (!remaining.empty()) {
    parallel_for(remaining.size(), threads, [&](std::size_t i, unsigned t) {
      auto const v = remaining[i];
      if (dirty[v] != 0U) prio[v] = b.priority(ws[t], v);
    });
    // strict local minima of (priority, tie) among live neighbours
    std::vector<std::uint8_t> pick(remaining.size(), 0U);
    parallel_for(remaining.size(), threads, [&](std::size_t i, unsigned) {
      auto const v = remaining[i];
      auto const key = std::make_pair(prio[v], tie(v));
      auto const lower = [&](std::vector<arc> const& adj) {
        return std::none_of(adj.begin(), adj.end(), [&](arc const& a) {
          return b.removed[a.node] == 0U && std::make_pair(prio[a.node], tie(a.node)) <= key && a.node != v;
        });
      };
      pick[i] = lower(b.out[v]) && lower(b.in[v]) ? 1U : 0U;
    });
    std::vector<std::uint32_t> set, rest;
    for (std::size_t i = 0; i != remaining.size(); ++i) (pick[i] != 0U ? set : rest).push_back(remaining[i]);
    for (auto const v : set) b.removed[v] = 1U;
    for (auto& x : sc) x.clear();
    parallel_for(set.size(), threads, [&](std::size_t i, unsigned t) {
      b.contract(ws[t], set[i], 500U, &sc[t]);
    });
    // drop the arcs to the set's nodes: they stay in the set's own lists,
    // which form the overlay graph, and leave the live graph's lists short
    for (auto const v : set) {
      for (auto const& a : b.out[v]) b.prune(b.in, a.node);
      for (auto const& a : b.in[v]) b.prune(b.out, a.node);
    }
    for (auto const& x : sc) {
      for (auto const& [u, y, c] : x) b.add_edge(u, y, c);
    }
    std::fill(dirty.begin(), dirty.end(), 0U);
    for (auto const v : set) {
      ch.rank[v] = next_rank++;
      auto const touch = [&](std::uint32_t x) {
        if (b.removed[x] != 0U) return;
        ++b.deleted_nb[x];
        b.level[x] = std::max(b.level[x], b.level[v] + 1U);
        dirty[x] = 1U;
      };
      for (auto const& a : b.out[v]) touch(a.node);
      for (auto const& a : b.in[v]) touch(a.node);
    }
    // new shortcuts change their end points' neighbourhoods
    for (auto const& x : sc) {
      for (auto const& [u, y, c] : x) dirty[u] = dirty[y] = 1U;
    }
    remaining = std::move(rest);
  }
  ch.n_shortcuts = b.n_shortcuts;
  auto const& rank = ch.rank;
  make_csr(n, b.out, [&](std::uint32_t v, std::uint32_t u) { return rank[u] > rank[v]; }, ch.up_out);
  make_csr(n, b.in, [&](std::uint32_t v, std::uint32_t u) { return rank[u] > rank[v]; }, ch.down_in);
  std::vector<std::uint32_t> by_rank_desc(n);
  for (std::uint32_t v = 0; v != n; ++v) by_rank_desc[n - 1U - rank[v]] = v;
  make_sweep(n, by_rank_desc, ch.down_in, ch.sweep_fwd);
  make_sweep(n, by_rank_desc, ch.up_out, ch.sweep_bwd);
  return ch;
}

void ch_upward(contraction_hierarchy const& ch, bool fwd,
               std::vector<std::pair<std::uint32_t, std::uint16_t>> const& seeds, ch_search_scratch& s,
               std::vector<std::pair<std::uint32_t, std::uint16_t>>& out) {
  auto const& g = fwd ? ch.up_out : ch.down_in;
  if (s.dist.size() != ch.n) {
    s.dist.assign(ch.n, 0U);
    s.stamp.assign(ch.n, 0U);
    s.cur = 0U;
  }
  if (++s.cur == 0U) {
    std::fill(s.stamp.begin(), s.stamp.end(), 0U);
    s.cur = 1U;
  }
  auto const dist_of = [&](std::uint32_t v) { return s.stamp[v] == s.cur ? s.dist[v] : ~0U; };
  s.heap.clear();
  auto const cmp = std::greater<>{};
  for (auto const& [l, d] : seeds) {
    if (l >= ch.n || d >= kLbCutoff || d >= dist_of(l)) continue;
    s.stamp[l] = s.cur;
    s.dist[l] = d;
    s.heap.emplace_back(d, l);
    std::push_heap(s.heap.begin(), s.heap.end(), cmp);
  }
  out.clear();
  while (!s.heap.empty()) {
    std::pop_heap(s.heap.begin(), s.heap.end(), cmp);
    auto const [d, u] = s.heap.back();
    s.heap.pop_back();
    if (d != dist_of(u)) continue;
    out.emplace_back(u, static_cast<std::uint16_t>(d));
    for (auto e = g.off[u]; e != g.off[u + 1U]; ++e) {
      auto const nd = d + g.w[e];
      auto const t = g.node[e];
      if (nd < kLbCutoff && nd < dist_of(t)) {
        s.stamp[t] = s.cur;
        s.dist[t] = nd;
        s.heap.emplace_back(nd, t);
        std::push_heap(s.heap.begin(), s.heap.end(), cmp);
      }
    }
  }
}

void ch_sweep_seeds(contraction_hierarchy const& ch, bool fwd,
                    std::vector<std::pair<std::uint32_t, std::uint16_t>> const& seeds, ch_search_scratch& s,
                    std::vector<std::pair<std::uint32_t, std::uint16_t>>& out) {
  thread_local std::vector<std::pair<std::uint32_t, std::uint16_t>> up;
  ch_upward(ch, fwd, seeds, s, up);
  auto const& sw = fwd ? ch.sweep_fwd : ch.sweep_bwd;
  out.clear();
  for (auto const& l : up) {
    if (sw.on_host[l.first] == 0U) out.push_back(l);
  }
  // the host levels' sweep over the upward labels (s.dist where stamped):
  // every edge comes from a location of an earlier host level or of no
  // sweep level at all, so their values are final when read
  auto const value = [&](std::uint32_t v) -> std::uint32_t { return s.stamp[v] == s.cur ? s.dist[v] : 0xFFFFU; };
  for (std::uint32_t p = 0U; p != sw.level_off[sw.host_levels]; ++p) {
    auto const v = sw.order[p];
    auto d = value(v);
    for (auto e = sw.edge_off[p]; e != sw.edge_off[p + 1U]; ++e) d = std::min(d, value(sw.src[e]) + sw.w[e]);
    if (d < kLbCutoff) {  // (an upward label is below kLbCutoff: d is then too)
      s.stamp[v] = s.cur;
      s.dist[v] = d;
      out.emplace_back(v, static_cast<std::uint16_t>(d));
    }
  }
}

void ch_lower_bounds(contraction_hierarchy const& ch, bool fwd,
                     std::vector<std::pair<std::uint32_t, std::uint16_t>> const& seeds, ch_search_scratch& s,
                     std::vector<std::uint16_t>& lb) {
  std::vector<std::pair<std::uint32_t, std::uint16_t>> start;
  ch_sweep_seeds(ch, fwd, seeds, s, start);
  lb.assign(ch.n, 0xFFFFU);
  for (auto const& [v, d] : start) lb[v] = d;
  auto const& sw = fwd ? ch.sweep_fwd : ch.sweep_bwd;
  for (std::size_t i = sw.level_off[sw.host_levels]; i != sw.order.size(); ++i) {
    auto const v = sw.order[i];
    std::uint32_t d = lb[v];
    for (auto e = sw.edge_off[i]; e != sw.edge_off[i + 1U]; ++e) {
      d = std::min<std::uint32_t>(d, static_cast<std::uint32_t>(lb[sw.src[e]]) + sw.w[e]);
    }
    lb[v] = static_cast<std::uint16_t>(d < kLbCutoff ? d : 0xFFFFU);
  }
}

}  // namespace nigiri_amd
