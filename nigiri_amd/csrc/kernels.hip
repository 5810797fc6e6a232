// gfx950 round kernels (see kernels.h for the round structure).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "kernels.h"

namespace nigiri_amd {

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / 64;

template <bool Fwd>
struct dirop {
  static constexpr std::int16_t kInvalid = Fwd ? 32767 : -32768;
  __device__ static bool better(int a, int b) { return Fwd ? a < b : a > b; }
  __device__ static bool better_eq(int a, int b) { return Fwd ? a <= b : a >= b; }
  __device__ static std::int16_t best_of(std::int16_t a, std::int16_t b) {
    return better(a, b) ? a : b;
  }
  __device__ static int dir(int a) { return Fwd ? a : -a; }
};

// A value that is the same in every active lane, moved to a scalar register
// (lets the compiler use scalar loads and scalar branches on it).
__device__ inline std::uint32_t uni(std::uint32_t v) {
  return static_cast<std::uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(v)));
}

__device__ inline std::int16_t clamp16(int v) {
  return static_cast<std::int16_t>(v < -32768 ? -32768 : (v > 32767 ? 32767 : v));
}

// Atomic "keep the better value" on an int16 element (32-bit CAS on the
// aligned word).  Returns the previous value.
template <bool Fwd>
__device__ inline std::int16_t atomic_best16(std::int16_t* p, std::int16_t v) {
  auto const addr = reinterpret_cast<std::uintptr_t>(p);
  auto* w = reinterpret_cast<unsigned int*>(addr & ~std::uintptr_t{3});
  unsigned const shift = (addr & 2U) != 0U ? 16U : 0U;
  unsigned old = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (true) {
    auto const cur = static_cast<std::int16_t>(static_cast<std::uint16_t>(old >> shift));
    if (!dirop<Fwd>::better(v, cur)) return cur;
    unsigned const nw = (old & ~(0xFFFFU << shift)) |
                        (static_cast<unsigned>(static_cast<std::uint16_t>(v)) << shift);
    unsigned const prev = atomicCAS(w, old, nw);
    if (prev == old) return cur;
    old = prev;
  }
}

__device__ inline bool bit_test(std::uint64_t const* bf, std::uint32_t idx, std::uint32_t day) {
  if (day >= 512U) return false;  // out-of-range day index: not active
  return ((bf[static_cast<std::size_t>(idx) * 8U + day / 64U] >> (day % 64U)) & 1ULL) != 0ULL;
}

__device__ inline std::uint32_t clamp_count(std::uint32_t n, std::size_t cap) {
  return n < cap ? n : static_cast<std::uint32_t>(cap);
}

__device__ inline void list_push(std::uint64_t* list, std::uint32_t* count, std::size_t cap,
                                 std::uint64_t v, std::uint32_t* overflow) {
  auto const i = atomicAdd(count, 1U);
  if (i < cap) {
    list[i] = v;
  } else {
    atomicOr(overflow, 1U);
  }
}

// Work items of a wavefront from a shared cursor, `chunk` at a time:
// wavefronts that start late (their CU was busy) simply take fewer items.
struct item_claim {
  std::uint32_t* cursor;
  std::uint32_t chunk;
  std::uint32_t next = 0U, end = 0U;
  __device__ std::uint32_t get() {
    if (next == end) {
      std::uint32_t v = 0U;
      if ((threadIdx.x & 63U) == 0U) v = atomicAdd(cursor, chunk);
      next = static_cast<std::uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(v)));
      end = next + chunk;
    }
    return next++;
  }
};

// adds a per-lane count to a global counter with one atomic per wavefront
__device__ inline void wave_count(unsigned long long* counter, unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) {
    v += __shfl_xor(v, off, 64);
  }
  if ((threadIdx.x & 63U) == 0U && v != 0ULL) {
    atomicAdd(counter, v);
  }
}

// ---- state indexing ---------------------------------------------------
template <int W>
struct lay {
  __device__ static std::size_t col(dev_state const& s, std::uint32_t g, std::uint32_t l,
                                    std::uint32_t lane) {
    return (static_cast<std::size_t>(g) * s.L + l) * W + lane;
  }
  __device__ static std::size_t rt(dev_state const& s, std::uint32_t g, std::uint32_t k,
                                   std::uint32_t l, std::uint32_t lane) {
    return ((static_cast<std::size_t>(g) * s.K + k) * s.L + l) * W + lane;
  }
  __device__ static std::size_t lword(dev_state const& s, std::uint32_t g, std::uint32_t l) {
    return W == 64 ? static_cast<std::size_t>(g) * s.L + l
                   : static_cast<std::size_t>(g) * s.LW + l / 64U;
  }
  __device__ static std::uint64_t lbit(std::uint32_t l, std::uint32_t lane) {
    return W == 64 ? (1ULL << lane) : (1ULL << (l % 64U));
  }
  __device__ static std::size_t rword(dev_state const& s, std::uint32_t g, std::uint32_t r) {
    return W == 64 ? static_cast<std::size_t>(g) * s.R + r
                   : static_cast<std::size_t>(g) * s.RW + r / 64U;
  }
  __device__ static std::uint64_t rbit(std::uint32_t r, std::uint32_t lane) {
    return W == 64 ? (1ULL << lane) : (1ULL << (r % 64U));
  }
  // lane mask (W == 64) or bit 0 (W == 1) of the marked lanes at location l
  __device__ static std::uint64_t lanes_at(std::uint64_t word, std::uint32_t l) {
    return W == 64 ? word : ((word >> (l % 64U)) & 1ULL);
  }
  __device__ static bool test(std::uint64_t const* m, dev_state const& s, std::uint32_t g,
                              std::uint32_t l, std::uint32_t lane) {
    return (m[lword(s, g, l)] & lbit(l, lane)) != 0ULL;
  }
};

// Batched engine: round_times[k][l] / best[l] of (g, lane) may now hold a
// label (the validity bitmaps that let resets and the best sweep skip the
// rest); no-ops for the plain engine, whose resets are dense.
template <int W>
__device__ inline void note_rt(dev_state const& s, std::uint32_t g, std::uint32_t k, std::uint32_t l,
                               std::uint32_t lane) {
  if constexpr (W == 64) {
    atomicOr(reinterpret_cast<unsigned long long*>(&s.rt_valid[(static_cast<std::size_t>(g) * s.K + k) * s.L + l]),
             1ULL << lane);
  }
}
template <int W>
__device__ inline void note_bt(dev_state const& s, std::uint32_t g, std::uint32_t l, std::uint32_t lane) {
  if constexpr (W == 64) {
    atomicOr(reinterpret_cast<unsigned long long*>(&s.bt_valid[static_cast<std::size_t>(g) * s.L + l]), 1ULL << lane);
  }
}

// Sets the mark of (g, l, lane); appends (g, l) to the list on the first
// mark of the location word.
template <int W>
__device__ inline void mark_one(std::uint64_t* marks, std::uint64_t* list, std::uint32_t* count,
                                std::size_t cap, std::uint32_t* overflow, dev_state const& s,
                                std::uint32_t g, std::uint32_t l, std::uint32_t lane) {
  auto const bit = lay<W>::lbit(l, lane);
  auto const old = atomicOr(reinterpret_cast<unsigned long long*>(&marks[lay<W>::lword(s, g, l)]),
                            static_cast<unsigned long long>(bit));
  bool const first = W == 64 ? old == 0ULL : (old & bit) == 0ULL;
  if (first) {
    list_push(list, count, cap, (static_cast<std::uint64_t>(g) << 32) | l, overflow);
  }
}

// Marks of the batched engine: the lane mask word and a one-bit-per-location
// summary word, both set with non-returning atomics; a compaction kernel
// turns the summary into the next phase's work list.
__device__ inline void mark_wave64(std::uint64_t* marks, std::uint8_t* flags, std::uint32_t g,
                                   std::uint32_t l, std::uint64_t lanes, std::uint32_t n_per_group) {
  auto const i = static_cast<std::size_t>(g) * n_per_group + l;
  atomicOr(reinterpret_cast<unsigned long long*>(&marks[i]), static_cast<unsigned long long>(lanes));
  flags[i] = 1U;
}

// Route-phase mark of (g, l, lane): a byte per lane for the batched engine
// (set with plain stores by the route scan), a bitset for the plain engine.
template <int W>
__device__ inline bool marked_r(dev_state const& s, std::uint32_t g, std::uint32_t l, std::uint32_t lane) {
  if constexpr (W == 64) {
    return s.mark_rb[(static_cast<std::size_t>(g) * s.L + l) * 64U + lane] != 0U;
  } else {
    return lay<W>::test(s.mark_r, s, g, l, lane);
  }
}

// ---- transfer-time settings (transfer_time_settings.h:21-31) ------------
template <typename T>
__device__ inline T adjusted(slot_param const& p, T const d) {
  if (p.tts_default != 0U) return d;
  auto const a = static_cast<T>(p.tts_add);
  auto const m = static_cast<T>(p.tts_min);
  auto const f = static_cast<T>(static_cast<float>(d) * p.tts_factor);
  return static_cast<T>(a + (m > f ? m : f));
}

template <bool Fwd>
__device__ inline void update_tad(std::int16_t* tad, std::int32_t k, std::int16_t t) {
  for (int i = k; i < kRowsMax; ++i) {
    tad[i] = dirop<Fwd>::best_of(tad[i], t);
  }
}

// time_at_dest in force at sequence position `pos` of the transfer phase
__device__ inline std::int16_t tad_at(dev_state const& s, std::uint32_t slot, std::uint32_t pos) {
  auto v = s.tad0[slot];
  auto const n = s.n_chg[slot];
  for (std::uint32_t j = 0; j < n; ++j) {
    if (s.chg_pos[slot * kMaxChanges + j] < pos) {
      v = s.chg_tad[slot * kMaxChanges + j];
    } else {
      break;
    }
  }
  return v;
}

// within_bounds (raptor.h:458-489, Vias = 0): with bounds set (a pong
// search), a new label must not be better than the bound of the other
// direction's search for the remaining rounds plus the transfer time.
// Plain engine only; the batched engine runs no bounded searches.
template <bool Fwd, int W>
__device__ inline bool within_bounds(dev_image const& im, dev_state const& s, std::uint32_t slot,
                                     std::uint32_t k, std::uint32_t l, int t) {
  if constexpr (W != 1) {
    return true;
  } else {
    auto const& p = s.param[slot];
    if (p.bounds_last_k == 0U) return true;
    int const b = s.bounds[lay<1>::rt(s, slot, p.bounds_last_k - k, l, 0U)];
    return dirop<Fwd>::better_eq(
        t, b + dirop<Fwd>::dir(adjusted<int>(p, static_cast<int>(im.transfer_time[l]))));
  }
}

// ======================================================================
// round_end_check: slots whose rounds are exhausted (k == end_k) stop
template <int W>
__global__ void k_round_end_check(round_args a, std::uint32_t n_slots) {
  auto const& s = a.st;
  auto const slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= n_slots) return;
  auto const g = slot / W, lane = slot % W;
  auto const lbit = 1ULL << lane;
  if ((s.grp_active[g] & lbit) == 0ULL) return;
  auto const k = static_cast<std::uint32_t>(s.kslot[slot]) + 1U;
  s.kslot[slot] = static_cast<std::uint8_t>(k);
  if (k >= s.param[slot].end_k) {  // (a slot is retired after its last round)
    atomicAnd(reinterpret_cast<unsigned long long*>(&s.grp_active[g]), ~lbit);
  }
}

// round_finish: a slot whose next round would reach end_k ends its epoch
// now (raptor.h:218: rounds k = 1 .. end_k - 1)
template <int W>
__global__ void k_round_finish(round_args a, std::uint32_t n_slots) {
  auto const& s = a.st;
  auto const slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= n_slots) return;
  auto const g = slot / W, lane = slot % W;
  auto const lbit = 1ULL << lane;
  if ((s.grp_active[g] & lbit) == 0ULL) return;
  if (static_cast<std::uint32_t>(s.kslot[slot]) + 1U >= s.param[slot].end_k) {
    atomicAnd(reinterpret_cast<unsigned long long*>(&s.grp_active[g]), ~lbit);
  }
}

// round_begin: time_at_dest from the destination labels (raptor.h:224-226)
template <bool Fwd, int W>
__global__ void k_round_begin(round_args a, std::uint32_t n_slots) {
  auto const& s = a.st;
  auto const slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= n_slots) return;
  auto const g = slot / W, lane = slot % W;
  if ((s.grp_active[g] & (1ULL << lane)) == 0ULL) return;
  auto const& p = s.param[slot];
  auto* tad = s.tad + static_cast<std::size_t>(slot) * kRowsMax;
  auto const k = static_cast<std::int32_t>(s.kslot[slot]);
  if (p.intermodal != 0U) {
    update_tad<Fwd>(tad, k, s.best[lay<W>::col(s, g, 1U, lane)]);
  } else {
    for (std::uint32_t j = 0; j < p.n_dest; ++j) {
      auto const d = s.dest[slot * kMaxDestBits + j];
      update_tad<Fwd>(tad, k, s.best[lay<W>::col(s, g, d, lane)]);
    }
  }
}

// best = min(round_times[k], best) for groups holding later-epoch slots
// (raptor.h:219-223); sweep[0] is the number of groups, sweep[1..] their
// indices (in device memory, so that a captured round replays with the
// current list).  Batched engine: one thread per (group, station); the
// group's active lanes are split by their round k, and for each such round
// the thread reads the station's rt_valid word, so that only the labels of
// earlier epochs are read (the word loads are coalesced, the label accesses
// sparse); the lanes that lower best are added to bt_valid.
template <bool Fwd, int W>
__global__ void k_best_sweep(round_args a, std::uint32_t const* sweep) {
  auto const& s = a.st;
  auto const n_groups = sweep[0];
  if constexpr (W == 64) {
    auto const tid = blockIdx.x * blockDim.x + threadIdx.x;
    auto const stride = gridDim.x * blockDim.x;
    for (std::uint32_t gi = 0U; gi < n_groups; ++gi) {  // uniform
      auto const g = sweep[1U + gi];
      // the group's active lanes per round k: lane i of the wavefront looks
      // at slot lane i (the same masks in every wavefront)
      auto const me = threadIdx.x & 63U;
      bool const on = ((s.grp_active[g] >> me) & 1ULL) != 0ULL;
      auto const my_k = static_cast<std::uint32_t>(s.kslot[g * 64U + me]);
      std::uint64_t at_k[kRowsMax];
#pragma unroll
      for (std::uint32_t k = 0U; k < kRowsMax; ++k) at_k[k] = __ballot(on && my_k == k);
      for (std::uint32_t l = tid; l < s.L; l += stride) {
        std::uint64_t wrote = 0ULL;
        // the station's rt_valid words of the rounds in use, in flight together
        std::uint64_t vw[kRowsMax];
#pragma unroll
        for (std::uint32_t k = 1U; k < kRowsMax; ++k) {
          vw[k] = (k < s.K && at_k[k] != 0ULL)  // uniform
                      ? s.rt_valid[(static_cast<std::size_t>(g) * s.K + k) * s.L + l] & at_k[k]
                      : 0ULL;
        }
#pragma unroll
        for (std::uint32_t k = 1U; k < kRowsMax; ++k) {
          auto x = vw[k];
          while (x != 0ULL) {
            auto const lane = static_cast<std::uint32_t>(__ffsll(static_cast<long long>(x)) - 1);
            x &= x - 1ULL;
            auto const v = s.round_times[lay<64>::rt(s, g, k, l, lane)];
            auto const c = lay<64>::col(s, g, l, lane);
            if (dirop<Fwd>::better(v, s.best[c])) {
              s.best[c] = v;
              wrote |= 1ULL << lane;
            }
          }
        }
        if (wrote != 0ULL) s.bt_valid[static_cast<std::size_t>(g) * s.L + l] |= wrote;
      }
    }
    return;
  }
  auto const per_group = s.L * static_cast<std::uint32_t>(W);  // < 2^32 (checked at engine creation)
  auto const stride = gridDim.x * blockDim.x;
  for (std::uint32_t gi = 0U; gi < n_groups; ++gi) {  // uniform: no per-element division
    auto const g = sweep[1U + gi];
    auto const act = s.grp_active[g];
    for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < per_group; i += stride) {
      auto const l = i / static_cast<std::uint32_t>(W), lane = i % static_cast<std::uint32_t>(W);
      if ((act & (1ULL << lane)) == 0ULL) continue;
      auto const c = lay<W>::col(s, g, l, lane);
      auto const v = s.round_times[lay<W>::rt(s, g, s.kslot[g * W + lane], l, lane)];
      if (dirop<Fwd>::better(v, s.best[c])) s.best[c] = v;
    }
  }
}

// mark_routes: station marks (mark_t) -> route marks
template <int W>
__global__ void k_mark_routes(round_args a) {
  auto const& im = a.im;
  auto const& s = a.st;
  auto* overflow = &s.counts[4];
  auto const n_items = clamp_count(s.counts[0], s.cap_t);
  if constexpr (W == 64) {
    // flag_route[rank][g] = 1 for every route of a marked station (plain
    // stores; rank: the route's place in the image's work order, so that
    // the compacted list runs through the routes in that order);
    // route_pack rebuilds the lane masks from the station marks.
    // A station serves a few routes: 8 lanes per (group, station) item,
    // 8 items per wavefront.
    constexpr std::uint32_t kSub = 8U;
    auto const lane = threadIdx.x & 63U;
    auto const sl = lane % kSub;
    std::uint32_t c_st = 0U, c_lr = 0U;
    auto const wave = (blockIdx.x * blockDim.x + threadIdx.x) / 64U;
    for (std::uint32_t base = wave * (64U / kSub); base < n_items;
         base += gridDim.x * kWavesPerBlock * (64U / kSub)) {
      auto const item = base + lane / kSub;
      if (item >= n_items) continue;
      auto const e = s.list_t[item];
      auto const g = static_cast<std::uint32_t>(e >> 32), l = static_cast<std::uint32_t>(e);
      auto const word = s.mark_t[static_cast<std::size_t>(g) * s.L + l] & s.grp_active[g];
      if (word == 0ULL) continue;
      auto const r0 = im.location_route_offsets[l], r1 = im.location_route_offsets[l + 1U];
      if (sl == 0U) {
        auto const nq = static_cast<std::uint32_t>(__popcll(word));
        c_st += nq;
        c_lr += nq * (r1 - r0);
      }
      // route-major flags: [rank][group]
      for (auto j = r0 + sl; j < r1; j += kSub) {
        s.flag_route[static_cast<std::size_t>(im.lr_rank[j]) * s.G + g] = 1U;
      }
      if (r1 != r0) {
        for (auto b = sl; b < 64U; b += kSub) {
          if (((word >> b) & 1ULL) != 0ULL) s.any_route[g * 64U + b] = 1U;
        }
      }
    }
    wave_count(&s.work[kWorkMarkedStations], c_st);
    wave_count(&s.work[kWorkLocationRoutes], c_lr);
  } else {
    for (std::uint32_t item = blockIdx.x * blockDim.x + threadIdx.x; item < n_items;
         item += gridDim.x * blockDim.x) {
      auto const e = s.list_t[item];
      auto const g = static_cast<std::uint32_t>(e >> 32), l = static_cast<std::uint32_t>(e);
      if (!lay<W>::test(s.mark_t, s, g, l, 0U) || (s.grp_active[g] & 1ULL) == 0ULL) continue;
      auto const r0 = im.location_route_offsets[l], r1 = im.location_route_offsets[l + 1U];
      atomicAdd(&s.work[kWorkMarkedStations], 1ULL);
      atomicAdd(&s.work[kWorkLocationRoutes], static_cast<unsigned long long>(r1 - r0));
      for (auto j = r0; j < r1; ++j) {
        auto const r = im.location_routes[j];
        auto const bit = lay<W>::rbit(r, 0U);
        auto const old = atomicOr(
            reinterpret_cast<unsigned long long*>(&s.route_mark[lay<W>::rword(s, g, r)]),
            static_cast<unsigned long long>(bit));
        if ((old & bit) == 0ULL) {
          list_push(s.list_route, &s.counts[2], s.cap_route,
                    (static_cast<std::uint64_t>(g) << 32) | r, overflow);
        }
      }
      if (r1 != r0) s.any_route[g] = 1U;
    }
  }
}

// ---- update_route for one query (one lane) ---------------------------------
template <bool Fwd>
struct et_t {
  std::uint32_t t_off;
  std::uint16_t day;
  __device__ bool valid() const { return day != 0xFFFFU; }
};

template <bool Fwd>
__device__ inline std::int16_t to_delta(std::uint16_t day, int mam, std::uint16_t base) {
  return clamp16((static_cast<int>(day) - static_cast<int>(base)) * 1440 + mam);
}

struct et_counters {
  unsigned long long events = 0, route_days = 0, transport_tests = 0;
};

// Traffic-day bitfields of one route scan: the route's 8 words and either
// the staged 8 words per transport (LDS) or the transports' bitfield
// indices into the global table.
struct route_bits {
  std::uint64_t const* route;   // 8 words
  std::uint64_t const* staged;  // n_t x 8 words (transport_active<true>)
  std::uint32_t const* tbf;     // bitfield index per transport (LDS or global)
  std::uint64_t const* table;   // global bitfield table
  bool use_staged;
  __device__ bool route_active(std::uint32_t day) const {
    return day < 512U && ((route[day / 64U] >> (day % 64U)) & 1ULL) != 0ULL;
  }
  // Staged: the transports' words are in LDS (route scans of small routes)
  template <bool Staged>
  __device__ bool transport_active(std::uint32_t t_off, std::uint32_t day) const {
    if (day >= 512U) return false;  // out-of-range day index: not active
    auto const w = Staged ? staged[t_off * 8U + day / 64U]
                          : table[static_cast<std::size_t>(tbf[t_off]) * 8U + day / 64U];
    return ((w >> (day % 64U)) & 1ULL) != 0ULL;
  }
};

// get_earliest_transport (raptor.h:1351-1442).  `ev` points at the event
// block of the boarding stop (one delta per transport, in LDS or global).
// The first-day seek (linear_lb, linear_lower_bound.h:7-15) is a branch-free
// binary search with `top` = the largest power of two <= n_t steps: the
// minutes of day of one stop block are non-decreasing because routes keep
// FIFO order modulo 1440 (checked when a timetable is built), which also
// makes the first day's "departure not before the arrival" test always
// pass after the seek.  The six-day loop and the event scan are one flat
// loop (one event or one day per iteration), so that lanes leave it
// independently.  The counters count what the reference's linear scan
// examines.
template <bool Fwd, bool Staged>
__device__ inline et_t<Fwd> earliest_transport(route_bits const& bits, std::uint16_t const* ev,
                                               std::uint32_t n_t, std::uint32_t top,
                                               std::uint16_t day_at_stop, int mam_at_stop,
                                               int lb_l, int tad_k, std::uint16_t base,
                                               et_counters& c) {
  using D = dirop<Fwd>;
  constexpr std::uint32_t kDays = 6U;  // kMaxTravelTime / 1 day + 1
  auto day = day_at_stop;
  std::uint32_t i = 0U;
  std::uint32_t pos = n_t;
  ++c.route_days;
  if (bits.route_active(day)) {
    std::uint32_t p = 0U;
    for (auto step = top; step != 0U; step >>= 1U) {
      auto const np = p + step;
      if (np <= n_t) {
        auto const t_off = Fwd ? np - 1U : n_t - np;
        if (D::better(ev[t_off] >> 5, mam_at_stop)) p = np;
      }
    }
    pos = p;
    c.events += p < n_t ? p + 1U : n_t;
  }
  et_t<Fwd> res{~0U, 0xFFFFU};
  std::uint32_t n_ev = 0U, n_days = 0U, n_tt = 0U;
  bool go = true;
  while (go) {
    if (pos >= n_t) {
      ++i;
      if (i == kDays) {
        go = false;
      } else {
        day = static_cast<std::uint16_t>(Fwd ? day + 1U : day - 1U);
        ++n_days;
        pos = bits.route_active(day) ? 0U : n_t;
      }
    } else {
      ++n_ev;
      auto const t_off = Fwd ? pos : n_t - 1U - pos;
      auto const e = ev[t_off];
      if (!D::better_eq(to_delta<Fwd>(day, e >> 5, base) + D::dir(lb_l), tad_k)) {
        go = false;
      } else {
        auto const start_day = static_cast<std::uint16_t>(static_cast<int>(day) - (e & 31));
        ++n_tt;
        if (bits.transport_active<Staged>(t_off, start_day)) {
          res = {t_off, start_day};
          go = false;
        } else {
          ++pos;
        }
      }
    }
  }
  c.events += n_ev;
  c.route_days += n_days;
  c.transport_tests += n_tt;
  return res;
}

template <bool Fwd>
__device__ inline std::int16_t time_at_stop(std::uint16_t const* ev, std::uint32_t n_t,
                                            et_t<Fwd> et, std::uint32_t stop_idx, bool arr,
                                            std::uint16_t base) {
  auto const e = ev[n_t * (stop_idx * 2U - (arr ? 1U : 0U)) + et.t_off];
  return to_delta<Fwd>(et.day, (e & 31) * 1440 + (e >> 5), base);
}

// Per-route data of one route scan: stop words, transport bitfield indices
// and the route's event block, either in LDS or in the global image.
struct route_view {
  std::uint32_t const* stops;
  std::uint16_t const* ev;
  route_bits bits;
  std::uint32_t n_stops, n_t;
};

__device__ inline void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// update_route (raptor.h:1181-1349), plain engine: one route per lane, all
// data read from the global image and state.
template <bool Fwd, int W>
__device__ inline void scan_route(route_view const& rv, dev_image const& im, dev_state const& s,
                                  std::uint32_t g, std::uint32_t r, std::uint32_t k,
                                  std::uint32_t lane_id, bool on, std::uint16_t base, int tad_k,
                                  unsigned long long& c_stops, unsigned long long& c_tmp,
                                  et_counters& c_et, bool& any) {
  using D = dirop<Fwd>;
  auto* overflow = &s.counts[4];
  auto const n_stops = rv.n_stops;
  auto const n_t = rv.n_t;
  et_t<Fwd> et{~0U, 0xFFFFU};
  bool done = !on;
  for (std::uint32_t i = 0U; i != n_stops; ++i) {
    auto const stop_idx = Fwd ? i : n_stops - i - 1U;
    auto const stp = rv.stops[stop_idx];
    auto const l = stp & ((1U << 28) - 1U);
    auto const mw = s.mark_t[lay<W>::lword(s, g, l)];
    auto const c = lay<W>::col(s, g, l, lane_id);
    int const lbv = done ? 0 : static_cast<int>(s.lb[c]);
    std::int16_t const prev = (!done && (mw & lay<W>::lbit(l, lane_id)) != 0ULL)
                                  ? s.round_times[lay<W>::rt(s, g, k - 1U, l, lane_id)]
                                  : D::kInvalid;
    auto const in_ok = ((stp >> 28) & 1U) != 0U, out_ok = ((stp >> 29) & 1U) != 0U;
    bool const can_finish = Fwd ? out_ok : in_ok;
    bool const can_start = Fwd ? in_ok : out_ok;
    if (!done) ++c_stops;
    bool ride = false;
    if (!done && et.valid() && can_finish) {
      ride = true;
      auto const by = time_at_stop<Fwd>(rv.ev, n_t, et, stop_idx, Fwd, base);
      if (D::better_eq(by, tad_k) && lbv != kUnreachableLb && D::better_eq(by + D::dir(lbv), tad_k) &&
          within_bounds<Fwd, W>(im, s, g * W + lane_id, k, l, by)) {
        if constexpr (Fwd) {
          atomicMin(&s.tmp[c], static_cast<int>(by));
        } else {
          atomicMax(&s.tmp[c], static_cast<int>(by));
        }
        ++c_tmp;
        any = true;
        mark_one<W>(s.mark_r, s.list_r, &s.counts[1], s.cap_r, overflow, s, g, l, 0U);
      }
    }
    bool const is_last = i == n_stops - 1U;
    if (done || is_last || !can_start || (mw & lay<W>::lbit(l, lane_id)) == 0ULL) continue;
    if (lbv == kUnreachableLb) {
      done = true;  // break (raptor.h:1309-1311)
      continue;
    }
    auto const et_dep = et.valid() ? time_at_stop<Fwd>(rv.ev, n_t, et, stop_idx, !Fwd, base)
                                   : D::kInvalid;
    if (prev != D::kInvalid && D::better_eq(prev, et_dep)) {
      int const minutes = static_cast<int>(base) * 1440 + prev;
      auto const day = static_cast<std::uint16_t>(minutes / 1440);
      int const mam = static_cast<std::int16_t>(minutes % 1440);
      auto const* ev_stop = rv.ev + n_t * (stop_idx * 2U - (Fwd ? 0U : 1U));
      auto const new_et = earliest_transport<Fwd, false>(rv.bits, ev_stop, n_t, 1U << (31 - __clz(n_t)),
                                                         day, mam, lbv, tad_k, base, c_et);
      if (new_et.valid()) {
        // current_best is only read when the new trip departs later (see
        // scan_route64)
        if (D::better_eq(time_at_stop<Fwd>(rv.ev, n_t, new_et, stop_idx, !Fwd, base), et_dep) ||
            !(ride || s.best[c] != D::kInvalid || s.tmp[c] != D::kInvalid)) {
          et = new_et;
        }
      }
    }
  }
}

// ---- batched route scan (one wavefront per (64-query group, route)) -------
//
// LDS per wavefront: the route's traffic-day bitfield and, for routes with at
// most kStagedTransports transports, the transports' bitfields (staged once
// per route); then, chunk by chunk, the stop words, station-mark words, the
// lower-bound and round_times[k-1] rows of the 64 lanes and the event blocks
// (arrival, departure) of the chunk's stops.  All staging loads of a chunk
// are issued together; the scan of the chunk then reads LDS only, except for
// transport bitfields of large routes and the rare current_best check.
constexpr std::uint32_t kChunkMax = 4U;
constexpr std::uint32_t kStagedTransports = 16U;
constexpr std::uint32_t kChunkEvBytes = 1024U;
// stop words staged per window (one per lane); a route scan starts at the
// scan index its packing found (16 bits in the packed item encodings)
constexpr std::uint32_t kStopWindow = 64U;
constexpr std::uint32_t kMarkBatch = 16U;  // mark words in flight per lane
constexpr std::uint32_t kMaxScanStart = 0xFFFFU;
static_assert(kChunkEvBytes / 4U >= kBatchedMaxRouteTransports, "a chunk holds at least one stop");
static_assert(kStagedTransports * 8U * 8U >= kBatchedMaxRouteTransports * 4U,
              "the bitfield staging area holds the bitfield indices of a staged route");

struct wave_lds {
  std::uint64_t rbits[8];
  std::uint64_t tbits[kStagedTransports * 8U];
  std::uint32_t stop[kStopWindow];
  std::uint16_t lb[kChunkMax][64];
  std::int16_t rt[kChunkMax][64];
  std::uint16_t ev[kChunkEvBytes / 2U];  // [chunk stop][arr, dep][n_t]
};
constexpr std::uint32_t kWaveLdsBytes = sizeof(wave_lds);
constexpr std::uint32_t kLocMask = (1U << 28) - 1U;

// One query of a packed wavefront: group g, lane ln of that group.
struct lane_query {
  std::uint32_t g, ln;
  bool on;
};

// update_route (raptor.h:1181-1349) for one route and up to 64 queries of
// any groups (one per lane).  Route data (stops, events, bitfields) is
// uniform and staged in LDS; each lane reads its own query's state.
template <bool Fwd>
__device__ inline void scan_route64(wave_lds& x, dev_image const& im, dev_state const& s,
                                    std::uint32_t r, std::uint32_t start, std::uint32_t k,
                                    std::uint32_t lane, lane_query const q, std::uint16_t base,
                                    int tad_k, std::uint32_t& c_stops, std::uint32_t& c_skip,
                                    std::uint32_t& c_tmp, std::uint32_t& c_ev, std::uint32_t& c_rd,
                                    std::uint32_t& c_tt, bool& any, bool& started) {
  using D = dirop<Fwd>;
  auto const L = s.L;
  auto const s0 = uni(im.route_stop_offsets[r]);
  auto const n_stops = uni(im.route_stop_offsets[r + 1U]) - s0;
  auto const t_from = uni(im.route_transport_from[r]);
  auto const n_t = uni(im.route_transport_from[r + 1U]) - t_from;
  auto const* stops_g = im.route_stops + s0;
  auto const* ev_g = im.event_times + uni(im.route_stop_time_from[r]);
  auto const* tbf_g = im.transport_traffic_days + t_from;
  // this lane's query: element l of a [G][L][64] array is at col0 + 64 l
  // (G * L * 64 < 2^32 is checked when the engine is created); the half of
  // the station-mark word (g, l) that holds its bit is u32 word mrow + 2 l
  std::uint32_t const col0 = q.g * L * 64U + q.ln;
  std::uint32_t const mrow = 2U * q.g * L + (q.ln >> 5);
  std::uint32_t const mbit = 1U << (q.ln & 31U);
  auto const* rt_q = s.round_times + (static_cast<std::size_t>(q.g) * s.K + (k - 1U)) * L * 64U + q.ln;
  auto const* mark_t32 = reinterpret_cast<std::uint32_t const*>(s.mark_t);

  // route-level staging: traffic-day bitfields; for routes of more than
  // kStagedTransports transports, the same LDS area holds the transports'
  // bitfield indices instead (at most 256 x 4 B), so that a traffic-day test
  // makes one global load instead of two dependent ones
  auto const rbf = uni(im.route_traffic_days[r]);
  bool const staged_t = n_t <= kStagedTransports;
  auto* tidx = reinterpret_cast<std::uint32_t*>(x.tbits);
  if (lane < 8U) x.rbits[lane] = im.bitfields[static_cast<std::size_t>(rbf) * 8U + lane];
  if (staged_t) {
    for (auto j = lane; j < 8U * n_t; j += 64U) {
      x.tbits[j] = im.bitfields[static_cast<std::size_t>(tbf_g[j / 8U]) * 8U + j % 8U];
    }
  } else {
    for (auto j = lane; j < n_t; j += 64U) tidx[j] = tbf_g[j];
  }
  route_bits const bits{x.rbits, x.tbits, tidx, im.bitfields, staged_t};
  auto const chunk = min(kChunkMax, kChunkEvBytes / (4U * n_t));
  auto const top = 1U << (31 - __clz(n_t));  // largest power of two <= n_t
  // the stop words of scan indices [ws, ws + kStopWindow), in scan order
  auto const stage_stops = [&](std::uint32_t w0) {
    if (w0 + lane < n_stops) {
      auto const si = w0 + lane;
      x.stop[lane] = stops_g[Fwd ? si : n_stops - si - 1U];
    }
  };
  std::uint32_t ws = start;
  stage_stops(ws);

  et_t<Fwd> et{~0U, 0xFFFFU};
  // this lane's station marks at the stops of the staged window, bit i for
  // scan index ws + i: loaded for the whole window at once, kMarkBatch stops
  // in flight, so that a chunk knows its marks before it loads anything
  std::uint64_t win_marks = 0ULL;
  auto const load_marks = [&](bool live) {
    win_marks = 0ULL;
    if (!live) return;
    auto const n_win = min(kStopWindow, n_stops - ws);
    for (std::uint32_t i0 = 0U; i0 < n_win; i0 += kMarkBatch) {
      std::uint32_t w[kMarkBatch];
#pragma unroll
      for (std::uint32_t j = 0U; j < kMarkBatch; ++j) {
        w[j] = i0 + j < n_win ? mark_t32[mrow + 2U * (x.stop[i0 + j] & kLocMask)] : 0U;
      }
#pragma unroll
      for (std::uint32_t j = 0U; j < kMarkBatch; ++j) {
        win_marks |= static_cast<std::uint64_t>((w[j] & mbit) != 0U ? 1U : 0U) << (i0 + j);
      }
    }
  };
  // `started`: one of the route's stops is marked for this query.  Before
  // its first marked stop a query has no trip (boarding needs a mark), so
  // nothing happens; its visited stops count only if the route was marked
  // for it (the packing admits only such queries, so on lanes start).  The
  // stops before `start` are marked for none of the wavefront's queries:
  // they are not visited, and count as visited stops of nigiri's loop.
  bool done = !q.on;
  started = false;
  std::uint32_t my_stops = done ? 0U : start;
  wave_sync_lds();  // the window's stop words
  load_marks(!done);
  for (std::uint32_t c0 = start; c0 < n_stops; c0 += chunk) {
    auto const nc = min(chunk, n_stops - c0);
    wave_sync_lds();  // readers of the previous chunk are done
    if (c0 + nc > ws + kStopWindow) {
      ws = c0;
      stage_stops(ws);
      wave_sync_lds();
      load_marks(!done);
    }
    auto const* cstop = x.stop + (c0 - ws);
    // bit j: this query is marked at chunk stop j
    auto const mbits = done ? 0U : static_cast<std::uint32_t>(win_marks >> (c0 - ws)) & ((1U << nc) - 1U);
    started = started || mbits != 0U;
    // A chunk in which no live lane rides and no live lane is marked changes
    // nothing: rides need a trip, boarding and the lower-bound break need a
    // mark.  Such a chunk only counts its stops.
    if (__ballot(!done && (et.valid() || mbits != 0U)) == 0ULL) {
      my_stops += done ? 0U : nc;
      continue;
    }
    // event blocks: for chunk stop j, [arr at stop_idx | dep at stop_idx];
    // the arrival row of stop_idx directly precedes its departure row in the
    // route's event array (stop-pair layout), so both rows are one run of
    // 2 n_t events starting at n_t (2 stop_idx - 1); a missing row (no
    // arrival at the first stop, no departure at the last) is zeroed.
    auto const stage_events = [&]() {
      for (std::uint32_t j = 0U; j < nc; ++j) {
        auto const si = c0 + j;
        auto const stop_idx = Fwd ? si : n_stops - si - 1U;
        auto const lo = stop_idx == 0U ? n_t : 0U;                  // first valid element
        auto const hi = stop_idx + 1U == n_stops ? n_t : 2U * n_t;  // end of valid elements
        auto const* src = ev_g + static_cast<std::ptrdiff_t>(n_t) * (2 * static_cast<std::ptrdiff_t>(stop_idx) - 1);
        auto* dst = x.ev + 2U * j * n_t;
        for (auto e = lane; e < 2U * n_t; e += 64U) {
          dst[e] = (e >= lo && e < hi) ? src[e] : std::uint16_t{0U};
        }
      }
    };
    // the lane values of the chunk and its event blocks, in flight together:
    // a lane that rides into the chunk needs lower bounds at every stop, one
    // that does not only from its first marked stop on, and round_times[k-1]
    // only at marked stops
    stage_events();
    bool const rides_in = !done && et.valid();
    std::uint16_t lbr[kChunkMax];
    std::int16_t rtr[kChunkMax];
    {
      auto const first = mbits != 0U ? static_cast<std::uint32_t>(__builtin_ctz(mbits)) : kChunkMax;
#pragma unroll
      for (std::uint32_t j = 0U; j < kChunkMax; ++j) {
        lbr[j] = 0U;
        rtr[j] = D::kInvalid;
        if (j < nc && !done) {
          auto const l = cstop[j] & kLocMask;
          if (rides_in || j >= first) lbr[j] = s.lb[col0 + l * 64U];
          if (((mbits >> j) & 1U) != 0U) rtr[j] = rt_q[static_cast<std::size_t>(l) * 64U];
        }
      }
    }
#pragma unroll
    for (std::uint32_t j = 0U; j < kChunkMax; ++j) {
      if (j < nc) {
        x.lb[j][lane] = lbr[j];
        x.rt[j][lane] = rtr[j];
      }
    }
    wave_sync_lds();

    for (std::uint32_t j = 0U; j < nc; ++j) {
      auto const i = c0 + j;
      auto const stp = uni(cstop[j]);
      auto const l = stp & kLocMask;
      bool const marked = ((mbits >> j) & 1U) != 0U;
      int const lbv = x.lb[j][lane];
      std::int16_t const prev = x.rt[j][lane];
      auto const* ev_arr = x.ev + (2U * j) * n_t;
      auto const* ev_dep = ev_arr + n_t;
      auto const in_ok = ((stp >> 28) & 1U) != 0U, out_ok = ((stp >> 29) & 1U) != 0U;
      bool const can_finish = Fwd ? out_ok : in_ok;
      bool const can_start = Fwd ? in_ok : out_ok;
      auto const off = col0 + l * 64U;
      my_stops += done ? 0U : 1U;

      bool ride = false;
      if (!done && et.valid() && can_finish) {
        ride = true;
        auto const e = (Fwd ? ev_arr : ev_dep)[et.t_off];
        auto const by = to_delta<Fwd>(et.day, (e & 31) * 1440 + (e >> 5), base);
        if (D::better_eq(by, tad_k) && lbv != kUnreachableLb && D::better_eq(by + D::dir(lbv), tad_k)) {
          if constexpr (Fwd) {
            atomicMin(&s.tmp[off], static_cast<int>(by));
          } else {
            atomicMax(&s.tmp[off], static_cast<int>(by));
          }
          s.mark_rb[off] = 1U;
          s.flag_r[static_cast<std::size_t>(l) * s.G + q.g] = 1U;
          any = true;
          ++c_tmp;
        }
      }

      bool const is_last = i == n_stops - 1U;
      if (done || is_last || !can_start || !marked) continue;
      if (lbv == kUnreachableLb) {
        done = true;  // break (raptor.h:1309-1311)
        continue;
      }
      auto const* ev_board = Fwd ? ev_dep : ev_arr;
      std::int16_t et_dep = D::kInvalid;
      if (et.valid()) {
        auto const e = ev_board[et.t_off];
        et_dep = to_delta<Fwd>(et.day, (e & 31) * 1440 + (e >> 5), base);
      }
      if (prev != D::kInvalid && D::better_eq(prev, et_dep)) {
        int const minutes = static_cast<int>(base) * 1440 + prev;
        auto const day = static_cast<std::uint16_t>(minutes / 1440);
        int const mam = static_cast<std::int16_t>(minutes % 1440);
        et_counters ce;
        auto const new_et =
            staged_t ? earliest_transport<Fwd, true>(bits, ev_board, n_t, top, day, mam, lbv, tad_k, base, ce)
                     : earliest_transport<Fwd, false>(bits, ev_board, n_t, top, day, mam, lbv, tad_k, base, ce);
        c_ev += static_cast<std::uint32_t>(ce.events);
        c_rd += static_cast<std::uint32_t>(ce.route_days);
        c_tt += static_cast<std::uint32_t>(ce.transport_tests);
        if (new_et.valid()) {
          // et = new_et if current_best (raptor.h:1328-1329) is invalid or
          // the new trip departs no later.  The search scans departures in
          // time order from prev <= et_dep, so a found trip departs no
          // later than et: the state loads of current_best are only made
          // when that test fails.  current_best is valid whenever a ride
          // was evaluated at this stop (it then includes the valid
          // round_times[k-1] label) or best / tmp hold a label.
          auto const e = ev_board[new_et.t_off];
          auto const new_dep = to_delta<Fwd>(new_et.day, (e & 31) * 1440 + (e >> 5), base);
          if (D::better_eq(new_dep, et_dep) ||
              !(ride || s.best[off] != D::kInvalid || s.tmp[off] != D::kInvalid)) {
            et = new_et;
          }
        }
      }
    }
  }
  started = started && q.on;
  if (started) {
    c_stops += my_stops;
    c_skip += start;
  }
}

// update_route (raptor.h:1181-1349), one lane per (route, query): for routes
// whose stop events do not fit the LDS staging of scan_route64 (more than
// kBatchedMaxRouteTransports transports) and for the loose route items.
// Every lane scans its route for its own query, reading the image and the
// state from global memory stop by stop.
template <bool Fwd>
__device__ inline void scan_route_lane64(dev_image const& im, dev_state const& s, std::uint32_t r,
                                         std::uint32_t start, std::uint32_t k, lane_query const q,
                                         std::uint16_t base, int tad_k, std::uint32_t& c_stops,
                                         std::uint32_t& c_skip, std::uint32_t& c_tmp, std::uint32_t& c_ev,
                                         std::uint32_t& c_rd, std::uint32_t& c_tt, bool& any, bool& started) {
  using D = dirop<Fwd>;
  auto const L = s.L;
  auto const s0 = im.route_stop_offsets[r];
  auto const n_stops = im.route_stop_offsets[r + 1U] - s0;
  auto const t_from = im.route_transport_from[r];
  auto const n_t = im.route_transport_from[r + 1U] - t_from;
  auto const* stops = im.route_stops + s0;
  auto const* ev = im.event_times + im.route_stop_time_from[r];
  route_bits const bits{im.bitfields + static_cast<std::size_t>(im.route_traffic_days[r]) * 8U, nullptr,
                        im.transport_traffic_days + t_from, im.bitfields, false};
  auto const top = 1U << (31 - __clz(n_t));
  std::uint32_t const col0 = q.g * L * 64U + q.ln;
  std::uint32_t const mrow = 2U * q.g * L + (q.ln >> 5);
  std::uint32_t const mbit = 1U << (q.ln & 31U);
  auto const* mark_t32 = reinterpret_cast<std::uint32_t const*>(s.mark_t);
  auto const* rt_q = s.round_times + (static_cast<std::size_t>(q.g) * s.K + (k - 1U)) * L * 64U + q.ln;
  started = q.on;  // the packing admits only queries for which the route is marked
  if (!q.on) return;
  // no stop before `start` is marked for this query: nothing happens there
  // (counted as visited stops of nigiri's loop)
  c_stops += start;
  c_skip += start;
  et_t<Fwd> et{~0U, 0xFFFFU};
  et_counters ce;
  for (std::uint32_t i = start; i < n_stops; ++i) {
    auto const stop_idx = Fwd ? i : n_stops - i - 1U;
    auto const stp = stops[stop_idx];
    auto const l = stp & kLocMask;
    auto const off = col0 + l * 64U;
    bool const marked = (mark_t32[mrow + 2U * l] & mbit) != 0U;
    ++c_stops;
    // without a trip and a mark nothing happens at this stop (its lower
    // bound is not needed)
    if (!et.valid() && !marked) continue;
    int const lbv = s.lb[off];
    auto const in_ok = ((stp >> 28) & 1U) != 0U, out_ok = ((stp >> 29) & 1U) != 0U;
    bool const can_finish = Fwd ? out_ok : in_ok;
    bool const can_start = Fwd ? in_ok : out_ok;
    bool ride = false;
    if (et.valid() && can_finish) {
      ride = true;
      auto const by = time_at_stop<Fwd>(ev, n_t, et, stop_idx, Fwd, base);
      if (D::better_eq(by, tad_k) && lbv != kUnreachableLb && D::better_eq(by + D::dir(lbv), tad_k)) {
        if constexpr (Fwd) {
          atomicMin(&s.tmp[off], static_cast<int>(by));
        } else {
          atomicMax(&s.tmp[off], static_cast<int>(by));
        }
        s.mark_rb[off] = 1U;
        s.flag_r[static_cast<std::size_t>(l) * s.G + q.g] = 1U;
        any = true;
        ++c_tmp;
      }
    }
    if (i == n_stops - 1U || !can_start || !marked) continue;
    if (lbv == kUnreachableLb) break;  // raptor.h:1309-1311
    auto const prev = rt_q[static_cast<std::size_t>(l) * 64U];
    auto const et_dep = et.valid() ? time_at_stop<Fwd>(ev, n_t, et, stop_idx, !Fwd, base) : D::kInvalid;
    if (prev != D::kInvalid && D::better_eq(prev, et_dep)) {
      int const minutes = static_cast<int>(base) * 1440 + prev;
      auto const day = static_cast<std::uint16_t>(minutes / 1440);
      int const mam = static_cast<std::int16_t>(minutes % 1440);
      auto const* ev_board = ev + n_t * (stop_idx * 2U - (Fwd ? 0U : 1U));
      auto const new_et = earliest_transport<Fwd, false>(bits, ev_board, n_t, top, day, mam, lbv, tad_k, base, ce);
      if (new_et.valid()) {
        // current_best is only read when the new trip departs later (see
        // scan_route64)
        if (D::better_eq(time_at_stop<Fwd>(ev, n_t, new_et, stop_idx, !Fwd, base), et_dep) ||
            !(ride || s.best[off] != D::kInvalid || s.tmp[off] != D::kInvalid)) {
          et = new_et;
        }
      }
    }
  }
  c_ev += static_cast<std::uint32_t>(ce.events);
  c_rd += static_cast<std::uint32_t>(ce.route_days);
  c_tt += static_cast<std::uint32_t>(ce.transport_tests);
}

// route_pack (batched engine): the flagged (route, group) pairs, compacted
// in route-major order, become packed wavefronts that each hold up to 64
// queries (group, lane) of one route.  A query takes part in a route when
// it is active and one of the route's stops is marked for it: those are the
// queries for which mark_routes (raptor.h:228-241) marked the route.
// Pairs are packed greedily in list order; one wavefront packs a window of
// 64 list entries.  Each pair also yields its start: the scan index of the
// route's first stop marked for one of its queries.  Before that stop none
// of them has a trip (boarding needs a mark), so update_route changes
// nothing there and the scan begins at the start (a packed wavefront at
// the smallest start of its pairs).
template <bool Fwd>
__global__ void k_route_pack(round_args a) {
  auto const& im = a.im;
  auto const& s = a.st;
  __shared__ std::uint32_t fills_all[kWavesPerBlock][64];
  // per packed wavefront of the window: its slot among the kept ones (or
  // ~0 when it goes loose) / its offset among the loose lanes; the window's
  // two bases
  __shared__ std::uint32_t kidx_all[kWavesPerBlock][64], loff_all[kWavesPerBlock][64];
  __shared__ std::uint32_t bases_all[kWavesPerBlock][2];
  __shared__ std::uint32_t starts_all[kWavesPerBlock][64];
  auto* fills = fills_all[threadIdx.x / 64U];
  auto* starts = starts_all[threadIdx.x / 64U];
  auto* kidx = kidx_all[threadIdx.x / 64U];
  auto* loff = loff_all[threadIdx.x / 64U];
  auto* bases = bases_all[threadIdx.x / 64U];
  auto const lane = threadIdx.x & 63U;
  auto const n_items = clamp_count(s.counts[2], s.cap_route);
  auto const wave = (blockIdx.x * blockDim.x + threadIdx.x) / 64U;
  for (std::uint32_t w0 = wave * 64U; w0 < n_items; w0 += gridDim.x * kWavesPerBlock * 64U) {
    auto const item = w0 + lane;
    std::uint32_t r = ~0U, g = 0U, start = 0U;
    std::uint64_t mask = 0ULL;
    if (item < n_items) {
      auto const e = s.list_route[item];
      r = im.rank_route[static_cast<std::uint32_t>(e >> 32)];
      g = static_cast<std::uint32_t>(e);
      auto const act = s.grp_active[g];
      if (act != 0ULL) {
        auto const* mt = s.mark_t + static_cast<std::size_t>(g) * s.L;
        auto const i0 = im.route_stop_offsets[r], i1 = im.route_stop_offsets[r + 1U];
        // first and last marked stop (array order) of the active queries
        std::uint32_t first = i1, last = i0;
        auto i = i0;
        // eight stops at a time: their stop words, then their mark words, in
        // flight together
        constexpr std::uint32_t kPackStops = 8U;
        for (; i + kPackStops <= i1; i += kPackStops) {
          std::uint32_t a[kPackStops];
          std::uint64_t m[kPackStops];
#pragma unroll
          for (std::uint32_t j = 0U; j < kPackStops; ++j) a[j] = im.route_stops[i + j] & kLocMask;
#pragma unroll
          for (std::uint32_t j = 0U; j < kPackStops; ++j) m[j] = mt[a[j]] & act;
          std::uint64_t any = 0ULL;
#pragma unroll
          for (std::uint32_t j = 0U; j < kPackStops; ++j) {
            if (m[j] != 0ULL) {
              if (first == i1) first = i + j;
              last = i + j;
            }
            any |= m[j];
          }
          mask |= any;
        }
        for (; i < i1; ++i) {
          auto const m0 = mt[im.route_stops[i] & kLocMask] & act;
          if (m0 != 0ULL) {
            if (first == i1) first = i;
            last = i;
            mask |= m0;
          }
        }
        // scan order runs backward through the stops for backward searches
        if (mask != 0ULL) start = min(Fwd ? first - i0 : i1 - 1U - last, kMaxScanStart);
      }
    }
    auto const c = static_cast<std::uint32_t>(__popcll(mask));
    // greedy assignment in item order (uniform loop over the 64 entries)
    // (at most max_entries list entries, i.e. groups, per wavefront)
    auto const max_entries = a.tune.pack_max_entries;
    std::uint32_t my_w = 0U, my_o = 0U, n_w = 0U, fill = 0U, cur_r = ~0U, entries = 0U, w_start = 0U;
    for (std::uint32_t t = 0U; t < 64U; ++t) {
      auto const ct = static_cast<std::uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(c), static_cast<int>(t)));
      if (ct == 0U) continue;
      auto const rt = static_cast<std::uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(r), static_cast<int>(t)));
      auto const st = static_cast<std::uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(start), static_cast<int>(t)));
      if (n_w == 0U || rt != cur_r || fill + ct > 64U || entries == max_entries) {
        if (n_w != 0U && lane == 0U) {
          fills[n_w - 1U] = fill;
          starts[n_w - 1U] = w_start;
        }
        ++n_w;
        fill = 0U;
        entries = 0U;
        cur_r = rt;
        w_start = st;
      }
      ++entries;
      w_start = min(w_start, st);
      if (lane == t) {
        my_w = n_w - 1U;
        my_o = fill;
      }
      fill += ct;
    }
    if (n_w != 0U && lane == 0U) {
      fills[n_w - 1U] = fill;
      starts[n_w - 1U] = w_start;
    }
    wave_sync_lds();
    // wavefronts of fewer than loose_lanes queries go loose: their queries
    // are scanned one (route, query) per lane; the kept ones take packed
    // slots without gaps
    if (lane == 0U) {
      std::uint32_t n_keep = 0U, n_loose = 0U;
      for (std::uint32_t w = 0U; w < n_w; ++w) {
        if (fills[w] < a.tune.loose_lanes) {
          kidx[w] = ~0U;
          loff[w] = n_loose;
          n_loose += fills[w];
        } else {
          kidx[w] = n_keep++;
        }
      }
      bases[0] = n_keep != 0U ? atomicAdd(&s.counts[5], n_keep) : 0U;
      bases[1] = n_loose != 0U ? atomicAdd(&s.counts[8], n_loose) : 0U;
    }
    wave_sync_lds();
    if (c != 0U) {
      if (kidx[my_w] != ~0U) {
        auto const pw = static_cast<std::size_t>(bases[0]) + kidx[my_w];
        if (pw < s.cap_pack) {
          if (my_o == 0U) {
            s.packed_info[pw] = (static_cast<std::uint64_t>(r) << 32) | (starts[my_w] << 16) | fills[my_w];
          }
          auto* out = s.packed_lanes + pw * 64U + my_o;
          auto m = mask;
          for (std::uint32_t j = 0U; m != 0ULL; ++j) {
            auto const b = static_cast<std::uint32_t>(__ffsll(static_cast<long long>(m)) - 1);
            m &= m - 1ULL;
            out[j] = static_cast<std::uint16_t>((g << 6) | b);
          }
        } else {
          atomicOr(&s.counts[4], 1U);
        }
      } else {
        auto const at = static_cast<std::size_t>(bases[1]) + loff[my_w] + my_o;
        if (at + c <= s.cap_loose) {
          auto* out = s.loose + at;
          auto m = mask;
          for (std::uint32_t j = 0U; m != 0ULL; ++j) {
            auto const b = static_cast<std::uint32_t>(__ffsll(static_cast<long long>(m)) - 1);
            m &= m - 1ULL;
            out[j] = (static_cast<std::uint64_t>(r) << 32) | (start << 16) | (g << 6) | b;
          }
        } else {
          atomicOr(&s.counts[4], 1U);
        }
      }
    }
    wave_sync_lds();  // fills is reused by the next window
  }
}

// route_scan: loop_routes + update_route (raptor.h:496-575, 1181-1349).
// The route phase reads only round_times[k-1], best, lb, time_at_dest and
// the previous marks, and writes tmp by minimum and marks by OR: route
// order does not matter, so routes run in parallel.  Batched engine: one
// packed wavefront per route (up to 64 queries of any groups, route data
// staged in LDS once); plain engine: one route per lane.
template <bool Fwd, int W>
__global__ void __launch_bounds__(256, 8) k_route_scan(round_args a) {
  auto const& im = a.im;
  auto const& s = a.st;
  if constexpr (W == 64) {
    __shared__ wave_lds lds_all[kWavesPerBlock];
    auto& x = lds_all[threadIdx.x / 64U];
    std::uint32_t const lane = threadIdx.x & 63U;
    auto const n_items = clamp_count(s.counts[5], s.cap_pack);
    std::uint32_t c_stops = 0U, c_skip = 0U, c_tmp = 0U, c_ev = 0U, c_rd = 0U, c_tt = 0U, c_ws = 0U, c_lws = 0U,
                  c_li = 0U;
    item_claim claim{&s.counts[6], a.tune.route_claim};
    for (std::uint32_t item = claim.get(); item < n_items; item = claim.get()) {
      auto const info = s.packed_info[item];
      auto const r = uni(static_cast<std::uint32_t>(info >> 32));
      auto const n_lanes = uni(static_cast<std::uint32_t>(info) & 0xFFFFU);
      auto const start = uni((static_cast<std::uint32_t>(info) >> 16) & 0xFFFFU);
      lane_query q{0U, 0U, false};
      if (lane < n_lanes) {
        auto const d = s.packed_lanes[static_cast<std::size_t>(item) * 64U + lane];
        q.g = static_cast<std::uint32_t>(d) >> 6;
        q.ln = static_cast<std::uint32_t>(d) & 63U;
        q.on = ((s.grp_active[q.g] >> q.ln) & 1ULL) != 0ULL;
      }
      auto const slot = q.g * 64U + q.ln;
      auto const base = s.param[slot].base;
      auto const k = static_cast<std::uint32_t>(s.kslot[slot]);  // this query's round
      int const tad_k = s.tad[static_cast<std::size_t>(slot) * kRowsMax + k];
      bool any = false, started = false;
      wave_sync_lds();  // the previous item's readers are done
      auto const n_t = uni(im.route_transport_from[r + 1U]) - uni(im.route_transport_from[r]);
      if (n_t <= kBatchedMaxRouteTransports) {
        scan_route64<Fwd>(x, im, s, r, start, k, lane, q, base, tad_k, c_stops, c_skip, c_tmp, c_ev, c_rd,
                          c_tt, any, started);
      } else {
        scan_route_lane64<Fwd>(im, s, r, start, k, q, base, tad_k, c_stops, c_skip, c_tmp, c_ev, c_rd, c_tt,
                               any, started);
      }
      if (__ballot(started) != 0ULL) {
        c_ws += im.route_stop_offsets[r + 1U] - im.route_stop_offsets[r] - start;
      }
      if (any) s.any_tmp[slot] = 1U;
    }
    // loose items: 64 per wavefront, each lane scans its own route for its
    // own query from global memory
    auto const n_loose = clamp_count(s.counts[8], s.cap_loose);
    while (true) {
      std::uint32_t b = 0U;
      if (lane == 0U) b = atomicAdd(&s.counts[9], 64U);
      b = uni(b);
      if (b >= n_loose) break;
      auto const it = b + lane;
      lane_query q{0U, 0U, false};
      std::uint32_t r = 0U, n_stops = 0U, start = 0U;
      if (it < n_loose) {
        auto const v = s.loose[it];
        r = static_cast<std::uint32_t>(v >> 32);
        start = (static_cast<std::uint32_t>(v) >> 16) & 0xFFFFU;
        q.g = (static_cast<std::uint32_t>(v) >> 6) & 1023U;
        q.ln = static_cast<std::uint32_t>(v) & 63U;
        q.on = ((s.grp_active[q.g] >> q.ln) & 1ULL) != 0ULL;
        n_stops = im.route_stop_offsets[r + 1U] - im.route_stop_offsets[r];
      }
      auto const slot = q.g * 64U + q.ln;
      auto const k = static_cast<std::uint32_t>(s.kslot[slot]);
      int const tad_k = s.tad[static_cast<std::size_t>(slot) * kRowsMax + k];
      bool any = false, started = false;
      if (it < n_loose) {
        scan_route_lane64<Fwd>(im, s, r, start, k, q, s.param[slot].base, tad_k, c_stops, c_skip, c_tmp, c_ev,
                               c_rd, c_tt, any, started);
      }
      // wavefront iterations: the longest scan of the 64
      auto ws = started ? n_stops - start : 0U;
      for (int off = 32; off > 0; off >>= 1) ws = max(ws, static_cast<std::uint32_t>(__shfl_xor(static_cast<int>(ws), off, 64)));
      c_ws += ws;
      c_lws += ws;
      c_li += min(64U, n_loose - b);
      if (any) s.any_tmp[slot] = 1U;
    }
    wave_count(&s.work[kWorkRouteStops], c_stops);
    wave_count(&s.work[kWorkSkippedStops], c_skip);
    wave_count(&s.work[kWorkTmpWrites], c_tmp);
    wave_count(&s.work[kWorkEvents], c_ev);
    wave_count(&s.work[kWorkRouteDayTests], c_rd);
    wave_count(&s.work[kWorkTransportTests], c_tt);
    if (lane == 0U && c_ws != 0U) atomicAdd(&s.work[kWorkWaveStops], static_cast<unsigned long long>(c_ws));
    if (lane == 0U && c_lws != 0U) {
      atomicAdd(&s.work[kWorkLooseWaveStops], static_cast<unsigned long long>(c_lws));
      atomicAdd(&s.work[kWorkLooseItems], static_cast<unsigned long long>(c_li));
    }
  } else {
    auto const n_items = clamp_count(s.counts[2], s.cap_route);
    unsigned long long c_stops = 0ULL, c_tmp = 0ULL;
    et_counters c_et;
    for (std::uint32_t item = blockIdx.x * blockDim.x + threadIdx.x; item < n_items;
         item += gridDim.x * blockDim.x) {
      auto const e = s.list_route[item];
      auto const g = static_cast<std::uint32_t>(e >> 32), r = static_cast<std::uint32_t>(e);
      auto const rw = lay<W>::rword(s, g, r);
      auto const mword = s.route_mark[rw];
      auto const lanes = ((mword >> (r % 64U)) & 1ULL) & s.grp_active[g] &
                         (s.any_route[g] != 0U ? 1ULL : 0ULL);
      // several routes share a word; clear only this route's bit
      atomicAnd(reinterpret_cast<unsigned long long*>(&s.route_mark[rw]),
                ~static_cast<unsigned long long>(lay<W>::rbit(r, 0U)));
      if (lanes == 0ULL) continue;
      auto const slot = g;
      auto const base = s.param[slot].base;
      auto const k = static_cast<std::uint32_t>(s.kslot[slot]);
      int const tad_k = s.tad[static_cast<std::size_t>(slot) * kRowsMax + k];
      auto const s0 = im.route_stop_offsets[r];
      auto const t_from = im.route_transport_from[r];
      auto const rbf = im.route_traffic_days[r];
      route_view const rv{im.route_stops + s0, im.event_times + im.route_stop_time_from[r],
                          route_bits{im.bitfields + static_cast<std::size_t>(rbf) * 8U, nullptr,
                                     im.transport_traffic_days + t_from, im.bitfields, false},
                          im.route_stop_offsets[r + 1U] - s0, im.route_transport_from[r + 1U] - t_from};
      bool any = false;
      scan_route<Fwd, W>(rv, im, s, g, r, k, 0U, true, base, tad_k, c_stops, c_tmp, c_et, any);
      if (any) s.any_tmp[g] = 1U;
    }
    wave_count(&s.work[kWorkRouteStops], c_stops);
    wave_count(&s.work[kWorkTmpWrites], c_tmp);
    wave_count(&s.work[kWorkEvents], c_et.events);
    wave_count(&s.work[kWorkRouteDayTests], c_et.route_days);
    wave_count(&s.work[kWorkTransportTests], c_et.transport_tests);
  }
}

// Compaction (batched engine): flag bytes of a [major][minor] array -> work
// list entries (major << 32 | minor); clears the flags.  Each lane tests
// kCompactWords 64-bit words (8 flags each); a workgroup scan places the
// entries, with one atomic per workgroup and pass, so the entries of one
// workgroup's kCompactWords * 256 words are in ascending order.
constexpr std::uint32_t kCompactWords = 4U;

__global__ void __launch_bounds__(kBlock) k_compact(std::uint8_t* flags, std::uint32_t n_per_group,
                                                    std::size_t n_words, std::uint64_t* list,
                                                    std::uint32_t* count, std::size_t cap,
                                                    std::uint32_t* overflow) {
  __shared__ std::uint32_t wave_total[kWavesPerBlock];
  __shared__ std::uint32_t block_base;
  auto* words = reinterpret_cast<std::uint64_t*>(flags);
  auto const lane = threadIdx.x & 63U, wave = threadIdx.x / 64U;
  constexpr std::size_t kPerBlock = static_cast<std::size_t>(kBlock) * kCompactWords;
  for (std::size_t b0 = blockIdx.x * kPerBlock; b0 < n_words; b0 += gridDim.x * kPerBlock) {
    auto const i0 = b0 + threadIdx.x * static_cast<std::size_t>(kCompactWords);
    std::uint64_t v[kCompactWords];
    unsigned n = 0U;
#pragma unroll
    for (std::uint32_t j = 0U; j < kCompactWords; ++j) {
      v[j] = i0 + j < n_words ? words[i0 + j] : 0ULL;
      n += static_cast<unsigned>(__popcll(v[j]));  // flags are 0 or 1
    }
    // the workgroup's flags in ascending order: wave scans, then the waves'
    // totals (uniform branches: every wave of the block takes them)
    unsigned incl = n;
    for (int off = 1; off < 64; off <<= 1) {
      auto const t = __shfl_up(incl, off, 64);
      if (static_cast<int>(lane) >= off) incl += t;
    }
    if (lane == 63U) wave_total[wave] = incl;
    __syncthreads();
    unsigned before = 0U, total = 0U;
    for (std::uint32_t w = 0U; w < kWavesPerBlock; ++w) {
      auto const t = wave_total[w];
      if (w < wave) before += t;
      total += t;
    }
    if (threadIdx.x == 0U) block_base = total != 0U ? atomicAdd(count, total) : 0U;
    __syncthreads();
    auto pos = block_base + before + incl - n;
    bool const narrow = n_words <= (std::size_t{1} << 29);  // flag indices below 2^32
#pragma unroll
    for (std::uint32_t j = 0U; j < kCompactWords; ++j) {
      auto x = v[j];
      if (x == 0ULL) continue;
      words[i0 + j] = 0ULL;
      auto const f0 = (i0 + j) * 8U;
      while (x != 0ULL) {
        auto const b = static_cast<std::uint32_t>(__ffsll(static_cast<long long>(x)) - 1);
        x &= x - 1ULL;
        auto const idx = f0 + b / 8U;
        if (pos < cap) {
          std::uint64_t e = 0ULL;
          if (narrow) {
            auto const i32 = static_cast<std::uint32_t>(idx);
            e = (static_cast<std::uint64_t>(i32 / n_per_group) << 32) | (i32 % n_per_group);
          } else {
            e = (static_cast<std::uint64_t>(idx / n_per_group) << 32) | (idx % n_per_group);
          }
          list[pos] = e;
        } else {
          atomicOr(overflow, 1U);
        }
        ++pos;
      }
    }
    __syncthreads();  // wave_total and block_base are reused by the next pass
  }
}

// clear the marks listed in `list` (all lanes)
template <int W>
__global__ void k_clear_marks(dev_state s, int which) {
  auto const n = which == 0   ? clamp_count(s.counts[0], s.cap_t)
                 : which == 1 ? clamp_count(s.counts[1], s.cap_r)
                              : clamp_count(s.counts[3], s.cap_hz);
  auto* marks = which == 0 ? s.mark_t : (which == 1 ? s.mark_r : s.hz_mask);
  auto const* list = which == 0 ? s.list_t : (which == 1 ? s.list_r : s.list_hz);
  if (W == 64 && which == 1) {
    // batched route-phase marks: 64 bytes per (g, l), cleared as 8 words by
    // 8 consecutive lanes; route-phase list entries are (l << 32 | g).  The
    // marked lanes (their tmp was written) are added to bt_valid.
    auto* words = reinterpret_cast<std::uint64_t*>(s.mark_rb);
    auto const lane = threadIdx.x & 63U;
    auto const total = n * 8U;
    // wavefront-uniform loop: the stride is a multiple of 64, so i % 8 == lane % 8
    for (std::uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~63U); base < total;
         base += gridDim.x * blockDim.x) {
      auto const i = base + lane;
      std::uint64_t m = 0ULL;
      std::size_t gl = 0U;
      if (i < total) {
        auto const e = list[i / 8U];
        gl = static_cast<std::size_t>(static_cast<std::uint32_t>(e)) * s.L + (e >> 32);
        auto const w = words[gl * 8U + lane % 8U];
        if (w != 0ULL) {
          words[gl * 8U + lane % 8U] = 0ULL;
          for (std::uint32_t b = 0U; b < 8U; ++b) {
            if (((w >> (8U * b)) & 0xFFULL) != 0ULL) m |= 1ULL << (8U * (lane % 8U) + b);
          }
        }
      }
      for (int off = 1; off < 8; off <<= 1) {
        m |= static_cast<std::uint64_t>(__shfl_xor(static_cast<unsigned long long>(m), off, 64));
      }
      if (i < total && lane % 8U == 0U && m != 0ULL) s.bt_valid[gl] |= m;
    }
    return;
  }
  for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    auto const e = list[i];
    auto const w = lay<W>::lword(s, static_cast<std::uint32_t>(e >> 32), static_cast<std::uint32_t>(e));
    if constexpr (W == 64) {
      // station marks of the transfer phase (best was written for the marked
      // lanes) and of the starts: added to bt_valid
      if (which == 0) {
        auto const m = marks[w];
        if (m != 0ULL) s.bt_valid[w] |= m;
      }
    }
    marks[w] = 0ULL;
  }
}

// group_update: slots without marked routes or without route improvements
// end their epoch (raptor.h:243-246, 370-373).  One thread per slot; the
// batched engine gathers a group's 64 slots with one ballot.
__global__ void k_group_update(dev_state s) {
  auto const slot = blockIdx.x * blockDim.x + threadIdx.x;
  auto const n_slots = s.G * s.W;
  bool const keep = slot < n_slots && s.any_route[slot] != 0U && s.any_tmp[slot] != 0U;
  if (s.W == 64U) {
    auto const m = __ballot(keep);
    if (slot < n_slots && (threadIdx.x & 63U) == 0U) s.grp_active[slot / 64U] &= m;
  } else if (slot < n_slots && !keep) {
    s.grp_active[slot] = 0ULL;
  }
  if (slot < n_slots) {
    s.any_route[slot] = 0U;
    s.any_tmp[slot] = 0U;
  }
}

__global__ void k_deactivate_all(dev_state s) {
  for (std::uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < s.G; g += gridDim.x * blockDim.x) {
    s.grp_active[g] = 0ULL;
  }
  for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < s.G * s.W;
       i += gridDim.x * blockDim.x) {
    s.any_route[i] = 0U;
    s.any_tmp[i] = 0U;
  }
}

// clears the work-list counts selected by `mask` (bit i: counts[i]; bit 4,
// the overflow flags, is never cleared here)
__global__ void k_reset_counts(dev_state s, std::uint32_t mask) {
  if (threadIdx.x == 0U && blockIdx.x == 0U) {
    for (int i = 0; i < 16; ++i) {
      if (i != 4 && ((mask >> i) & 1U) != 0U) s.counts[i] = 0U;
    }
  }
}

// ---- destination events (exact sequential simulation per query) ---------

template <bool Fwd, int W>
__global__ void k_dest_pass(round_args a, std::uint32_t n_slots) {
  using D = dirop<Fwd>;
  auto const& im = a.im;
  auto const& s = a.st;
  auto const slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= n_slots) return;
  auto const k = static_cast<std::uint32_t>(s.kslot[slot]);
  auto const g = slot / W, lane = slot % W;
  auto* tad = s.tad + static_cast<std::size_t>(slot) * kRowsMax;
  s.n_chg[slot] = 0U;
  s.tad0[slot] = tad[k < kRowsMax ? k : kRowsMax - 1];
  if ((s.grp_active[g] & (1ULL << lane)) == 0ULL) return;
  auto const& p = s.param[slot];

  auto* ev = s.dest_ev + static_cast<std::size_t>(slot) * kMaxDestEvents;
  std::uint32_t n = 0U;
  auto const add = [&](std::uint32_t pos, std::uint32_t target, int x, std::uint8_t kind) {
    if (n < kMaxDestEvents) {
      ev[n++] = dest_event{pos, target, clamp16(x), kind};
    } else {
      atomicOr(&s.counts[4], 2U);
    }
  };
  if (p.intermodal != 0U) {
    for (std::uint32_t j = 0; j < p.n_dist; ++j) {
      auto const i = s.dist_loc[slot * kMaxDist + j];
      if (!marked_r<W>(s, g, i, lane)) continue;
      auto const t = s.tmp[lay<W>::col(s, g, i, lane)];
      if (t == D::kInvalid) continue;
      add(kPosIntermodal | i, 1U, t + D::dir(s.dist_min[slot * kMaxDist + j]), 1U);
    }
  } else {
    for (std::uint32_t j = 0; j < p.n_dest; ++j) {
      auto const d = s.dest[slot * kMaxDestBits + j];
      if (marked_r<W>(s, g, d, lane)) {
        auto const t = s.tmp[lay<W>::col(s, g, d, lane)];
        if (t != D::kInvalid) add(kPosTransfer | d, d, t, 0U);  // transfer time 0 at dest
      }
      // footpaths ending in d, at their position in the source's list
      auto const* off = Fwd ? im.fp_in_offsets : im.fp_out_offsets;
      auto const* fps = Fwd ? im.fp_in : im.fp_out;
      auto const* map = Fwd ? im.fp_in_to_out : im.fp_out_to_in;
      for (auto e = off[d]; e != off[d + 1U]; ++e) {
        auto const i = fps[e] & ((1U << 23) - 1U);
        if (!marked_r<W>(s, g, i, lane)) continue;
        auto const t = s.tmp[lay<W>::col(s, g, i, lane)];
        if (t == D::kInvalid) continue;
        auto const dur = static_cast<std::int16_t>(fps[e] >> 23);
        add(kPosFootpath | map[e], d, t + D::dir(adjusted<std::int16_t>(p, dur)), 0U);
      }
    }
  }
  // insertion sort by position
  for (std::uint32_t i = 1U; i < n; ++i) {
    auto const v = ev[i];
    auto j = i;
    while (j > 0U && ev[j - 1U].pos > v.pos) {
      ev[j] = ev[j - 1U];
      --j;
    }
    ev[j] = v;
  }
  std::uint32_t n_chg = 0U;
  for (std::uint32_t i = 0U; i < n; ++i) {
    auto const& e = ev[i];
    auto const before = tad[k];
    auto const c = lay<W>::col(s, g, e.target, lane);
    auto const rti = lay<W>::rt(s, g, k, e.target, lane);
    if (e.kind == 1U) {
      // update_intermodal_footpaths (raptor.h:1009-1025)
      if (D::better_eq(e.x, tad[k])) {
        s.round_times[rti] = e.x;
        s.best[c] = e.x;
        note_rt<W>(s, g, k, e.target, lane);
        note_bt<W>(s, g, e.target, lane);
        update_tad<Fwd>(tad, k, e.x);
      }
    } else {
      // update_transfers / update_footpaths for a destination target
      if (p.bounds_last_k == 0U && D::better(e.x, s.best[c])) {
        s.round_times[rti] = D::best_of(e.x, s.round_times[rti]);
        note_rt<W>(s, g, k, e.target, lane);
      }
      if (D::better(e.x, s.best[c]) && D::better_eq(e.x, tad[k])) {
        int const lbv = s.lb[c];
        if (lbv == kUnreachableLb || !D::better_eq(e.x + D::dir(lbv), tad[k])) continue;
        if (!within_bounds<Fwd, W>(im, s, slot, k, e.target, e.x)) continue;
        s.round_times[rti] = e.x;
        s.best[c] = e.x;
        note_rt<W>(s, g, k, e.target, lane);  // best: through the station mark
        if constexpr (W == 64) {
          mark_wave64(s.mark_t, s.flag_t, g, e.target, 1ULL << lane, s.L);
        } else {
          mark_one<W>(s.mark_t, s.list_t, &s.counts[0], s.cap_t, &s.counts[4], s, g, e.target, lane);
        }
        update_tad<Fwd>(tad, k, e.x);
      }
    }
    if (tad[k] != before) {
      if (n_chg < kMaxChanges) {
        s.chg_pos[slot * kMaxChanges + n_chg] = e.pos;
        s.chg_tad[slot * kMaxChanges + n_chg] = tad[k];
        ++n_chg;
      } else {
        atomicOr(&s.counts[4], 4U);
      }
    }
  }
  s.n_chg[slot] = n_chg;
}

// One relaxation (update_transfers raptor.h:699-740 / update_footpaths
// raptor.h:785-833); relaxations into the query's destination stations are
// k_dest_pass's and do nothing here.  Returns kRelaxImproved when it
// improved best (the station becomes marked), kRelaxHazard when it was a
// pruned pre-write after a time_at_dest change (exact replay needed if the
// station also has a passing relaxation), else 0.
constexpr int kRelaxImproved = 1;
constexpr int kRelaxHazard = 2;

template <bool Fwd, int W>
__device__ inline int relax(dev_image const& im, dev_state const& s, std::uint32_t g,
                            std::uint32_t lane, std::uint32_t slot, std::uint32_t k,
                            std::uint32_t t, int x, std::uint32_t pos, std::int16_t tad0) {
  using D = dirop<Fwd>;
  auto const c = lay<W>::col(s, g, t, lane);
  auto const xv = clamp16(x);
  if (!D::better(xv, s.best[c])) return 0;
  // destination stations are relaxed by k_dest_pass (tested only now: a
  // relaxation that does not beat best changes nothing either way)
  if (lay<W>::test(s.dest_mask, s, g, t, lane)) return 0;
  auto const tad_e = tad_at(s, slot, pos);
  int const lbv = s.lb[c];
  bool const elig = D::better_eq(xv, tad_e) && lbv != kUnreachableLb &&
                    D::better_eq(xv + D::dir(lbv), tad_e) &&
                    within_bounds<Fwd, W>(im, s, slot, k, t, xv);
  if (elig) {
    auto const old = atomic_best16<Fwd>(&s.best[c], xv);
    return D::better(xv, old) ? kRelaxImproved : 0;
  }
  // no pre-write while bounds are set (raptor.h:709, 790)
  if (W == 1 && s.param[slot].bounds_last_k != 0U) return 0;
  auto const old = atomic_best16<Fwd>(&s.round_times[lay<W>::rt(s, g, k, t, lane)], xv);  // pre-write
  if (D::better(xv, old)) note_rt<W>(s, g, k, t, lane);
  return tad_e != tad0 ? kRelaxHazard : 0;
}

// transfer_pack (batched engine): the (station, group) pairs flagged by the
// route phase, compacted station by station, become packed wavefronts that
// each hold up to 64 queries (group, lane) marked at one station; every
// lane then relaxes that station's transfer and footpaths for its own query
// (update_transfers / update_footpaths, raptor.h:680-845).  Packing as in
// k_route_pack; the packed arrays are free once the route scan is done.
__global__ void k_transfer_pack(round_args a) {
  auto const& s = a.st;
  __shared__ std::uint32_t fills_all[kWavesPerBlock][64];
  auto* fills = fills_all[threadIdx.x / 64U];
  auto const lane = threadIdx.x & 63U;
  auto const n_items = clamp_count(s.counts[1], s.cap_r);
  auto const wave = (blockIdx.x * blockDim.x + threadIdx.x) / 64U;
  for (std::uint32_t w0 = wave * 64U; w0 < n_items; w0 += gridDim.x * kWavesPerBlock * 64U) {
    auto const item = w0 + lane;
    std::uint32_t l = ~0U, g = 0U;
    std::uint64_t mask = 0ULL;
    if (item < n_items) {
      auto const e = s.list_r[item];
      l = static_cast<std::uint32_t>(e >> 32);
      g = static_cast<std::uint32_t>(e);
      // the group's 64 route-phase mark bytes (0 or 1) -> a lane mask: for
      // bytes b_0..b_7 of a word, (w * 0x0102040810204080) >> 56 has bit k = b_k
      auto const* mb = reinterpret_cast<std::uint64_t const*>(s.mark_rb + (static_cast<std::size_t>(g) * s.L + l) * 64U);
#pragma unroll
      for (std::uint32_t j = 0U; j < 8U; ++j) {
        mask |= ((mb[j] * 0x0102040810204080ULL) >> 56) << (8U * j);
      }
      mask &= s.grp_active[g];
    }
    auto const c = static_cast<std::uint32_t>(__popcll(mask));
    auto const max_entries = a.tune.transfer_pack_max_entries;
    std::uint32_t my_w = 0U, my_o = 0U, n_w = 0U, fill = 0U, cur_l = ~0U, entries = 0U;
    for (std::uint32_t t = 0U; t < 64U; ++t) {
      auto const ct = static_cast<std::uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(c), static_cast<int>(t)));
      if (ct == 0U) continue;
      auto const lt = static_cast<std::uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(l), static_cast<int>(t)));
      if (n_w == 0U || lt != cur_l || fill + ct > 64U || entries == max_entries) {
        if (n_w != 0U && lane == 0U) fills[n_w - 1U] = fill;
        ++n_w;
        fill = 0U;
        entries = 0U;
        cur_l = lt;
      }
      ++entries;
      if (lane == t) {
        my_w = n_w - 1U;
        my_o = fill;
      }
      fill += ct;
    }
    if (n_w != 0U && lane == 0U) fills[n_w - 1U] = fill;
    wave_sync_lds();
    std::uint32_t base = 0U;
    if (lane == 0U && n_w != 0U) base = atomicAdd(&s.counts[5], n_w);
    base = static_cast<std::uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(base)));
    if (c != 0U) {
      auto const pw = static_cast<std::size_t>(base) + my_w;
      if (pw < s.cap_pack) {
        if (my_o == 0U) s.packed_info[pw] = (static_cast<std::uint64_t>(l) << 32) | fills[my_w];
        auto* out = s.packed_lanes + pw * 64U + my_o;
        auto m = mask;
        for (std::uint32_t j = 0U; m != 0ULL; ++j) {
          auto const b = static_cast<std::uint32_t>(__ffsll(static_cast<long long>(m)) - 1);
          m &= m - 1ULL;
          out[j] = static_cast<std::uint16_t>((g << 6) | b);
        }
      } else {
        atomicOr(&s.counts[4], 1U);
      }
    }
    wave_sync_lds();  // fills is reused by the next window
  }
}

// update_transfers + update_footpaths (raptor.h:680-845), batched engine:
// one packed wavefront per station, one query (group, lane) per lane.
// Relaxations and their checks as in k_transfer_fp1; marks are set per lane.
template <bool Fwd>
__global__ void k_transfer_fp64(round_args a) {
  using D = dirop<Fwd>;
  auto const& im = a.im;
  auto const& s = a.st;
  auto const n_items = clamp_count(s.counts[5], s.cap_pack);
  auto const lane = threadIdx.x & 63U;
  item_claim claim{&s.counts[7], a.tune.transfer_claim};
  auto const* off = Fwd ? im.fp_out_offsets : im.fp_in_offsets;
  auto const* fps = Fwd ? im.fp_out : im.fp_in;
  unsigned long long c_st = 0ULL, c_fp = 0ULL, c_items = 0ULL;
  for (std::uint32_t item = claim.get(); item < n_items; item = claim.get()) {
    auto const info = s.packed_info[item];
    auto const i = uni(static_cast<std::uint32_t>(info >> 32));
    auto const n_lanes = uni(static_cast<std::uint32_t>(info));
    if (lane == 0U) ++c_items;
    std::uint32_t g = 0U, ln = 0U;
    bool on = false;
    if (lane < n_lanes) {
      auto const d = s.packed_lanes[static_cast<std::size_t>(item) * 64U + lane];
      g = static_cast<std::uint32_t>(d) >> 6;
      ln = static_cast<std::uint32_t>(d) & 63U;
      on = ((s.grp_active[g] >> ln) & 1ULL) != 0ULL;
    }
    auto const slot = g * 64U + ln;
    auto const& p = s.param[slot];
    auto const k = static_cast<std::uint32_t>(s.kslot[slot]);
    auto const tmp_i = on ? s.tmp[lay<64>::col(s, g, i, ln)] : static_cast<int>(D::kInvalid);
    bool const valid = on && tmp_i != D::kInvalid;
    auto const tad0 = s.tad0[slot];
    auto const fe0 = off[i], fe1 = off[i + 1U];
    if (on) {
      ++c_st;
      c_fp += fe1 - fe0;
    }
    auto const bit = 1ULL << ln;
    auto const emit = [&](std::uint32_t t, int code) {
      auto const gt = static_cast<std::size_t>(g) * s.L + t;
      if (code == kRelaxImproved) {
        atomicOr(reinterpret_cast<unsigned long long*>(&s.mark_t[gt]), static_cast<unsigned long long>(bit));
        s.flag_t[gt] = 1U;
      } else if (code == kRelaxHazard) {
        atomicOr(reinterpret_cast<unsigned long long*>(&s.hz_mask[gt]), static_cast<unsigned long long>(bit));
        s.flag_hz[gt] = 1U;
      }
    };
    // transfer at i (destination stations are handled by k_dest_pass)
    if (valid) {
      auto const tt = adjusted<std::uint8_t>(p, im.transfer_time[i]);
      emit(i, relax<Fwd, 64>(a.im, s, g, ln, slot, k, i, tmp_i + D::dir(tt), kPosTransfer | i, tad0));
    }
    // footpaths of i in list order
    for (auto fe = fe0; fe != fe1; ++fe) {
      auto const fp = fps[fe];
      auto const t = fp & ((1U << 23) - 1U);
      if (valid) {
        auto const dur = static_cast<std::int16_t>(fp >> 23);
        emit(t, relax<Fwd, 64>(a.im, s, g, ln, slot, k, t, tmp_i + D::dir(adjusted<std::int16_t>(p, dur)),
                               kPosFootpath | fe, tad0));
      }
    }
  }
  wave_count(&s.work[kWorkTransferStations], c_st);
  wave_count(&s.work[kWorkFootpaths], c_fp);
  wave_count(&s.work[kWorkTransferItems], c_items);
}

// update_transfers + update_footpaths (raptor.h:680-845), plain engine:
// one thread per (slot, station) marked by the route phase.
template <bool Fwd>
__global__ void k_transfer_fp1(round_args a) {
  using D = dirop<Fwd>;
  constexpr int W = 1;
  auto const& im = a.im;
  auto const& s = a.st;
  auto* overflow = &s.counts[4];
  auto const n_items = clamp_count(s.counts[1], s.cap_r);
  auto const* off = Fwd ? im.fp_out_offsets : im.fp_in_offsets;
  auto const* fps = Fwd ? im.fp_out : im.fp_in;
  unsigned long long c_st = 0ULL, c_fp = 0ULL;
  for (std::uint32_t item = blockIdx.x * blockDim.x + threadIdx.x; item < n_items;
       item += gridDim.x * blockDim.x) {
    auto const e = s.list_r[item];
    auto const g = static_cast<std::uint32_t>(e >> 32), i = static_cast<std::uint32_t>(e);
    if ((lay<W>::lanes_at(s.mark_r[lay<W>::lword(s, g, i)], i) & s.grp_active[g]) == 0ULL) continue;
    auto const slot = g;
    auto const& p = s.param[slot];
    auto const k = static_cast<std::uint32_t>(s.kslot[slot]);
    auto const tmp_i = s.tmp[lay<W>::col(s, g, i, 0U)];
    bool const valid = tmp_i != D::kInvalid;
    auto const tad0 = s.tad0[slot];
    ++c_st;
    c_fp += off[i + 1U] - off[i];
    auto const emit = [&](std::uint32_t t, int code) {
      if (code == kRelaxImproved) mark_one<W>(s.mark_t, s.list_t, &s.counts[0], s.cap_t, overflow, s, g, t, 0U);
      if (code == kRelaxHazard) mark_one<W>(s.hz_mask, s.list_hz, &s.counts[3], s.cap_hz, overflow, s, g, t, 0U);
    };
    // transfer at i (destination stations are handled by k_dest_pass)
    if (valid) {
      auto const tt = adjusted<std::uint8_t>(p, im.transfer_time[i]);
      emit(i, relax<Fwd, W>(a.im, s, g, 0U, slot, k, i, tmp_i + D::dir(tt), kPosTransfer | i, tad0));
    }
    // footpaths of i in list order
    for (auto fe = off[i]; fe != off[i + 1U]; ++fe) {
      auto const fp = fps[fe];
      auto const t = fp & ((1U << 23) - 1U);
      if (valid) {
        auto const dur = static_cast<std::int16_t>(fp >> 23);
        emit(t, relax<Fwd, W>(a.im, s, g, 0U, slot, k, t, tmp_i + D::dir(adjusted<std::int16_t>(p, dur)),
                              kPosFootpath | fe, tad0));
      }
    }
  }
  wave_count(&s.work[kWorkTransferStations], c_st);
  wave_count(&s.work[kWorkFootpaths], c_fp);
}

// round_times[k] = best for every station improved in the transfer phase
// (a passing relaxation overwrites round_times, raptor.h:734, 827)
template <bool Fwd, int W>
__global__ void k_rt_fixup(round_args a) {
  auto const& s = a.st;
  auto const n = clamp_count(s.counts[0], s.cap_t);
  // one thread per (group, station) item, over its improved lanes
  for (std::uint32_t h = blockIdx.x * blockDim.x + threadIdx.x; h < n; h += gridDim.x * blockDim.x) {
    auto const e = s.list_t[h];
    auto const g = static_cast<std::uint32_t>(e >> 32), t = static_cast<std::uint32_t>(e);
    std::uint64_t m;
    if constexpr (W == 64) {
      auto const gt = static_cast<std::size_t>(g) * s.L + t;
      m = s.mark_t[gt] & s.grp_active[g] & ~s.dest_mask[gt];
    } else {
      m = (lay<W>::test(s.mark_t, s, g, t, 0U) && (s.grp_active[g] & 1ULL) != 0ULL &&
           !lay<W>::test(s.dest_mask, s, g, t, 0U))
              ? 1ULL
              : 0ULL;
    }
    while (m != 0ULL) {
      auto const lane = static_cast<std::uint32_t>(__ffsll(static_cast<long long>(m)) - 1);
      m &= m - 1ULL;
      auto const k = static_cast<std::uint32_t>(s.kslot[g * W + lane]);
      s.round_times[lay<W>::rt(s, g, k, t, lane)] = s.best[lay<W>::col(s, g, t, lane)];
      if constexpr (W == 64) {
        // this thread owns the (g, t) items of this round: plain update
        s.rt_valid[(static_cast<std::size_t>(g) * s.K + k) * s.L + t] |= 1ULL << lane;
      }
    }
  }
}

// Exact replay for hazard stations: last passing relaxation P is the first
// eligible one with x == best; later pre-writes lower round_times below it.
// Replay of one hazard station t for one query (group g, lane): the last
// passing relaxation P is the first eligible one with x == best; the
// pre-writes after it lower round_times[k][t] below best.
template <bool Fwd, int W>
__device__ inline void replay_hazard(round_args const& a, std::uint32_t g, std::uint32_t t,
                                     std::uint32_t lane) {
  using D = dirop<Fwd>;
  auto const& im = a.im;
  auto const& s = a.st;
  auto const slot = g * W + lane;
  auto const& p = s.param[slot];
  auto const k = static_cast<std::uint32_t>(s.kslot[slot]);
  auto const c = lay<W>::col(s, g, t, lane);
  auto const bt = s.best[c];
  int const lbv = s.lb[c];
  // (a slot with bounds makes no pre-writes, so it has no hazards)
  auto const eligible = [&](std::int16_t x, std::uint32_t pos) {
    auto const tad_e = tad_at(s, slot, pos);
    return D::better_eq(x, tad_e) && lbv != kUnreachableLb && D::better_eq(x + D::dir(lbv), tad_e) &&
           within_bounds<Fwd, W>(im, s, slot, k, t, x);
  };
  // enumerate events into t: (pos, x)
  auto const* off = Fwd ? im.fp_in_offsets : im.fp_out_offsets;
  auto const* fps = Fwd ? im.fp_in : im.fp_out;
  auto const* map = Fwd ? im.fp_in_to_out : im.fp_out_to_in;
  auto const for_each_event = [&](auto&& fn) {
    if (marked_r<W>(s, g, t, lane)) {
      auto const tmp_t = s.tmp[c];
      if (tmp_t != D::kInvalid) {
        fn(kPosTransfer | t, clamp16(tmp_t + D::dir(adjusted<std::uint8_t>(p, im.transfer_time[t]))));
      }
    }
    for (auto fe = off[t]; fe != off[t + 1U]; ++fe) {
      auto const src = fps[fe] & ((1U << 23) - 1U);
      if (!marked_r<W>(s, g, src, lane)) continue;
      auto const tmp_s = s.tmp[lay<W>::col(s, g, src, lane)];
      if (tmp_s == D::kInvalid) continue;
      auto const dur = static_cast<std::int16_t>(fps[fe] >> 23);
      fn(kPosFootpath | map[fe], clamp16(tmp_s + D::dir(adjusted<std::int16_t>(p, dur))));
    }
  };
  std::uint32_t pos_p = 0xFFFFFFFFU;
  for_each_event([&](std::uint32_t pos, std::int16_t x) {
    if (x == bt && eligible(x, pos) && pos < pos_p) pos_p = pos;
  });
  if (pos_p == 0xFFFFFFFFU) return;
  auto v = bt;
  for_each_event([&](std::uint32_t pos, std::int16_t x) {
    if (pos > pos_p) v = D::best_of(x, v);
  });
  s.round_times[lay<W>::rt(s, g, k, t, lane)] = v;
}

// Exact replay for hazard stations (see replay_hazard).  Batched engine: one
// thread per (group, station) item, looping over its hazard lanes (about one
// per item); plain engine: one thread per item.
template <bool Fwd, int W>
__global__ void k_hazards(round_args a) {
  auto const& s = a.st;
  auto const n = clamp_count(s.counts[3], s.cap_hz);
  unsigned long long c_items = 0ULL, c_lanes = 0ULL;
  for (std::uint32_t h = blockIdx.x * blockDim.x + threadIdx.x; h < n; h += gridDim.x * blockDim.x) {
    auto const e = s.list_hz[h];
    auto const g = static_cast<std::uint32_t>(e >> 32), t = static_cast<std::uint32_t>(e);
    ++c_items;
    std::uint64_t m;
    if constexpr (W == 64) {
      // lanes with a hazard, active, with a passing relaxation, not a destination
      auto const gt = static_cast<std::size_t>(g) * s.L + t;
      m = s.hz_mask[gt] & s.grp_active[g] & s.mark_t[gt] & ~s.dest_mask[gt];
    } else {
      m = (lay<W>::test(s.hz_mask, s, g, t, 0U) && (s.grp_active[g] & 1ULL) != 0ULL &&
           lay<W>::test(s.mark_t, s, g, t, 0U) && !lay<W>::test(s.dest_mask, s, g, t, 0U))
              ? 1ULL
              : 0ULL;
    }
    while (m != 0ULL) {
      auto const lane = static_cast<std::uint32_t>(__ffsll(static_cast<long long>(m)) - 1);
      m &= m - 1ULL;
      ++c_lanes;
      replay_hazard<Fwd, W>(a, g, t, lane);
    }
  }
  wave_count(&s.work[kWorkHazardItems], c_items);
  wave_count(&s.work[kWorkHazardLanes], c_lanes);
}

// ======================================================================
// epoch begin / end

// reset_arrivals / next_start_time (raptor.h:174-188) for the lanes of the
// reset masks.  Batched engine: only the labels the validity bitmaps list
// are reset, since every other element already holds the direction's
// invalid value, except for the lanes of reset_full_mask (first use of the
// slot in this direction), which are reset densely; the station marks left
// by an epoch that ended at its last round count as written best labels.
template <int W>
__global__ void k_lane_reset(epoch_args a) {
  auto const& s = a.st;
  auto const stride = gridDim.x * blockDim.x;
  auto const tid = blockIdx.x * blockDim.x + threadIdx.x;
  if constexpr (W == 64) {
    for (std::uint32_t gi = 0U; gi < a.n_reset_groups; ++gi) {  // uniform
      auto const g = a.reset_groups[gi];
      auto const full = a.reset_full_mask[g];
      auto const rtm = a.reset_rt_mask[g] & ~full, btm = a.reset_bt_mask[g] & ~full;
      if ((rtm | btm | full) == 0ULL) continue;
      auto const gl0 = static_cast<std::size_t>(g) * s.L;
      if (full != 0ULL) {
        // dense: every element of the lanes' columns
        auto const per_group = s.L * 64U;
        for (std::uint32_t i = tid; i < per_group; i += stride) {
          auto const l = i / 64U, lane = i % 64U;
          if (((full >> lane) & 1ULL) == 0ULL) continue;
          auto const c = lay<64>::col(s, g, l, lane);
          s.best[c] = a.invalid;
          s.tmp[c] = a.invalid;
          for (std::uint32_t k = 0; k < s.K; ++k) s.round_times[lay<64>::rt(s, g, k, l, lane)] = a.invalid;
        }
      }
      if ((btm | full) != 0ULL) {
        auto const clr = btm | full;
        for (std::uint32_t l = tid; l < s.L; l += stride) {
          auto const mk = s.mark_t[gl0 + l], v = s.bt_valid[gl0 + l];
          if ((mk & clr) != 0ULL) s.mark_t[gl0 + l] = mk & ~clr;
          if ((v & clr) != 0ULL) s.bt_valid[gl0 + l] = v & ~clr;
          auto x = (mk | v) & btm;
          while (x != 0ULL) {
            auto const lane = static_cast<std::uint32_t>(__ffsll(static_cast<long long>(x)) - 1);
            x &= x - 1ULL;
            auto const c = lay<64>::col(s, g, l, lane);
            s.best[c] = a.invalid;
            s.tmp[c] = a.invalid;
          }
        }
      }
      if ((rtm | full) != 0ULL) {
        auto const clr = rtm | full;
        auto const n = static_cast<std::size_t>(s.K) * s.L;
        auto* valid = s.rt_valid + gl0 * s.K;
        for (std::size_t i = tid; i < n; i += stride) {
          auto const v = valid[i];
          if ((v & clr) == 0ULL) continue;
          valid[i] = v & ~clr;
          auto x = v & rtm;
          auto const k = static_cast<std::uint32_t>(i / s.L), l = static_cast<std::uint32_t>(i % s.L);
          while (x != 0ULL) {
            auto const lane = static_cast<std::uint32_t>(__ffsll(static_cast<long long>(x)) - 1);
            x &= x - 1ULL;
            s.round_times[lay<64>::rt(s, g, k, l, lane)] = a.invalid;
          }
        }
      }
    }
    return;
  }
  // station marks left by an epoch that ended at its last round are cleared
  // for the restarting lanes (the words may hold running lanes' marks)
  auto const words = W == 64 ? s.L : s.LW;
  for (std::uint32_t gi = 0U; gi < a.n_reset_groups; ++gi) {  // uniform
    auto const g = a.reset_groups[gi];
    auto const m = W == 64 ? a.reset_bt_mask[g] : (a.reset_bt_mask[g] != 0ULL ? ~0ULL : 0ULL);
    if (m == 0ULL) continue;
    auto* mt = s.mark_t + static_cast<std::size_t>(g) * words;
    for (std::uint32_t i = tid; i < words; i += stride) {
      auto const w = mt[i];
      if ((w & m) != 0ULL) mt[i] = w & ~m;
    }
  }
  auto const per_group = s.L * static_cast<std::uint32_t>(W);
  for (std::uint32_t gi = 0U; gi < a.n_reset_groups; ++gi) {
    auto const g = a.reset_groups[gi];
    auto const rtm = a.reset_rt_mask[g], btm = a.reset_bt_mask[g];
    if ((rtm | btm) == 0ULL) continue;
    for (std::uint32_t i = tid; i < per_group; i += stride) {
      auto const l = i / static_cast<std::uint32_t>(W), lane = i % static_cast<std::uint32_t>(W);
      auto const bit = 1ULL << lane;
      if ((rtm & bit) != 0ULL) {
        for (std::uint32_t k = 0; k < s.K; ++k) {
          s.round_times[lay<W>::rt(s, g, k, l, lane)] = a.invalid;
        }
      }
      if ((btm & bit) != 0ULL) {
        auto const c = lay<W>::col(s, g, l, lane);
        s.best[c] = a.invalid;
        s.tmp[c] = a.invalid;
      }
    }
  }
}

// ---- device lower bounds (PHAST) ---------------------------------------------
constexpr std::uint32_t kLbCutoff = 7201U;  // dial buckets of dijkstra.h:52

// every selected bound = unreachable
template <int W>
__global__ void k_lb_fill(lb_phast_args a) {
  auto const per = static_cast<std::size_t>(a.L) * W;
  auto const total = per * a.n_sel;
  for (std::size_t i = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x; i < total;
       i += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    auto const k = static_cast<std::uint32_t>(i / per);
    auto const r = i % per;
    if (W == 64 && ((a.masks[k] >> (r % 64U)) & 1ULL) == 0ULL) continue;
    a.dist[static_cast<std::size_t>(a.sel[k]) * per + r] = kUnreachableLb;
  }
}

// the upward searches' labels
__global__ void k_lb_seed(lb_phast_args a) {
  for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n_seeds; i += gridDim.x * blockDim.x) {
    a.dist[a.seed_idx[i]] = a.seed_min[i];
  }
}

// one level of the downward sweep: bound(v) = min(bound(v), bound(u) + w)
// over v's edges from higher-ranked u; values of kLbCutoff or more are
// unreachable.  W = 64: one wavefront per (location, group), one query per
// lane, each edge one 128 B row load; W = 1: one thread per (location, row).
// `first` / `stride`: this thread's first item and the item stride, in
// wavefronts (W = 64) or threads (W = 1).
template <int W>
__device__ inline void lb_level_items(lb_phast_args const& a, std::uint32_t b, std::uint32_t e,
                                      std::uint32_t first, std::uint32_t stride) {
  auto const& sw = a.sw;
  auto const n_items = (e - b) * a.n_sel;
  if constexpr (W == 64) {
    auto const lane = threadIdx.x & 63U;
    for (std::uint32_t item = first; item < n_items; item += stride) {
      auto const p = b + item / a.n_sel;
      auto const k = item % a.n_sel;
      auto const v = uni(sw.order[p]);
      auto const e0 = uni(sw.edge_off[p]), e1 = uni(sw.edge_off[p + 1U]);
      auto* col = a.dist + static_cast<std::size_t>(uni(a.sel[k])) * a.L * 64U + lane;
      std::uint32_t d = col[static_cast<std::size_t>(v) * 64U];
#pragma unroll 4
      for (auto x = e0; x < e1; ++x) {
        auto const u = uni(sw.src[x]);
        d = min(d, static_cast<std::uint32_t>(col[static_cast<std::size_t>(u) * 64U]) + sw.w[x]);
      }
      if (((a.masks[k] >> lane) & 1ULL) != 0ULL) {
        col[static_cast<std::size_t>(v) * 64U] = static_cast<std::uint16_t>(d < kLbCutoff ? d : kUnreachableLb);
      }
    }
  } else {
    for (std::uint32_t item = first; item < n_items; item += stride) {
      auto const p = b + item / a.n_sel;
      auto const v = sw.order[p];
      auto* row = a.dist + static_cast<std::size_t>(a.sel[item % a.n_sel]) * a.L;
      std::uint32_t d = row[v];
      for (auto x = sw.edge_off[p]; x < sw.edge_off[p + 1U]; ++x) {
        d = min(d, static_cast<std::uint32_t>(row[sw.src[x]]) + sw.w[x]);
      }
      row[v] = static_cast<std::uint16_t>(d < kLbCutoff ? d : kUnreachableLb);
    }
  }
}

template <int W>
__global__ void k_lb_level(lb_phast_args a, std::uint32_t b, std::uint32_t e) {
  auto const t = blockIdx.x * blockDim.x + threadIdx.x;
  if constexpr (W == 64) {
    lb_level_items<64>(a, b, e, t / 64U, gridDim.x * (blockDim.x / 64U));
  } else {
    lb_level_items<1>(a, b, e, t, gridDim.x * blockDim.x);
  }
}

// consecutive small levels [l0, l1) of the sweep in one workgroup, a barrier
// between levels (a level's bounds are final before the next level reads
// them): saves a launch per level near the top of the hierarchy
template <int W>
__global__ void __launch_bounds__(1024) k_lb_levels(lb_phast_args a, std::uint32_t l0, std::uint32_t l1) {
  for (auto lv = l0; lv < l1; ++lv) {
    auto const b = a.sw.level_off[lv], e = a.sw.level_off[lv + 1U];
    if constexpr (W == 64) {
      lb_level_items<64>(a, b, e, threadIdx.x / 64U, blockDim.x / 64U);
    } else {
      lb_level_items<1>(a, b, e, threadIdx.x, blockDim.x);
    }
    __syncthreads();
  }
}

// get_fastest_direct, one wavefront per query (see kernels.h).  Every node
// holds one live label (minutes, insertion number); popping the smallest
// (minutes, insertion) pair replays the reference's queue order, stale
// entries included (a node improved after its push leaves the queue at its
// improved label; its older entry would be skipped there).  The lanes pick
// the smallest label; lane 0 relaxes the popped node's footpaths in list
// order, which fixes the insertion numbers.
__global__ void __launch_bounds__(64) k_fastest_direct(fd_args a) {
  constexpr std::uint32_t kNone = 32767U;  // get_or_create's default distance
  __shared__ std::uint32_t dist[kFdMaxNodes], ins[kFdMaxNodes];
  __shared__ std::uint8_t queued[kFdMaxNodes];
  auto const lane = threadIdx.x;
  for (std::uint32_t qi = blockIdx.x; qi < a.n; qi += gridDim.x) {
    auto const q = a.q[qi];
    auto const n = q.node_end - q.node_begin;
    auto const* nodes = a.nodes + q.node_begin;
    // the starts enter the queue in ascending location order
    std::uint32_t n_starts = 0U;
    for (std::uint32_t b = 0U; b < n; b += 64U) {
      auto const i = b + lane;
      bool const st = i < n && nodes[i].is_start != 0U;
      auto const m = __ballot(st);
      if (i < n) {
        dist[i] = st ? nodes[i].start : kNone;
        ins[i] = n_starts + static_cast<std::uint32_t>(__popcll(m & ((1ULL << lane) - 1ULL)));
        queued[i] = st ? 1U : 0U;
      }
      n_starts += static_cast<std::uint32_t>(__popcll(m));
    }
    __syncthreads();
    slot_param tts{};
    tts.tts_default = q.tts_default;
    tts.tts_min = q.tts_min;
    tts.tts_add = q.tts_add;
    tts.tts_factor = q.tts_factor;
    auto const* off = q.fwd != 0U ? a.im.fp_out_offsets : a.im.fp_in_offsets;
    auto const* fps = q.fwd != 0U ? a.im.fp_out : a.im.fp_in;
    std::uint32_t end_dist = kNone, next_ins = n_starts;
    while (true) {
      // smallest (minutes, insertion) of the queued nodes
      auto key = ~0ULL;
      for (auto i = lane; i < n; i += 64U) {
        if (queued[i] != 0U) {
          auto const k = (static_cast<unsigned long long>(dist[i]) << 42) |
                         (static_cast<unsigned long long>(ins[i]) << 10) | i;
          key = k < key ? k : key;
        }
      }
      for (int o = 32; o > 0; o >>= 1) {
        auto const other = static_cast<unsigned long long>(__shfl_xor(static_cast<long long>(key), o, 64));
        key = other < key ? other : key;
      }
      if (key == ~0ULL) break;
      auto const v = static_cast<std::uint32_t>(key & 1023ULL);
      auto const d = dist[v];
      auto const l = nodes[v].loc;
      __syncthreads();
      if (l == 1U) break;  // special_station::kEnd
      if (lane == 0U) {
        queued[v] = 0U;
        for (auto e = off[l]; e != off[l + 1U]; ++e) {
          auto const nd = d + adjusted<std::uint16_t>(tts, static_cast<std::uint16_t>(fps[e] >> 23));
          if (nd > 65535U) continue;
          auto const t = fps[e] & ((1U << 23) - 1U);
          // the target's node (a destination), by binary search
          std::uint32_t lo = 0U, hi = n;
          while (lo < hi) {
            auto const mid = (lo + hi) / 2U;
            if (nodes[mid].loc < t) {
              lo = mid + 1U;
            } else {
              hi = mid;
            }
          }
          if (lo == n || nodes[lo].loc != t || nodes[lo].is_dest == 0U) continue;
          if (nd < dist[lo] && nd < 65535U) {
            dist[lo] = nd;
            ins[lo] = next_ins++;
            queued[lo] = 1U;
          }
        }
        if (nodes[v].is_dest != 0U) {
          auto const nd = d + nodes[v].off;
          if (nd < 65535U && nd < end_dist) end_dist = nd;
        }
      }
      __syncthreads();
    }
    if (lane == 0U) {
      a.out[qi] = end_dist == kNone ? static_cast<std::int16_t>(kNone)
                                    : static_cast<std::int16_t>(ceil(end_dist * q.factor));
    }
    __syncthreads();  // the next query reuses the LDS arrays
  }
}

// PHAST rows [n_rows][L] into the lower-bound lanes of their slots
// (batched engine); row blockIdx.y
__global__ void k_lb_scatter64(dev_state s, std::uint16_t const* rows, std::uint32_t const* slots) {
  auto const r = blockIdx.y;
  auto const slot = slots[r];
  auto const* row = rows + static_cast<std::size_t>(r) * s.L;
  for (std::uint32_t l = blockIdx.x * blockDim.x + threadIdx.x; l < s.L; l += gridDim.x * blockDim.x) {
    s.lb[lay<64>::col(s, slot / 64U, l, slot % 64U)] = row[l];
  }
}

template <int W>
__global__ void k_lb_put(dev_state s, std::uint32_t slot, std::uint16_t const* row) {
  for (std::uint32_t l = blockIdx.x * blockDim.x + threadIdx.x; l < s.L; l += gridDim.x * blockDim.x) {
    s.lb[lay<W>::col(s, slot / W, l, slot % W)] = row[l];
  }
}

template <int W>
__global__ void k_dest_ops(epoch_args a) {
  auto const& s = a.st;
  for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n_dest_ops; i += gridDim.x * blockDim.x) {
    auto const op = a.dest_ops[i];
    auto const slot = static_cast<std::uint32_t>(op >> 32), l = static_cast<std::uint32_t>(op);
    auto const g = slot / W, lane = slot % W;
    auto const bit = lay<W>::lbit(l, lane);
    auto* w = reinterpret_cast<unsigned long long*>(&s.dest_mask[lay<W>::lword(s, g, l)]);
    if (a.dest_set[i] != 0U) {
      atomicOr(w, static_cast<unsigned long long>(bit));
    } else {
      atomicAnd(w, ~static_cast<unsigned long long>(bit));
    }
  }
}

template <bool Fwd, int W>
__global__ void k_slot_init(epoch_args a) {
  auto const& s = a.st;
  auto const slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= a.n_slots) return;
  auto const f = a.slot_flags[slot];
  if ((f & 1U) == 0U) return;
  auto* tad = s.tad + static_cast<std::size_t>(slot) * kRowsMax;
  if ((f & 2U) != 0U) {
    for (int i = 0; i < kRowsMax; ++i) tad[i] = a.invalid;  // reset_arrivals
  }
  auto const w = a.slot_worst[slot];
  for (int i = 0; i < kRowsMax; ++i) tad[i] = dirop<Fwd>::best_of(w, tad[i]);
  s.kslot[slot] = 0U;
  auto const g = slot / W, lane = slot % W;
  atomicOr(reinterpret_cast<unsigned long long*>(&s.grp_active[g]), 1ULL << lane);
}

// add_start (raptor.h:190-203)
template <bool Fwd, int W>
__global__ void k_add_starts(epoch_args a) {
  auto const& s = a.st;
  for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n_starts; i += gridDim.x * blockDim.x) {
    auto const slot = a.start_slot[i], l = a.start_loc[i];
    auto const d = a.start_delta[i];
    auto const g = slot / W, lane = slot % W;
    atomic_best16<Fwd>(&s.best[lay<W>::col(s, g, l, lane)], d);
    atomic_best16<Fwd>(&s.round_times[lay<W>::rt(s, g, 0U, l, lane)], d);
    note_rt<W>(s, g, 0U, l, lane);
    note_bt<W>(s, g, l, lane);
    // appended to the station work list of the next round unless a lane
    // already marked the word there (other slots may be mid-epoch)
    mark_one<W>(s.mark_t, s.list_t, &s.counts[0], s.cap_t, &s.counts[4], s, g, l, lane);
  }
}

// journey extraction input: round_times[k][dest bit] of the listed slots
template <int W>
__global__ void k_extract(dev_state s, std::uint32_t const* slots, std::uint32_t n_slots) {
  auto const total = static_cast<std::size_t>(n_slots) * kMaxDestBits * s.K;
  for (std::size_t ii = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x; ii < total;
       ii += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    auto const slot = slots[ii / (kMaxDestBits * s.K)];
    auto const rem = static_cast<std::uint32_t>(ii % (kMaxDestBits * s.K));
    auto const i = ii;  // rows of the listed slots, in list order
    auto const j = rem / s.K, k = rem % s.K;
    auto const& p = s.param[slot];
    auto const nbits = p.intermodal != 0U ? 1U : p.n_dest;
    if (j >= nbits) continue;
    auto const d = p.intermodal != 0U ? 1U : s.dest[slot * kMaxDestBits + j];
    s.dest_rows[i] = s.round_times[lay<W>::rt(s, slot / W, k, d, slot % W)];
  }
}

// copies one slot's state into contiguous rows (state download)
// fill_bounds (raptor.h:117-170, Vias = 0, no td footpaths) of the plain
// engine: bounds row k of slot dst is the best round time of slot src over
// rounds 0..k; Fwd is the direction of src's search.
template <bool Fwd>
__global__ void k_fill_bounds(dev_state s, std::uint32_t src, std::uint32_t dst, std::uint32_t n_rows) {
  for (std::uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < s.L; x += gridDim.x * blockDim.x) {
    auto v = s.round_times[lay<1>::rt(s, src, 0U, x, 0U)];
    s.bounds[lay<1>::rt(s, dst, 0U, x, 0U)] = v;
    for (std::uint32_t k = 1U; k < n_rows; ++k) {
      v = dirop<Fwd>::best_of(s.round_times[lay<1>::rt(s, src, k, x, 0U)], v);
      s.bounds[lay<1>::rt(s, dst, k, x, 0U)] = v;
    }
  }
}

template <int W>
__global__ void k_gather_slot(dev_state s, std::uint32_t slot, std::int16_t* rt, std::int16_t* best,
                              std::int16_t* tmp, std::uint16_t* lb) {
  auto const g = slot / W, lane = slot % W;
  auto const total = static_cast<std::size_t>(s.K + 3U) * s.L;
  for (std::size_t i = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x; i < total;
       i += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    auto const row = static_cast<std::uint32_t>(i / s.L), l = static_cast<std::uint32_t>(i % s.L);
    if (row < s.K) {
      rt[i] = s.round_times[lay<W>::rt(s, g, row, l, lane)];
    } else if (row == s.K) {
      best[l] = s.best[lay<W>::col(s, g, l, lane)];
    } else if (row == s.K + 1U) {
      tmp[l] = static_cast<std::int16_t>(s.tmp[lay<W>::col(s, g, l, lane)]);
    } else {
      lb[l] = s.lb[lay<W>::col(s, g, l, lane)];
    }
  }
}

int grid_for(std::size_t work, int per_block, int cap) {
  auto const b = (work + static_cast<std::size_t>(per_block) - 1U) / static_cast<std::size_t>(per_block);
  return static_cast<int>(b < 1U ? 1U : (b > static_cast<std::size_t>(cap) ? static_cast<std::size_t>(cap) : b));
}

void compact(std::uint8_t* flags, std::uint32_t n_minor, std::uint32_t n_major, std::uint64_t* list,
             std::uint32_t* count, std::size_t cap, dev_state const& s, hipStream_t st) {
  auto const n_per_group = n_minor;
  auto const n_words = flag_words(static_cast<std::size_t>(n_minor) * n_major);
  hipLaunchKernelGGL(k_compact, dim3(grid_for(n_words, kBlock * kCompactWords, 2048)), dim3(kBlock), 0, st, flags,
                     n_per_group, n_words, list, count, cap, &s.counts[4]);
}

// One round (raptor.h:218-387) as five segments of kernels.  The route scan
// (segment 1) and the transfer kernel (segment 3) are segments of their
// own, so that profiling can record events between segments.
constexpr int kRoundSegments = 5;

template <bool Fwd, int W>
void round_segment(int seg, round_args const& a, std::uint32_t const* sweep, bool has_sweep,
                   std::uint32_t n_slots, int grid_waves, hipStream_t st) {
  auto const slot_blocks = grid_for(n_slots, kBlock, 1 << 20);
  auto const item_blocks = W == 64 ? (grid_waves + kWavesPerBlock - 1) / kWavesPerBlock
                                   : (grid_waves * 64 + kBlock - 1) / kBlock;
  switch (seg) {
    case 0:
      hipLaunchKernelGGL((k_round_end_check<W>), dim3(slot_blocks), dim3(kBlock), 0, st, a, n_slots);
      if (has_sweep) {
        hipLaunchKernelGGL((k_best_sweep<Fwd, W>), dim3(2048), dim3(kBlock), 0, st, a, sweep);
      }
      hipLaunchKernelGGL((k_round_begin<Fwd, W>), dim3(slot_blocks), dim3(kBlock), 0, st, a, n_slots);
      hipLaunchKernelGGL((k_mark_routes<W>), dim3(item_blocks), dim3(kBlock), 0, st, a);
      if constexpr (W == 64) {
        compact(a.st.flag_route, a.st.G, a.st.R, a.st.list_route, &a.st.counts[2], a.st.cap_route, a.st, st);
        hipLaunchKernelGGL((k_route_pack<Fwd>), dim3(item_blocks), dim3(kBlock), 0, st, a);
      }
      break;
    case 1:
      hipLaunchKernelGGL((k_route_scan<Fwd, W>), dim3(item_blocks), dim3(kBlock), 0, st, a);
      break;
    case 2:
      hipLaunchKernelGGL((k_clear_marks<W>), dim3(1024), dim3(kBlock), 0, st, a.st, 0);
      hipLaunchKernelGGL((k_group_update), dim3(grid_for(n_slots, kBlock, 1 << 20)), dim3(kBlock), 0, st, a.st);
      // list_t, routes, packed, loose items and their cursor
      hipLaunchKernelGGL((k_reset_counts), dim3(1), dim3(64), 0, st, a.st, 1U | 4U | 32U | 256U | 512U);
      if constexpr (W == 64) {
        // station-major flags: list entries (l << 32 | g), then packed
        // transfer wavefronts (counts[5] was reset after the route scan)
        compact(a.st.flag_r, a.st.G, a.st.L, a.st.list_r, &a.st.counts[1], a.st.cap_r, a.st, st);
        hipLaunchKernelGGL(k_transfer_pack, dim3(item_blocks), dim3(kBlock), 0, st, a);
      }
      hipLaunchKernelGGL((k_dest_pass<Fwd, W>), dim3(grid_for(n_slots, 64, 1 << 20)), dim3(64), 0, st, a, n_slots);
      break;
    case 3:
      if constexpr (W == 64) {
        hipLaunchKernelGGL((k_transfer_fp64<Fwd>), dim3(item_blocks), dim3(kBlock), 0, st, a);
      } else {
        hipLaunchKernelGGL((k_transfer_fp1<Fwd>), dim3(item_blocks), dim3(kBlock), 0, st, a);
      }
      break;
    default:
      if constexpr (W == 64) {
        compact(a.st.flag_t, a.st.L, a.st.G, a.st.list_t, &a.st.counts[0], a.st.cap_t, a.st, st);
        compact(a.st.flag_hz, a.st.L, a.st.G, a.st.list_hz, &a.st.counts[3], a.st.cap_hz, a.st, st);
      }
      hipLaunchKernelGGL((k_rt_fixup<Fwd, W>), dim3(1024), dim3(kBlock), 0, st, a);
      hipLaunchKernelGGL((k_hazards<Fwd, W>), dim3(1024), dim3(kBlock), 0, st, a);
      hipLaunchKernelGGL((k_clear_marks<W>), dim3(1024), dim3(kBlock), 0, st, a.st, 1);
      hipLaunchKernelGGL((k_clear_marks<W>), dim3(256), dim3(kBlock), 0, st, a.st, 2);
      // list_r, hazards, packed wavefronts, work cursors of the route scan and transfers
      hipLaunchKernelGGL((k_reset_counts), dim3(1), dim3(64), 0, st, a.st, 2U | 8U | 32U | 64U | 128U);
      hipLaunchKernelGGL((k_round_finish<W>), dim3(slot_blocks), dim3(kBlock), 0, st, a, n_slots);
      break;
  }
}

void round_segment_any(bool fwd, std::uint32_t W, int seg, round_args const& a, std::uint32_t const* sweep,
                       bool has_sweep, std::uint32_t n_slots, int grid_waves, hipStream_t st) {
  if (W == 64U) {
    fwd ? round_segment<true, 64>(seg, a, sweep, has_sweep, n_slots, grid_waves, st)
        : round_segment<false, 64>(seg, a, sweep, has_sweep, n_slots, grid_waves, st);
  } else {
    fwd ? round_segment<true, 1>(seg, a, sweep, has_sweep, n_slots, grid_waves, st)
        : round_segment<false, 1>(seg, a, sweep, has_sweep, n_slots, grid_waves, st);
  }
}

// events around the route scan (0, 1) and the transfer kernel (2, 3)
void record_before(round_args const& a, int seg, hipStream_t st) {
  if (seg == 1 && a.ev[0] != nullptr) (void)hipEventRecord(a.ev[0], st);
  if (seg == 3 && a.ev[2] != nullptr) (void)hipEventRecord(a.ev[2], st);
}
void record_after(round_args const& a, int seg, hipStream_t st) {
  if (seg == 1 && a.ev[1] != nullptr) (void)hipEventRecord(a.ev[1], st);
  if (seg == 3 && a.ev[3] != nullptr) (void)hipEventRecord(a.ev[3], st);
}

template <bool Fwd, int W>
void epoch_begin_impl(epoch_args const& a, hipStream_t st) {
  if (a.n_reset_groups != 0U) {
    hipLaunchKernelGGL((k_lane_reset<W>), dim3(4096), dim3(kBlock), 0, st, a);
  }
  if (a.n_dest_ops != 0U) {
    hipLaunchKernelGGL((k_dest_ops<W>), dim3(grid_for(a.n_dest_ops, kBlock, 4096)), dim3(kBlock), 0, st, a);
  }
  hipLaunchKernelGGL((k_slot_init<Fwd, W>), dim3(grid_for(a.n_slots, kBlock, 1 << 20)), dim3(kBlock), 0, st, a);
  if (a.n_starts != 0U) {
    hipLaunchKernelGGL((k_add_starts<Fwd, W>), dim3(grid_for(a.n_starts, kBlock, 4096)), dim3(kBlock), 0, st, a);
  }
}


// ======================================================================
// Journey reconstruction (reconstruct.cc:162-877).  One wavefront per job;
// every search below runs 64 candidates at a time and takes the first hit
// in the reference's iteration order, so the result equals the sequential
// walk of get_transport -> get_route_transport -> find_start_in_prev_round.
template <bool Fwd, int W>
struct rc_walker {
  using D = dirop<Fwd>;
  dev_image const& im;
  dev_state const& st;
  slot_param const& p;
  std::uint32_t g, slane, tl;

  __device__ std::int16_t rt(std::uint32_t k, std::uint32_t l) const {
    return st.round_times[lay<W>::rt(st, g, k, l, slane)];
  }
  // unix_to_delta of a trip event: (traffic day - base day) * 1440 + event
  __device__ std::int16_t ev_delta(std::uint32_t r, std::uint32_t to, std::uint32_t i, bool arr,
                                   std::uint32_t day) const {
    auto const n_t = im.route_transport_from[r + 1U] - im.route_transport_from[r];
    auto const d = im.event_times[im.route_stop_time_from[r] + n_t * (2U * i - (arr ? 1U : 0U)) + to];
    auto const cnt = static_cast<int>(d & 31U) * 1440 + static_cast<int>(d >> 5);
    return clamp16((static_cast<int>(day) - static_cast<int>(p.base)) * 1440 + cnt);
  }

  // find_start_in_prev_round (reconstruct.cc:187-293): the stop nearest to
  // `finish` (before it forward, after it backward) where round k - 1 allows
  // boarding; -1 if none
  __device__ int find_start(std::uint32_t k, std::uint32_t r, std::uint32_t to, std::uint32_t day,
                            std::uint32_t finish) const {
    auto const s0 = im.route_stop_offsets[r];
    auto const n = im.route_stop_offsets[r + 1U] - s0;
    auto const n_cand = Fwd ? finish : n - 1U - finish;
    for (std::uint32_t c0 = 0U; c0 < n_cand; c0 += 64U) {
      auto const x = c0 + tl;
      bool ok = false;
      if (x < n_cand) {
        auto const j = Fwd ? finish - 1U - x : finish + 1U + x;
        auto const stp = im.route_stops[s0 + j];
        bool const allowed = ((stp >> (Fwd ? 28U : 29U)) & 1U) != 0U;
        if (allowed) {
          ok = D::better_eq(rt(k - 1U, stp & kLocMask), ev_delta(r, to, j, !Fwd, day));
        }
      }
      auto const m = __ballot(ok);
      if (m != 0ULL) {
        auto const x0 = c0 + static_cast<std::uint32_t>(__ffsll(static_cast<unsigned long long>(m)) - 1);
        return static_cast<int>(Fwd ? finish - 1U - x0 : finish + 1U + x0);
      }
    }
    return -1;
  }

  // get_transport + get_route_transport (reconstruct.cc:304-374, 376-521):
  // over the (route, stop) visits of l in the reference's order, the trips
  // whose event at the stop falls on (day, mam).  A stop's events are sorted
  // by minute of the day (FIFO modulo 1440, checked when the timetable is
  // built), so each visit's matching trips are one range found by binary
  // search; the lanes search 64 visits at a time, then the visits with
  // matches are tried in order, their trips in route order.
  __device__ bool transport(std::uint32_t k, std::uint32_t l, std::int16_t time, rc_step& out) const {
    auto const minutes = static_cast<int>(p.base) * 1440 + time;  // split_day_mam
    auto const day = minutes / 1440, mam = minutes % 1440;
    auto const v0 = im.lrs_offsets[l], v1 = im.lrs_offsets[l + 1U];
    for (auto c0 = v0; c0 < v1; c0 += 64U) {
      auto const v = c0 + tl;
      std::uint32_t r = 0U, i = 0U, lo = 0U, hi = 0U;
      if (v < v1 && mam >= 0) {
        r = im.lrs_route[v];
        i = im.lrs_stop[v];
        auto const s0 = im.route_stop_offsets[r];
        auto const n = im.route_stop_offsets[r + 1U] - s0;
        auto const stp = im.route_stops[s0 + i];
        bool const allowed = Fwd ? (i != 0U && ((stp >> 29) & 1U) != 0U)
                                 : (i != n - 1U && ((stp >> 28) & 1U) != 0U);
        if (allowed) {
          auto const n_t = im.route_transport_from[r + 1U] - im.route_transport_from[r];
          auto const* ev = im.event_times + im.route_stop_time_from[r] + n_t * (2U * i - (Fwd ? 1U : 0U));
          // lo = first trip with minute >= mam, hi = first with minute > mam
          std::uint32_t a = 0U, b = n_t;
          while (a < b) {
            auto const mid = (a + b) / 2U;
            if (static_cast<int>(ev[mid] >> 5) < mam) a = mid + 1U; else b = mid;
          }
          lo = a;
          b = n_t;
          while (a < b) {
            auto const mid = (a + b) / 2U;
            if (static_cast<int>(ev[mid] >> 5) <= mam) a = mid + 1U; else b = mid;
          }
          hi = a;
        }
      }
      auto m = __ballot(lo < hi);
      while (m != 0ULL) {
        auto const b = __ffsll(static_cast<unsigned long long>(m)) - 1;
        m &= m - 1ULL;
        auto const sr = static_cast<std::uint32_t>(__shfl(static_cast<int>(r), b, 64));
        auto const si = static_cast<std::uint32_t>(__shfl(static_cast<int>(i), b, 64));
        auto const slo = static_cast<std::uint32_t>(__shfl(static_cast<int>(lo), b, 64));
        auto const shi = static_cast<std::uint32_t>(__shfl(static_cast<int>(hi), b, 64));
        auto const t0 = im.route_transport_from[sr];
        auto const n_t = im.route_transport_from[sr + 1U] - t0;
        auto const blk = im.route_stop_time_from[sr] + n_t * (2U * si - (Fwd ? 1U : 0U));
        for (auto to = slo; to < shi; ++to) {
          auto const d = im.event_times[blk + to];
          auto const td = static_cast<std::uint16_t>(day - static_cast<int>(d & 31U));
          if (!bit_test(im.bitfields, im.transport_traffic_days[t0 + to], td)) continue;
          auto const j = find_start(k, sr, to, td, si);
          if (j >= 0) {
            out.transport = t0 + to;
            out.route = sr;
            out.day = td;
            out.stop_idx = static_cast<std::uint16_t>(j);
            out.finish_idx = static_cast<std::uint16_t>(si);
            return true;
          }
        }
      }
    }
    return false;
  }

  // check_fp (reconstruct.cc:523-649): a walk of `dur` minutes from `target`
  // to the round's location ends at curr
  __device__ bool walk(std::uint32_t k, std::int16_t curr, std::uint32_t target, std::int16_t dur,
                       rc_step& out) const {
    return transport(k, target, static_cast<std::int16_t>(curr - D::dir(dur)), out);
  }
};

template <bool Fwd, int W>
__global__ __launch_bounds__(64) void k_reconstruct(rc_args a) {
  auto const& im = a.im;
  auto const& st = a.st;
  auto const tl = threadIdx.x & 63U;
  for (auto job_i = blockIdx.x; job_i < a.n_jobs; job_i += gridDim.x) {
    auto const job = a.jobs[job_i];
    auto const& p = st.param[job.slot];
    rc_walker<Fwd, W> w{im, st, p, job.slot / W, job.slot % W, tl};
    auto& out = a.out[job_i];
    auto const top = static_cast<std::uint32_t>(job.transfers) + 1U;
    auto l = job.dest;
    std::int32_t status = 0;
    std::uint32_t n_steps = 0U;
    for (std::uint32_t i = 0U; i < top; ++i) {
      auto const k = top - i;
      auto const curr = w.rt(k, l);
      rc_step s{};
      std::int32_t walk_id = -1;
      bool found = false;
      if (job.intermodal != 0U && k == top) {
        // find_dest_leg (reconstruct.cc:651-679) per egress offset
        for (std::uint32_t o = 0U; o < job.n_off && !found; ++o) {
          auto const dur = a.off_dur[job.off_begin + o];
          if (dur >= 511) continue;  // footpath::kMaxDuration
          found = w.walk(k, curr, a.off_loc[job.off_begin + o], dur, s);
          walk_id = static_cast<std::int32_t>(o);
        }
      } else {
        // transfer at l (duration 0 into the destination), then footpaths
        auto const tr = static_cast<std::int16_t>(k == top ? 0 : adjusted<std::uint8_t>(p, im.transfer_time[l]));
        found = w.walk(k, curr, l, tr, s);
        auto const* off = Fwd ? im.fp_in_offsets : im.fp_out_offsets;
        auto const* fps = Fwd ? im.fp_in : im.fp_out;
        for (auto e = off[l]; e < off[l + 1U] && !found; ++e) {
          auto const fp = fps[e];
          auto const dur = adjusted<std::int16_t>(p, static_cast<std::int16_t>(fp >> 23));
          found = w.walk(k, curr, fp & ((1U << 23) - 1U), dur, s);
          walk_id = static_cast<std::int32_t>(e - off[l]);
        }
      }
      if (!found) {
        status = 1;
        break;
      }
      s.curr_time = curr;
      s.walk = walk_id;
      if (tl == 0U) out.steps[i] = s;
      ++n_steps;
      l = im.route_stops[im.route_stop_offsets[s.route] + s.stop_idx] & kLocMask;
    }
    if (tl == 0U) {
      out.status = status;
      out.n_steps = n_steps;
      out.start_loc = l;
      out.rt0 = w.rt(0U, l);
      out.pad = 0;
    }
  }
}

}  // namespace

void launch_round(bool fwd, std::uint32_t W, round_args const& a, std::uint32_t const* sweep,
                  bool has_sweep, std::uint32_t n_slots, int grid_waves, hipStream_t stream) {
  for (int seg = 0; seg != kRoundSegments; ++seg) {
    record_before(a, seg, stream);
    round_segment_any(fwd, W, seg, a, sweep, has_sweep, n_slots, grid_waves, stream);
    record_after(a, seg, stream);
  }
}

// Captured rounds.  A graph holds the kernels of one round for fixed
// arguments (direction, lanes, sweep on/off, grid, and the round_args
// bytes apart from the events); it is captured on first use and replayed
// while the arguments stay the same.  With events set, the round is held
// as one graph per segment and the events are recorded between them.
struct round_graphs {
  struct entry {
    bool fwd, has_sweep, segmented;
    std::uint32_t W, n_slots;
    int grid_waves;
    std::uint32_t const* sweep;
    round_args key;
    hipGraphExec_t exec[kRoundSegments];
    int n_exec;
  };
  std::vector<entry> entries;
  std::uint64_t captures = 0U;
  ~round_graphs() {
    for (auto& e : entries) {
      for (int i = 0; i != e.n_exec; ++i) (void)hipGraphExecDestroy(e.exec[i]);
    }
  }
};

round_graphs* round_graphs_create() { return new round_graphs{}; }
void round_graphs_destroy(round_graphs* g) { delete g; }
std::uint64_t round_graphs_captures(round_graphs const* g) { return g == nullptr ? 0U : g->captures; }

namespace {

hipError_t capture_segments(bool fwd, std::uint32_t W, int seg0, int seg1, round_args const& a,
                            std::uint32_t const* sweep, bool has_sweep, std::uint32_t n_slots,
                            int grid_waves, hipStream_t stream, hipGraphExec_t* out) {
  auto err = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
  if (err != hipSuccess) return err;
  for (int seg = seg0; seg != seg1; ++seg) {
    round_segment_any(fwd, W, seg, a, sweep, has_sweep, n_slots, grid_waves, stream);
  }
  auto const launch_err = hipGetLastError();
  hipGraph_t graph = nullptr;
  err = hipStreamEndCapture(stream, &graph);
  if (launch_err != hipSuccess) err = launch_err;
  if (err == hipSuccess) err = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
  if (graph != nullptr) (void)hipGraphDestroy(graph);
  return err;
}

}  // namespace

hipError_t launch_round_graph(round_graphs& g, bool fwd, std::uint32_t W, round_args const& a,
                              std::uint32_t const* sweep, bool has_sweep, std::uint32_t n_slots,
                              int grid_waves, hipStream_t stream) {
  bool const segmented = a.ev[0] != nullptr;
  round_args key = a;
  for (auto& ev : key.ev) ev = nullptr;
  round_graphs::entry* hit = nullptr;
  for (auto& e : g.entries) {
    if (e.fwd == fwd && e.has_sweep == has_sweep && e.segmented == segmented && e.W == W &&
        e.n_slots == n_slots && e.grid_waves == grid_waves && e.sweep == sweep) {
      if (std::memcmp(&e.key, &key, sizeof(round_args)) == 0) {
        hit = &e;
      } else {  // arguments changed: drop the stale graphs
        for (int i = 0; i != e.n_exec; ++i) (void)hipGraphExecDestroy(e.exec[i]);
        e.n_exec = 0;
        e.key = key;
        hit = &e;
      }
      break;
    }
  }
  if (hit == nullptr) {
    g.entries.push_back(round_graphs::entry{fwd, has_sweep, segmented, W, n_slots, grid_waves, sweep,
                                            key, {}, 0});
    hit = &g.entries.back();
  }
  if (hit->n_exec == 0) {
    ++g.captures;
    if (segmented) {
      for (int seg = 0; seg != kRoundSegments; ++seg) {
        auto const err = capture_segments(fwd, W, seg, seg + 1, a, sweep, has_sweep, n_slots, grid_waves,
                                          stream, &hit->exec[seg]);
        if (err != hipSuccess) return err;
        hit->n_exec = seg + 1;
      }
    } else {
      auto const err = capture_segments(fwd, W, 0, kRoundSegments, a, sweep, has_sweep, n_slots, grid_waves,
                                        stream, &hit->exec[0]);
      if (err != hipSuccess) return err;
      hit->n_exec = 1;
    }
  }
  for (int i = 0; i != hit->n_exec; ++i) {
    if (segmented) record_before(a, i, stream);
    auto const err = hipGraphLaunch(hit->exec[i], stream);
    if (err != hipSuccess) return err;
    if (segmented) record_after(a, i, stream);
  }
  return hipSuccess;
}

void launch_lb_phast(std::uint32_t W, lb_phast_args const& a, std::uint32_t const* level_off,
                     std::uint32_t first_level, std::uint32_t n_levels, hipStream_t stream) {
  if (a.n_sel == 0U) return;
  auto const fill = static_cast<std::size_t>(a.n_sel) * a.L * W;
  if (W == 64U) {
    hipLaunchKernelGGL((k_lb_fill<64>), dim3(grid_for(fill, kBlock, 8192)), dim3(kBlock), 0, stream, a);
  } else {
    hipLaunchKernelGGL((k_lb_fill<1>), dim3(grid_for(fill, kBlock, 8192)), dim3(kBlock), 0, stream, a);
  }
  if (a.n_seeds != 0U) {
    hipLaunchKernelGGL(k_lb_seed, dim3(grid_for(a.n_seeds, kBlock, 1024)), dim3(kBlock), 0, stream, a);
  }
  // levels of at most kSmall items (wavefronts for W = 64, threads for
  // W = 1) run in one 1024-thread workgroup, as many in a row as there are
  std::size_t const kSmall = W == 64U ? 64U : 4096U;
  for (std::uint32_t l = first_level; l < n_levels; ++l) {
    auto const b = level_off[l], e = level_off[l + 1U];
    if (b == e) continue;
    auto const items = static_cast<std::size_t>(e - b) * a.n_sel;
    if (items <= kSmall) {
      auto l1 = l + 1U;
      while (l1 != n_levels && static_cast<std::size_t>(level_off[l1 + 1U] - level_off[l1]) * a.n_sel <= kSmall) ++l1;
      if (W == 64U) {
        hipLaunchKernelGGL((k_lb_levels<64>), dim3(1), dim3(1024), 0, stream, a, l, l1);
      } else {
        hipLaunchKernelGGL((k_lb_levels<1>), dim3(1), dim3(1024), 0, stream, a, l, l1);
      }
      l = l1 - 1U;
      continue;
    }
    if (W == 64U) {
      hipLaunchKernelGGL((k_lb_level<64>), dim3(grid_for(items * 64U, kBlock, 8192)), dim3(kBlock), 0, stream,
                         a, b, e);
    } else {
      hipLaunchKernelGGL((k_lb_level<1>), dim3(grid_for(items, kBlock, 8192)), dim3(kBlock), 0, stream, a, b, e);
    }
  }
}

void launch_fastest_direct(fd_args const& a, hipStream_t stream) {
  if (a.n == 0U) return;
  hipLaunchKernelGGL(k_fastest_direct, dim3(std::min<std::uint32_t>(a.n, 4096U)), dim3(64), 0, stream, a);
}

void launch_lb_scatter64(dev_state const& st, std::uint16_t const* rows, std::uint32_t const* slots,
                         std::uint32_t n_rows, hipStream_t stream) {
  if (n_rows == 0U) return;
  hipLaunchKernelGGL(k_lb_scatter64, dim3(grid_for(st.L, kBlock, 64), n_rows), dim3(kBlock), 0, stream, st, rows,
                     slots);
}

void launch_lb_put(std::uint32_t W, dev_state const& st, std::uint32_t slot, std::uint16_t const* row,
                   hipStream_t stream) {
  auto const grid = dim3(grid_for(st.L, kBlock, 4096));
  if (W == 64U) {
    hipLaunchKernelGGL((k_lb_put<64>), grid, dim3(kBlock), 0, stream, st, slot, row);
  } else {
    hipLaunchKernelGGL((k_lb_put<1>), grid, dim3(kBlock), 0, stream, st, slot, row);
  }
}

void launch_epoch_begin(bool fwd, std::uint32_t W, epoch_args const& a, hipStream_t stream) {
  if (W == 64U) {
    fwd ? epoch_begin_impl<true, 64>(a, stream) : epoch_begin_impl<false, 64>(a, stream);
  } else {
    fwd ? epoch_begin_impl<true, 1>(a, stream) : epoch_begin_impl<false, 1>(a, stream);
  }
}

void launch_gather_slot(std::uint32_t W, dev_state const& st, std::uint32_t slot, std::int16_t* rt,
                        std::int16_t* best, std::int16_t* tmp, std::uint16_t* lb, hipStream_t stream) {
  if (W == 64U) {
    hipLaunchKernelGGL((k_gather_slot<64>), dim3(2048), dim3(kBlock), 0, stream, st, slot, rt, best, tmp, lb);
  } else {
    hipLaunchKernelGGL((k_gather_slot<1>), dim3(2048), dim3(kBlock), 0, stream, st, slot, rt, best, tmp, lb);
  }
}

void launch_extract(std::uint32_t W, dev_state const& st, std::uint32_t const* slots,
                    std::uint32_t n_slots, hipStream_t stream) {
  if (n_slots == 0U) return;
  auto const work = static_cast<std::size_t>(n_slots) * kMaxDestBits * st.K;
  if (W == 64U) {
    hipLaunchKernelGGL((k_extract<64>), dim3(grid_for(work, kBlock, 4096)), dim3(kBlock), 0, stream, st, slots, n_slots);
  } else {
    hipLaunchKernelGGL((k_extract<1>), dim3(grid_for(work, kBlock, 4096)), dim3(kBlock), 0, stream, st, slots, n_slots);
  }
}

void launch_fill_bounds(bool fwd, dev_state const& st, std::uint32_t src, std::uint32_t dst,
                        std::uint32_t n_rows, hipStream_t stream) {
  auto const grid = dim3(grid_for(st.L, kBlock, 4096));
  if (fwd) {
    hipLaunchKernelGGL((k_fill_bounds<true>), grid, dim3(kBlock), 0, stream, st, src, dst, n_rows);
  } else {
    hipLaunchKernelGGL((k_fill_bounds<false>), grid, dim3(kBlock), 0, stream, st, src, dst, n_rows);
  }
}

void launch_reconstruct(bool fwd, std::uint32_t W, rc_args const& a, hipStream_t stream) {
  if (a.n_jobs == 0U) return;
  auto const grid = dim3(std::min<std::uint32_t>(a.n_jobs, 8192U));
  if (W == 64U) {
    if (fwd) {
      hipLaunchKernelGGL((k_reconstruct<true, 64>), grid, dim3(64), 0, stream, a);
    } else {
      hipLaunchKernelGGL((k_reconstruct<false, 64>), grid, dim3(64), 0, stream, a);
    }
  } else {
    if (fwd) {
      hipLaunchKernelGGL((k_reconstruct<true, 1>), grid, dim3(64), 0, stream, a);
    } else {
      hipLaunchKernelGGL((k_reconstruct<false, 1>), grid, dim3(64), 0, stream, a);
    }
  }
}

}  // namespace nigiri_amd
