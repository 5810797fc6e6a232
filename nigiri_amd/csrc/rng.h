// Portable seeded RNG: xoshiro256** seeded through splitmix64, with explicit
// bounded draws so that generated timetables and queries do not depend on
// the standard library's distributions.
#pragma once

#include <cstdint>

namespace nigiri_amd {

inline std::uint64_t splitmix64(std::uint64_t& x) {
  std::uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

struct rng {
  explicit rng(std::uint64_t seed) {
    for (auto& w : s_) {
      w = splitmix64(seed);
    }
  }
  static std::uint64_t rotl(std::uint64_t x, int k) {
    return (x << k) | (x >> (64 - k));
  }
  std::uint64_t next() {
    auto const result = rotl(s_[1] * 5U, 7) * 9U;
    auto const t = s_[1] << 17;
    s_[2] ^= s_[0];
    s_[3] ^= s_[1];
    s_[1] ^= s_[2];
    s_[0] ^= s_[3];
    s_[2] ^= t;
    s_[3] = rotl(s_[3], 45);
    return result;
  }
  // uniform integer in [lo, hi] (inclusive)
  std::int64_t uniform(std::int64_t lo, std::int64_t hi) {
    auto const span = static_cast<std::uint64_t>(hi - lo) + 1U;
    return lo + static_cast<std::int64_t>(next() % span);
  }
  // uniform double in [0, 1)
  double unit() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
  bool coin(double p) { return unit() < p; }
  std::uint64_t s_[4];
};

}  // namespace nigiri_amd
