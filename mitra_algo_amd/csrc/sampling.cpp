// Host random streams shared by the stages: mt19937_64 over the caller's
// state and area-weighted vertex sampling (src/core/sampling.cpp:5-24).
#include <algorithm>
#include <cmath>
#include <utility>
#include <vector>

#include "common.hpp"

namespace mvs {

// std::mt19937_64 ([rand.predef]) advanced on a state in libstdc++ layout.
uint64_t mt64_next(mvs_mt64 *g) {
  constexpr int N = 312, M = 156;
  if (g->p >= N) {
    const uint64_t upper = ~0ULL << 31, lower = ~upper, a = 0xB5026F5AA96619E9ULL;
    for (int k = 0; k < N; ++k) {
      uint64_t y = (g->x[k] & upper) | (g->x[(k + 1) % N] & lower);
      g->x[k] = g->x[(k + M) % N] ^ (y >> 1) ^ ((y & 1) ? a : 0);
    }
    g->p = 0;
  }
  uint64_t z = g->x[g->p++];
  z ^= (z >> 29) & 0x5555555555555555ULL;
  z ^= (z << 17) & 0x71D67FFFEDA60000ULL;
  z ^= (z << 37) & 0xFFF7EEE000000000ULL;
  z ^= z >> 43;
  return z;
}

}  // namespace mvs

using namespace mvs;

extern "C" {

void mvs_mt64_seed(mvs_mt64 *g, uint64_t seed) {
  g->x[0] = seed;
  for (uint64_t i = 1; i < 312; ++i)
    g->x[i] = 6364136223846793005ULL * (g->x[i - 1] ^ (g->x[i - 1] >> 62)) + i;
  g->p = 312;
}

uint64_t mvs_mt64_next(mvs_mt64 *g) { return mt64_next(g); }

int mvs_sample_vertices(uint64_t nv, const double *weights, double ratio, mvs_mt64 *rng,
                        int32_t *out, uint64_t *count) {
  return guarded([&] {
    if (!rng || !count || (nv && (!weights || !out)))
      raise(MVS_EINVAL, "null argument");
    // Weighted sampling without replacement with exponential keys, one draw
    // per positive-weight vertex in index order.
    std::vector<std::pair<double, int>> keys;
    keys.reserve(nv);
    for (uint64_t i = 0; i < nv; ++i)
      if (weights[i] > 0) {
        double u = (static_cast<double>(mt64_next(rng) >> 11) + 0.5) / 9007199254740992.0;
        keys.push_back({-std::log(u) / weights[i], static_cast<int>(i)});
      }
    std::sort(keys.begin(), keys.end());
    uint64_t target = std::min<uint64_t>(keys.size(),
                                         static_cast<uint64_t>(std::ceil(nv * ratio)));
    for (uint64_t i = 0; i < target; ++i)
      out[i] = keys[i].second;
    std::sort(out, out + target);
    *count = target;
  });
}

}  // extern "C"
