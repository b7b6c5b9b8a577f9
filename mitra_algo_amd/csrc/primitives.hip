// Device primitives: exclusive scans and the LDS-staged stable LSD radix sort
// used for the curvature-signature sort.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.hpp"
#include "session.hpp"

namespace mvs {
namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;  // 2048

// Block-wide exclusive scan of one value per thread; returns the total.
template <class T> __device__ T block_exclusive_scan(T v, T *lds, T &total) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  T x = v;
  for (int o = 1; o < kWave; o <<= 1) {
    T y = __shfl_up(x, o, kWave);
    if (lane >= o)
      x += y;
  }
  if (lane == kWave - 1)
    lds[wave] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    T run = 0;
    for (int w = 0; w < kScanThreads / kWave; ++w) {
      T t = lds[w];
      lds[w] = run;
      run += t;
    }
    lds[kScanThreads / kWave] = run;
  }
  __syncthreads();
  T out = x - v + lds[wave];
  total = lds[kScanThreads / kWave];
  __syncthreads();
  return out;
}

template <class TIn>
__global__ void __launch_bounds__(kScanThreads)
    k_scan_reduce(const TIn *in, uint64_t n, uint64_t *partial) {
  __shared__ uint64_t lds[kScanThreads / kWave + 1];
  uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
  uint64_t sum = 0;
  for (int i = 0; i < kScanItems; ++i)
    if (base + i < n)
      sum += static_cast<uint64_t>(in[base + i]);
  uint64_t total;
  (void)block_exclusive_scan<uint64_t>(sum, lds, total);
  if (threadIdx.x == 0)
    partial[blockIdx.x] = total;
}

// Writes the exclusive scan of one tile.  The tile's offset is
// partial[blockIdx.x] (kSelfPrefix false: already scanned totals; nullptr
// for a single tile) or, with kSelfPrefix, the sum of the raw tile totals
// partial[0 .. blockIdx.x), which the block adds up itself: two launches per
// scan for up to kSelfPrefixTiles tiles.
constexpr unsigned kSelfPrefixTiles = 8192;
template <class TIn, bool kSelfPrefix>
__global__ void __launch_bounds__(kScanThreads)
    k_scan_apply(const TIn *in, uint64_t n, const uint64_t *partial, uint64_t *out,
                 uint64_t *total_also) {
  __shared__ uint64_t lds[kScanThreads / kWave + 1];
  uint64_t tile_off = 0;
  if (kSelfPrefix) {
    uint64_t pre = 0;
    for (unsigned i = threadIdx.x; i < blockIdx.x; i += kScanThreads)
      pre += partial[i];
    uint64_t tot;
    (void)block_exclusive_scan<uint64_t>(pre, lds, tot);
    tile_off = tot;
  } else if (partial) {
    tile_off = partial[blockIdx.x];
  }
  uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
  uint64_t v[kScanItems];
  uint64_t sum = 0;
  for (int i = 0; i < kScanItems; ++i) {
    v[i] = base + i < n ? static_cast<uint64_t>(in[base + i]) : 0;
    sum += v[i];
  }
  uint64_t total;
  uint64_t run = block_exclusive_scan<uint64_t>(sum, lds, total) + tile_off;
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i < n)
      out[base + i] = run;
    run += v[i];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
    out[n] = tile_off + total;
    if (total_also)
      *total_also = tile_off + total;
  }
}

__global__ void k_zero_u64(uint64_t *p, uint64_t *q) {
  *p = 0;
  if (q)
    *q = 0;
}

// Exclusive scan: one launch for one tile, two (tile totals, then tiles that
// add their own prefix) up to kSelfPrefixTiles tiles, recursive beyond.
// `pool` holds the partial arrays of all levels (carved from one scratch
// slot, sized by scan_pool_words).
template <class TIn>
void exclusive_scan_level(Session *s, const TIn *in, uint64_t *out, uint64_t n, uint64_t *pool,
                          uint64_t *total_also) {
  if (n == 0) {
    MVS_LAUNCH(s, "scan_zero", k_zero_u64, 1, 1, 0, out, total_also);
    return;
  }
  unsigned nb = blocks_for(n, kScanTile);
  uint64_t *part = pool, *part_scan = pool + nb;
  if (nb == 1) {
    MVS_LAUNCH(s, "scan_apply", (k_scan_apply<TIn, false>), 1, kScanThreads, 0, in, n, nullptr,
               out, total_also);
    return;
  }
  MVS_LAUNCH(s, "scan_reduce", (k_scan_reduce<TIn>), nb, kScanThreads, 0, in, n, part);
  if (nb <= kSelfPrefixTiles) {
    MVS_LAUNCH(s, "scan_apply", (k_scan_apply<TIn, true>), nb, kScanThreads, 0, in, n, part, out,
               total_also);
    return;
  }
  exclusive_scan_level<uint64_t>(s, part, part_scan, nb, pool + 2 * nb + 1, nullptr);
  MVS_LAUNCH(s, "scan_apply", (k_scan_apply<TIn, false>), nb, kScanThreads, 0, in, n, part_scan,
             out, total_also);
}
uint64_t scan_pool_words(uint64_t n) {
  uint64_t words = 0;
  while (true) {
    uint64_t nb = (n + kScanTile - 1) / kScanTile;
    if (nb == 0)
      nb = 1;
    words += 2 * nb + 1;
    if (nb == 1)
      return words;
    n = nb;
  }
}
template <class TIn>
void exclusive_scan_impl(Session *s, const TIn *in, uint64_t *out, uint64_t n, int slot,
                         uint64_t *total_also) {
  uint64_t *pool = scratch_as<uint64_t>(s, slot, scan_pool_words(n));
  exclusive_scan_level<TIn>(s, in, out, n, pool, total_also);
}

// ------------------------------------------------------------- radix sort
constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;  // 2048 keys per tile
constexpr int kSortWaves = kSortThreads / kWave;      // 4
constexpr int kRadix = 256;

__global__ void __launch_bounds__(kSortThreads)
    k_radix_hist(const uint64_t *keys, uint64_t n, int shift, unsigned tiles, uint32_t *hist) {
  __shared__ uint32_t h[kRadix];
  h[threadIdx.x] = 0;
  __syncthreads();
  uint64_t base = static_cast<uint64_t>(blockIdx.x) * kSortTile;
  for (int i = 0; i < kSortItems; ++i) {
    uint64_t j = base + i * kSortThreads + threadIdx.x;
    if (j < n)
      atomicAdd(&h[(keys[j] >> shift) & 0xFF], 1u);
  }
  __syncthreads();
  // digit-major layout: hist[d * tiles + tile]
  hist[static_cast<uint64_t>(threadIdx.x) * tiles + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of one tile.  Wave w owns keys [w*512, (w+1)*512) of the
// tile, processed in 8 rounds of 64; ranks are stable within the tile, the
// tile is re-ordered in LDS and written out in digit runs (coalesced).
__global__ void __launch_bounds__(kSortThreads)
    k_radix_scatter(const uint64_t *keys_in, const uint32_t *vals_in, uint64_t n, int shift,
                    unsigned tiles, const uint64_t *offsets, uint64_t *keys_out,
                    uint32_t *vals_out) {
  __shared__ uint32_t wave_hist[kSortWaves][kRadix];
  __shared__ uint32_t tile_off[kRadix];
  __shared__ uint64_t s_keys[kSortTile];
  __shared__ uint32_t s_vals[kSortTile];
  __shared__ uint64_t g_off[kRadix];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  for (int w = 0; w < kSortWaves; ++w)
    wave_hist[w][threadIdx.x] = 0;
  g_off[threadIdx.x] = offsets[static_cast<uint64_t>(threadIdx.x) * tiles + blockIdx.x];
  __syncthreads();
  const uint64_t tile_base = static_cast<uint64_t>(blockIdx.x) * kSortTile;
  uint64_t k[kSortItems];
  uint32_t v[kSortItems];
  uint32_t rank[kSortItems];
  uint32_t dig[kSortItems];
  for (int r = 0; r < kSortItems; ++r) {
    uint64_t j = tile_base + wave * (kSortItems * kWave) + r * kWave + lane;
    bool in = j < n;
    k[r] = in ? keys_in[j] : 0;
    v[r] = in ? vals_in[j] : 0;
    uint32_t d = in ? static_cast<uint32_t>((k[r] >> shift) & 0xFF) : 0u;
    dig[r] = d;
    // lanes holding the same digit (only valid lanes participate)
    uint64_t match = ballot(in);
    for (int b = 0; b < 8; ++b) {
      uint64_t bb = ballot(((d >> b) & 1u) != 0);
      match &= ((d >> b) & 1u) ? bb : ~bb;
    }
    uint32_t base = wave_hist[wave][d];
    uint32_t below = static_cast<uint32_t>(lanes_below(match));
    rank[r] = base + below;
    // wave executes in lockstep: all reads above precede this update
    if (in && below == 0)
      wave_hist[wave][d] = base + static_cast<uint32_t>(__popcll(match));
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  // per-digit offsets: tile exclusive scan over digits, then waves
  {
    const int d = threadIdx.x;
    uint32_t tot = 0;
    uint32_t pw[kSortWaves];
    for (int w = 0; w < kSortWaves; ++w) {
      pw[w] = tot;
      tot += wave_hist[w][d];
    }
    __syncthreads();
    for (int w = 0; w < kSortWaves; ++w)
      wave_hist[w][d] = pw[w];
    // exclusive scan of tot over digits (256 threads = 4 waves)
    uint32_t x = tot;
    for (int o = 1; o < kWave; o <<= 1) {
      uint32_t y = __shfl_up(x, o, kWave);
      if (lane >= o)
        x += y;
    }
    __shared__ uint32_t wsum[kSortWaves];
    if (lane == kWave - 1)
      wsum[wave] = x;
    __syncthreads();
    uint32_t add = 0;
    for (int w = 0; w < wave; ++w)
      add += wsum[w];
    tile_off[d] = x - tot + add;
  }
  __syncthreads();
  for (int r = 0; r < kSortItems; ++r) {
    uint64_t j = tile_base + wave * (kSortItems * kWave) + r * kWave + lane;
    if (j < n) {
      uint32_t p = tile_off[dig[r]] + wave_hist[wave][dig[r]] + rank[r];
      s_keys[p] = k[r];
      s_vals[p] = v[r];
    }
  }
  __syncthreads();
  uint64_t count = n - tile_base < kSortTile ? n - tile_base : kSortTile;
  for (int i = threadIdx.x; i < static_cast<int>(count); i += kSortThreads) {
    uint64_t key = s_keys[i];
    uint32_t d = static_cast<uint32_t>((key >> shift) & 0xFF);
    uint64_t dst = g_off[d] + (i - tile_off[d]);
    keys_out[dst] = key;
    vals_out[dst] = s_vals[i];
  }
}

// ------------------------------------------------ single-pass digit passes
// Each LSD pass is one launch (after one launch that counts every pass's
// digits): a tile takes a ticket in launch order, ranks its keys in LDS,
// publishes its per-digit counts, and finds the counts of the tiles before
// it by looking back over their published states (decoupled look-back,
// Merrill & Garland 2016; one pass per digit as in Adinets & Merrill 2022).
// A state word holds a flag (2 bits) and a count (30 bits): kStAgg, the
// tile's own counts; kStPre, the inclusive prefix over tiles 0..t.  Tickets
// follow the order in which blocks start, so every tile a block waits on
// belongs to a block that is already running and publishes its counts
// before it looks back itself: every wait ends.  The spin is also bounded
// (kStSpins windows); past the bound the tile sets the error word and goes
// on with what it has, so a broken invariant shows as wrong order, never as
// a hang.
constexpr uint32_t kStAgg = 1u << 30, kStPre = 2u << 30, kStCount = (1u << 30) - 1;
constexpr int kStWindow = 16;           // predecessor states read together
constexpr uint32_t kStSpins = 1u << 22;  // windows read before giving up
constexpr int kMaxPasses = 8;
// Up to kStMaxTiles tiles of 2048 keys, or of 4096 keys up to kStMaxTiles
// of those, a pass is one launch.  Tiles that run together all publish
// before any finds a prefix, so the look-back front spreads over about
// sqrt(2 tiles / kStWindow) rounds of reads; past a few hundred tiles that
// costs as much as the separate count and scan launches it saves.
constexpr unsigned kStMaxTiles = 256;
constexpr unsigned kStMaxTiles2 = 512;

__device__ inline uint32_t st_load(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_store(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Digit counts of every pass in one read of the keys: ghist[p * 256 + d]
// (cleared before).
__global__ void __launch_bounds__(kSortThreads)
    k_os_hist(const uint64_t *keys, uint64_t n, int passes, uint32_t *ghist) {
  __shared__ uint32_t h[kMaxPasses][kRadix];
  for (int p = 0; p < passes; ++p)
    h[p][threadIdx.x] = 0;
  __syncthreads();
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kSortThreads + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * kSortThreads) {
    const uint64_t k = keys[i];
    for (int p = 0; p < passes; ++p)
      atomicAdd(&h[p][(k >> (8 * p)) & 0xFF], 1u);
  }
  __syncthreads();
  for (int p = 0; p < passes; ++p)
    if (h[p][threadIdx.x])
      atomicAdd(&ghist[p * kRadix + threadIdx.x], h[p][threadIdx.x]);
}

// One pass: ghist the pass's digit counts, ticket its tile counter, state
// its tiles' state words (all cleared before the first pass).
template <int kItems>
__global__ void __launch_bounds__(kSortThreads)
    k_os_scatter(const uint64_t *keys_in, const uint32_t *vals_in, uint64_t n, int shift,
                 const uint32_t *ghist, unsigned *ticket, uint32_t *state, unsigned *err,
                 uint64_t *keys_out, uint32_t *vals_out) {
  __shared__ uint32_t wave_hist[kSortWaves][kRadix];
  __shared__ uint32_t tile_off[kRadix];
  constexpr int kTile = kSortThreads * kItems;
  __shared__ uint64_t s_keys[kTile];
  __shared__ uint32_t s_vals[kTile];
  __shared__ uint64_t g_off[kRadix];
  __shared__ uint32_t wsum[kSortWaves];
  __shared__ uint32_t s_tile;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  if (threadIdx.x == 0)
    s_tile = atomicAdd(ticket, 1u);
  for (int w = 0; w < kSortWaves; ++w)
    wave_hist[w][threadIdx.x] = 0;
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint64_t tile_base = static_cast<uint64_t>(tile) * kTile;
  uint64_t k[kItems];
  uint32_t v[kItems];
  uint32_t rank[kItems];
  uint32_t dig[kItems];
  for (int r = 0; r < kItems; ++r) {
    uint64_t j = tile_base + wave * (kItems * kWave) + r * kWave + lane;
    bool in = j < n;
    k[r] = in ? keys_in[j] : 0;
    v[r] = in ? vals_in[j] : 0;
    uint32_t d = in ? static_cast<uint32_t>((k[r] >> shift) & 0xFF) : 0u;
    dig[r] = d;
    uint64_t match = ballot(in);
    for (int b = 0; b < 8; ++b) {
      uint64_t bb = ballot(((d >> b) & 1u) != 0);
      match &= ((d >> b) & 1u) ? bb : ~bb;
    }
    uint32_t base = wave_hist[wave][d];
    uint32_t below = static_cast<uint32_t>(lanes_below(match));
    rank[r] = base + below;
    if (in && below == 0)
      wave_hist[wave][d] = base + static_cast<uint32_t>(__popcll(match));
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  const int d = threadIdx.x;
  uint32_t tot = 0;
  {
    uint32_t pw[kSortWaves];
    for (int w = 0; w < kSortWaves; ++w) {
      pw[w] = tot;
      tot += wave_hist[w][d];
    }
    // this tile's count of digit d, published before anything else
    uint32_t *mine = state + static_cast<uint64_t>(tile) * kRadix + d;
    if (tile == 0)
      st_store(mine, kStPre | tot);
    else
      st_store(mine, kStAgg | tot);
    __syncthreads();
    for (int w = 0; w < kSortWaves; ++w)
      wave_hist[w][d] = pw[w];
  }
  // tile offsets of the digits (exclusive scan of tot over d) and the pass's
  // digit starts (exclusive scan of ghist over d)
  uint32_t x = tot, gx = ghist[d];
  const uint32_t gd = gx;
  for (int o = 1; o < kWave; o <<= 1) {
    uint32_t y = __shfl_up(x, o, kWave), gy = __shfl_up(gx, o, kWave);
    if (lane >= o)
      x += y, gx += gy;
  }
  __shared__ uint32_t gsum[kSortWaves];
  if (lane == kWave - 1)
    wsum[wave] = x, gsum[wave] = gx;
  // tiles before this one: the look-back over their states
  uint32_t excl = 0;
  if (tile > 0) {
    int64_t j = static_cast<int64_t>(tile) - 1;
    uint32_t windows = 0;
    bool done = false;
    while (!done) {
      uint32_t w[kStWindow];
#pragma unroll
      for (int q = 0; q < kStWindow; ++q)
        w[q] = j - q >= 0 ? st_load(state + static_cast<uint64_t>(j - q) * kRadix + d) : kStPre;
      int q = 0;
      for (; q < kStWindow; ++q) {
        const uint32_t flag = w[q] & ~kStCount;
        if (flag == 0)
          break;  // not yet published: read again from here
        excl += w[q] & kStCount;
        if (flag == kStPre) {
          done = true;
          break;
        }
      }
      j -= q;
      if (!done && ++windows >= kStSpins) {
        atomicOr(err, 1u);
        break;
      }
    }
    st_store(state + static_cast<uint64_t>(tile) * kRadix + d, kStPre | (excl + tot));
  }
  __syncthreads();
  {
    uint32_t add = 0, gadd = 0;
    for (int w = 0; w < wave; ++w)
      add += wsum[w], gadd += gsum[w];
    tile_off[d] = x - tot + add;
    g_off[d] = static_cast<uint64_t>(gx - gd + gadd) + excl;
  }
  __syncthreads();
  for (int r = 0; r < kItems; ++r) {
    uint64_t j = tile_base + wave * (kItems * kWave) + r * kWave + lane;
    if (j < n) {
      uint32_t p = tile_off[dig[r]] + wave_hist[wave][dig[r]] + rank[r];
      s_keys[p] = k[r];
      s_vals[p] = v[r];
    }
  }
  __syncthreads();
  uint64_t count = n - tile_base < kTile ? n - tile_base : kTile;
  for (int i = threadIdx.x; i < static_cast<int>(count); i += kSortThreads) {
    uint64_t key = s_keys[i];
    uint32_t dd = static_cast<uint32_t>((key >> shift) & 0xFF);
    uint64_t dst = g_off[dd] + (i - tile_off[dd]);
    keys_out[dst] = key;
    vals_out[dst] = s_vals[i];
  }
}

__global__ void k_zero_bytes(unsigned char *p, uint64_t bytes) {
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  uint64_t words = bytes / 4;
  for (uint64_t w = i; w < words; w += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    reinterpret_cast<uint32_t *>(p)[w] = 0u;
  if (i < bytes - 4 * words)
    p[4 * words + i] = 0;
}

}  // namespace

// Kernel fill instead of hipMemsetAsync: stays correct inside captured graphs.
void zero_device(Session *s, void *ptr, size_t bytes) {
  if (!bytes)
    return;
  if (reinterpret_cast<uintptr_t>(ptr) % 4 != 0)
    raise(MVS_EINVAL, "zero_device needs 4-byte alignment");
  unsigned blocks = blocks_for((bytes / 4) ? bytes / 4 : 1, 256);
  if (blocks > 4096)
    blocks = 4096;
  MVS_LAUNCH(s, "zero", k_zero_bytes, blocks, 256, 0, static_cast<unsigned char *>(ptr),
             static_cast<uint64_t>(bytes));
}

void exclusive_scan_u64(Session *s, const uint64_t *in, uint64_t *out, uint64_t n,
                        uint64_t *total_also) {
  exclusive_scan_impl<uint64_t>(s, in, out, n, s->bank ? kSlotScanB : kSlotScan, total_also);
}
void exclusive_scan_u32(Session *s, const uint32_t *in, uint64_t *out, uint64_t n,
                        uint64_t *total_also) {
  exclusive_scan_impl<uint32_t>(s, in, out, n, s->bank ? kSlotScanB : kSlotScan, total_also);
}

void radix_sort_pairs_u64(Session *s, const uint64_t *keys_in, const uint32_t *vals_in,
                          uint64_t *keys_out, uint32_t *vals_out, uint64_t n, int bits) {
  if (n == 0)
    return;
  unsigned tiles = blocks_for(n, kSortTile);
  const bool b = s->bank != 0;
  uint64_t *k2 = scratch_as<uint64_t>(s, b ? kSlotSortB0 : kSlotSort0, n);
  uint32_t *v2 = scratch_as<uint32_t>(s, b ? kSlotSortB1 : kSlotSort1, n);
  const int passes = (bits + 7) / 8;
  const unsigned tiles2 = blocks_for(n, 2 * kSortTile);
  const bool one = tiles <= kStMaxTiles, two = !one && tiles2 <= kStMaxTiles2;
  if (passes <= kMaxPasses && (one || two)) {
    // one launch clears the counts, tickets and states, one counts every
    // pass's digits, then one launch per pass
    const unsigned ntile = one ? tiles : tiles2;
    const size_t words = static_cast<size_t>(passes) * kRadix + passes + 1 +
                         static_cast<size_t>(passes) * ntile * kRadix;
    uint32_t *st = scratch_as<uint32_t>(s, b ? kSlotSortBState : kSlotSortState, words);
    zero_device(s, st, words * sizeof(uint32_t));
    uint32_t *ghist = st, *ticket = st + passes * kRadix, *err = ticket + passes,
             *state = err + 1;
    MVS_LAUNCH(s, "radix_hist", k_os_hist, std::min<unsigned>(tiles, 1024u), kSortThreads, 0,
               keys_in, n, passes, ghist);
    const uint64_t *ksrc = keys_in;
    const uint32_t *vsrc = vals_in;
    for (int p = 0; p < passes; ++p) {
      bool last_to_out = ((passes - 1 - p) % 2) == 0;
      uint64_t *kdst = last_to_out ? keys_out : k2;
      uint32_t *vdst = last_to_out ? vals_out : v2;
      uint32_t *pstate = state + static_cast<size_t>(p) * ntile * kRadix;
      if (one)
        MVS_LAUNCH(s, "radix_scatter", k_os_scatter<kSortItems>, ntile, kSortThreads, 0, ksrc,
                   vsrc, n, 8 * p, ghist + p * kRadix, ticket + p, pstate, err, kdst, vdst);
      else
        MVS_LAUNCH(s, "radix_scatter", k_os_scatter<2 * kSortItems>, ntile, kSortThreads, 0,
                   ksrc, vsrc, n, 8 * p, ghist + p * kRadix, ticket + p, pstate, err, kdst, vdst);
      ksrc = kdst;
      vsrc = vdst;
    }
    return;
  }
  uint32_t *hist = scratch_as<uint32_t>(s, b ? kSlotSortBHist : kSlotSortHist,
                                        static_cast<size_t>(tiles) * kRadix);
  uint64_t *off = scratch_as<uint64_t>(s, b ? kSlotSortB2 : kSlotSort2,
                                       static_cast<size_t>(tiles) * kRadix + 1);
  // ping-pong so that the last pass writes keys_out/vals_out
  const uint64_t *ksrc = keys_in;
  const uint32_t *vsrc = vals_in;
  for (int p = 0; p < passes; ++p) {
    bool last_to_out = ((passes - 1 - p) % 2) == 0;
    uint64_t *kdst = last_to_out ? keys_out : k2;
    uint32_t *vdst = last_to_out ? vals_out : v2;
    int shift = 8 * p;
    MVS_LAUNCH(s, "radix_hist", k_radix_hist, tiles, kSortThreads, 0, ksrc, n, shift, tiles,
               hist);
    exclusive_scan_u32(s, hist, off, static_cast<uint64_t>(tiles) * kRadix);
    MVS_LAUNCH(s, "radix_scatter", k_radix_scatter, tiles, kSortThreads, 0, ksrc, vsrc, n, shift,
               tiles, off, kdst, vdst);
    ksrc = kdst;
    vsrc = vdst;
  }
}

}  // namespace mvs
