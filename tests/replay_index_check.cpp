// TEST INFRASTRUCTURE ONLY (host C++, built with AddressSanitizer and
// UndefinedBehaviorSanitizer by tests/test_replay_index.py).
//
// Replays, on the CPU, the index arithmetic of the batched mean-shift set-up
// that a captured batch graph runs with frozen capacities when the live
// per-mesh vote counts differ from the captured ones: the per-mesh point
// segments after the running-maximum clamp (k_monotone, batch.hip), the
// segment lookup (seg_of), the cell keys with the sentinel key for points
// past the last segment (k_cell_keys), the key width of the radix sort, the
// cell offsets (k_cell_offsets), the row ranges a query takes from them
// (k_ms_query_small / k_ms_query) and the stratified seed draw
// (k_pick_sorted, its modulo included), all in mitra_algo_amd/csrc/
// msbatch.hip.  Every buffer has the captured step's size and lives on the
// heap, so a read or write outside it is reported by AddressSanitizer.
//
// Usage: replay_index_check <seed> <trials> [old]
//   old: the segment lookup before the sentinel key, which reads the grid of
//   the segment past the last one for points outside every segment (the
//   cause of the replay fault); the check must then report it.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <stdexcept>
#include <vector>

namespace {

struct Grid {
  int dims[3];
  double lo[3], inv;
  int reach;
};

[[noreturn]] void fail(const char *what) {
  std::fprintf(stderr, "replay_index_check: %s\n", what);
  std::exit(3);
}

// seg_of: the last m with off[m] <= i (off: meshes + 1 starts)
uint32_t seg_of(const uint64_t *off, uint32_t meshes, uint64_t i) {
  uint32_t f = 0, c = meshes + 1;
  while (c) {
    const uint32_t h = c / 2, m = f + h;
    if (off[m] <= i) {
      f = m + 1;
      c -= h + 1;
    } else {
      c = h;
    }
  }
  return f - 1;
}

int cell_of(double v, double lo, double inv, int dim) {
  double c = (v - lo) * inv;
  c = c < 0 ? 0 : c;
  const double f = static_cast<double>(static_cast<long long>(c));
  return f >= dim ? dim - 1 : static_cast<int>(f);
}

// k_monotone: running maximum, capped
void monotone(std::vector<uint64_t> &off, uint64_t cap) {
  uint64_t carry = 0;
  for (auto &v : off) {
    v = std::min(v, cap);
    v = std::max(v, carry);
    carry = v;
  }
}

std::vector<uint64_t> exclusive(const std::vector<uint64_t> &c) {
  std::vector<uint64_t> o(c.size() + 1, 0);
  for (size_t k = 0; k < c.size(); ++k)
    o[k + 1] = o[k] + c[k];
  return o;
}

// One captured step (counts n_cap) replayed over live counts n_live.
void trial(std::mt19937_64 &g, bool old_lookup) {
  const uint32_t M = 1 + g() % 48;
  std::vector<uint64_t> n_cap(M), n_live(M);
  for (uint32_t m = 0; m < M; ++m) {
    n_cap[m] = g() % 3000;
    // shrunk, grown or equal live counts
    const int mode = static_cast<int>(g() % 3);
    n_live[m] = mode == 0 ? n_cap[m] / (1 + g() % 4) : (mode == 1 ? n_cap[m] + g() % 2000 : n_cap[m]);
  }
  const uint64_t nr = exclusive(n_cap)[M];  // frozen point count
  const uint64_t cells_cap = std::max<uint64_t>(2 * nr + 1024ULL * M, 1);
  int key_bits = 1;
  while (key_bits < 63 && (1ULL << key_bits) <= cells_cap)
    ++key_bits;
  // live segments, clamped as the captured step clamps them
  std::vector<uint64_t> seg = exclusive(n_live);
  monotone(seg, nr);
  // per-mesh grids over the live segment sizes: clamp(2n, 1024, 2^22) cells
  std::vector<Grid> *grids = new std::vector<Grid>(M);
  std::vector<uint64_t> ncell(M);
  for (uint32_t m = 0; m < M; ++m) {
    const uint64_t n = seg[m + 1] - seg[m];
    const uint64_t cap = n < (1ULL << 21) ? std::max<uint64_t>(2 * n, 1024) : (1ULL << 22);
    Grid &gr = (*grids)[m];
    gr.reach = 1 + static_cast<int>(g() % 2);
    uint64_t left = cap;
    for (int k = 0; k < 3; ++k) {
      const uint64_t d = 1 + g() % std::max<uint64_t>(1, k == 2 ? left : std::min<uint64_t>(left, 64));
      gr.dims[k] = static_cast<int>(std::min(d, left));
      left = std::max<uint64_t>(1, left / static_cast<uint64_t>(gr.dims[k]));
      gr.lo[k] = 0;
    }
    gr.inv = 1.0;
    ncell[m] = static_cast<uint64_t>(gr.dims[0]) * gr.dims[1] * gr.dims[2];
    if (ncell[m] > cap)
      fail("grid larger than its cell budget");
  }
  std::vector<uint64_t> cellbase = exclusive(ncell);
  const uint64_t total = cellbase[M];
  if (total > cells_cap)
    fail("live cells exceed the captured cell bound");
  // cell keys of the nr frozen points (coordinates anywhere, also outside
  // the grids: clamped to the edge cells)
  const uint64_t beyond = cells_cap;
  std::vector<uint64_t> keys(nr);
  std::uniform_real_distribution<double> coord(-20.0, 400.0);
  for (uint64_t i = 0; i < nr; ++i) {
    const uint32_t m = seg_of(seg.data(), M, i);
    if (m >= M && !old_lookup) {
      keys[i] = beyond;
      continue;
    }
    const Grid &gr = grids->data()[m];  // old lookup: m == M reads past the array
    const int x = cell_of(coord(g), gr.lo[0], gr.inv, gr.dims[0]);
    const int y = cell_of(coord(g), gr.lo[1], gr.inv, gr.dims[1]);
    const int z = cell_of(coord(g), gr.lo[2], gr.inv, gr.dims[2]);
    keys[i] = cellbase[m] + (static_cast<uint64_t>(z) * gr.dims[1] + y) * gr.dims[0] + x;
  }
  for (uint64_t i = 0; i < nr; ++i)
    if (keys[i] >> key_bits)
      fail("a cell key is wider than the radix sort's key bits");
  // the LSD radix sort orders by the low key_bits bits, stably
  std::vector<uint64_t> sorted(keys);
  std::stable_sort(sorted.begin(), sorted.end());
  // k_cell_offsets: offsets[c] = first sorted position with key >= c,
  // c in [0, total], into a buffer of cells_cap + 1 entries
  std::vector<uint64_t> *offsets = new std::vector<uint64_t>(cells_cap + 1);
  for (uint64_t c = 0; c <= total; ++c)
    offsets->data()[c] =
        static_cast<uint64_t>(std::lower_bound(sorted.begin(), sorted.end(), c) - sorted.begin());
  for (uint64_t c = 0; c < total; ++c)
    if (offsets->data()[c] > offsets->data()[c + 1])
      fail("cell offsets decrease");
  // query rows: every mesh's rows around a few cells, lengths non-negative
  // and inside the sorted records (plus the sentinel at [nr])
  for (uint32_t m = 0; m < M; ++m) {
    const Grid &gr = (*grids)[m];
    for (int probe = 0; probe < 4; ++probe) {
      const int cx = static_cast<int>(g() % gr.dims[0]), cy = static_cast<int>(g() % gr.dims[1]),
                cz = static_cast<int>(g() % gr.dims[2]);
      const int x0 = std::max(cx - gr.reach, 0), x1 = std::min(cx + gr.reach, gr.dims[0] - 1);
      for (int dz = -gr.reach; dz <= gr.reach; ++dz)
        for (int dy = -gr.reach; dy <= gr.reach; ++dy) {
          const int z = cz + dz, y = cy + dy;
          if (z < 0 || z >= gr.dims[2] || y < 0 || y >= gr.dims[1])
            continue;
          const uint64_t row = cellbase[m] + (static_cast<uint64_t>(z) * gr.dims[1] + y) * gr.dims[0];
          const uint64_t b = offsets->data()[row + x0], e = offsets->data()[row + x1 + 1];
          if (e < b || e > nr)
            fail("a query row range is negative or past the records");
        }
    }
  }
  // seeds: per mesh at most its live points (k_batch_seeds caps each kind's
  // seeds by its votes), starts clamped to the captured seed capacity
  const uint64_t seed_cap = 1 + nr / 5;
  std::vector<uint64_t> kseed(M);
  for (uint32_t m = 0; m < M; ++m)
    kseed[m] = std::min<uint64_t>(seg[m + 1] - seg[m], g() % 600);
  std::vector<uint64_t> soff = exclusive(kseed);
  monotone(soff, seed_cap);
  std::vector<uint32_t> *pos = new std::vector<uint32_t>(seed_cap);
  for (uint64_t j = 0; j < soff[M]; ++j) {
    const uint32_t m = seg_of(soff.data(), M, j);
    if (m >= M)
      fail("a drawn seed outside every mesh");
    const uint64_t jl = j - soff[m], km = soff[m + 1] - soff[m];
    const uint64_t base = seg[m], nm = seg[m + 1] - base;
    const uint64_t lo = jl * nm / km, hi = (jl + 1) * nm / km;
    if (hi <= lo)
      fail("an empty seed stratum (modulo by zero)");
    const uint64_t p = base + lo + g() % (hi - lo);
    if (p >= nr)
      fail("a seed past the sorted records");
    pos->data()[j] = static_cast<uint32_t>(p);
  }
  delete pos;
  delete offsets;
  delete grids;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: replay_index_check seed trials [old]\n");
    return 2;
  }
  std::mt19937_64 g(std::strtoull(argv[1], nullptr, 10));
  const int trials = std::atoi(argv[2]);
  const bool old_lookup = argc > 3 && std::strcmp(argv[3], "old") == 0;
  for (int t = 0; t < trials; ++t)
    trial(g, old_lookup);
  std::printf("ok %d trials\n", trials);
  return 0;
}
