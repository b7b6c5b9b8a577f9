// Mean-shift over transform votes.
//
// Sequential-seed mode (the reference's semantics, src/cpu/clustering.cpp:
// 32-208): one trajectory at a time, each run entirely inside one launch of
// k_trajectory (a single 1024-thread workgroup): per iteration the distances
// of the trajectory's candidate votes are evaluated in parallel, the
// in-support votes are compacted in vote order, and their weights and
// weighted descriptors are summed IN VOTE ORDER by one lane per component, so
// every sum rounds exactly as the reference's sequential loop does.  The
// candidates are the same-kind votes within 2h of an anchor mean, listed in
// vote order by a full pass over the votes and kept while the mean stays
// within h of the anchor: a superset of every support met meanwhile, so the
// pass over all votes runs about 1.6 times per trajectory instead of once
// per iteration.  The host keeps the reference's
// driver: seed = k-th unvisited vote (Fenwick tree instead of rebuilding the
// unvisited list), merge into the first same-kind centre within h^2/4,
// winners by largest support count, extraction, stable sort, truncation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "session.hpp"

namespace mvs {
namespace {

constexpr int kTrajThreads = 1024;
constexpr int kTrajWaves = kTrajThreads / kWave;
// entries per thread in one round of a scan: 4 x 1024 entries share one set
// of barriers, and each thread keeps 4 independent distance sums in flight
constexpr int kTrajPer = 4;
static_assert(kTrajPer * kTrajWaves == kWave, "one wave scans the round's counts");
// A trajectory's candidate list holds the same-kind votes within
// (1 + kAnchorRadius) h of the mean at which the list was built (its anchor);
// it is rebuilt by a full pass once the mean is kAnchorRadius h or farther
// from the anchor.  1.0 rebuilds about 1.6 times per trajectory.
constexpr double kAnchorRadius = 1.0;
// products staged per chunk of the ordered sums (13 x 512 doubles of LDS)
constexpr int kSumChunk = 512;
static_assert(kTrajThreads == 2 * kSumChunk, "two staging threads per chunk entry");

struct TrajState {
  double mean[12];
  int32_t converged, exhausted, iterations, touched;
};

// Slots of the selected entries of one scan round (entry q * kTrajThreads +
// tid of the round) in entry order, after the *count entries selected before
// the round; *count then includes the round.  Every thread of the workgroup
// calls it.
__device__ inline void ordered_round(const bool (&in)[kTrajPer], uint32_t (&pos)[kTrajPer],
                                     uint32_t *wsum, uint32_t *count) {
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  uint64_t mask[kTrajPer];
#pragma unroll
  for (int q = 0; q < kTrajPer; ++q) {
    mask[q] = ballot(in[q]);
    if (lane == 0)
      wsum[q * kTrajWaves + wave] = static_cast<uint32_t>(__popcll(mask[q]));
  }
  __syncthreads();
  if (wave == 0) {
    const uint32_t c = wsum[lane];
    uint32_t incl = c;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const uint32_t t = __shfl_up(incl, d);
      if (lane >= d)
        incl += t;
    }
    const uint32_t before = *count;
    wsum[lane] = before + incl - c;
    if (lane == kWave - 1)
      *count = before + incl;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kTrajPer; ++q)
    pos[q] = wsum[q * kTrajWaves + wave] + static_cast<uint32_t>(lanes_below(mask[q]));
  __syncthreads();
}

// Squared distance of vote j from m over the first dim components, summed
// in component order as the reference does.
__device__ inline double vote_d2(const double *desc, uint64_t V, uint64_t j, const double *m,
                                 int dim) {
  double s = 0;
  for (int k = 0; k < dim; ++k) {
    const double d = m[k] - desc[static_cast<uint64_t>(k) * V + j];
    s += d * d;
  }
  return s;
}

// support[] must be zero on entry for every vote; on exit the touched votes
// are listed in out_ids/out_counts and their support entries reset.  out,
// out_ids and out_counts are pinned host memory: the host reads them after
// one stream synchronisation, with no copies.
// Exactness of the candidate list: distances carry a relative rounding error
// below 2e-15, so with reach2 = (1 + r)^2 h^2 (1 + 1e-6) and keep2 =
// r^2 h^2 (1 - 1e-6) every vote whose computed distance from the mean is
// below h while the list is kept lies in the list (triangle inequality), and
// the in-support set, its vote order and every sum are those of a full pass.
__global__ void __launch_bounds__(kTrajThreads)
    k_trajectory(const double *desc, const uint8_t *kind, uint64_t V, int32_t seed, double h2,
                 double reach2, double keep2, int max_iterations, uint32_t *support,
                 int32_t *cand, int32_t *list_j, double *list_w, int32_t *touched_ids,
                 int32_t *out_ids, uint32_t *out_counts, TrajState *out) {
  __shared__ double mean[12], anchor[12];
  __shared__ double sums[13];
  __shared__ double stage[13][kSumChunk];
  __shared__ uint32_t wsum[kWave];
  __shared__ uint32_t n_in, n_cand, n_touched;
  __shared__ int flag_exhausted, flag_converged, iters, build;
  const int tid = threadIdx.x;
  const bool refl = kind[seed] != 0;
  const int dim = refl ? 4 : 12;
  if (tid < 12)
    mean[tid] = desc[static_cast<uint64_t>(tid) * V + seed];
  if (tid == 0) {
    n_touched = 0;
    flag_exhausted = flag_converged = 0;
    iters = 0;
  }
  bool anchored = false;  // meaningful in thread 0 only
  __syncthreads();
  for (int it = 0; it < max_iterations; ++it) {
    if (tid == 0) {
      n_in = 0;
      bool b = !anchored;
      if (anchored) {
        double drift = 0;
        for (int k = 0; k < dim; ++k) {
          const double d = mean[k] - anchor[k];
          drift += d * d;
        }
        b = !(drift <= keep2);
      }
      build = b;
      if (b)
        n_cand = 0;
      anchored = true;
    }
    __syncthreads();
    const bool rebuild = build;
    double m[12];
    for (int k = 0; k < 12; ++k)
      m[k] = mean[k];
    if (rebuild) {
      // candidate list: the same-kind votes within reach of the mean, in vote order
      for (uint64_t base = 0; base < V; base += kTrajPer * kTrajThreads) {
        bool in[kTrajPer];
        uint32_t pos[kTrajPer];
#pragma unroll
        for (int q = 0; q < kTrajPer; ++q) {
          const uint64_t j = base + static_cast<uint64_t>(q * kTrajThreads + tid);
          in[q] = j < V && (kind[j] != 0) == refl && vote_d2(desc, V, j, m, dim) < reach2;
        }
        ordered_round(in, pos, wsum, &n_cand);
#pragma unroll
        for (int q = 0; q < kTrajPer; ++q)
          if (in[q])
            cand[pos[q]] = static_cast<int32_t>(base + static_cast<uint64_t>(q * kTrajThreads + tid));
      }
      if (tid < 12)
        anchor[tid] = m[tid];
      __threadfence_block();
      __syncthreads();
    }
    // in-support votes among the candidates, compacted in vote order
    const uint32_t C = n_cand;
    for (uint32_t base = 0; base < C; base += kTrajPer * kTrajThreads) {
      bool in[kTrajPer];
      uint32_t pos[kTrajPer];
      int32_t jj[kTrajPer];
      double w[kTrajPer];
#pragma unroll
      for (int q = 0; q < kTrajPer; ++q) {
        const uint32_t e = base + static_cast<uint32_t>(q * kTrajThreads + tid);
        in[q] = false;
        jj[q] = 0;
        w[q] = 0;
        if (e < C) {
          jj[q] = cand[e];
          const double s = vote_d2(desc, V, static_cast<uint64_t>(jj[q]), m, dim);
          if (s < h2) {
            in[q] = true;
            w[q] = 1 - s / h2;
          }
        }
      }
      ordered_round(in, pos, wsum, &n_in);
#pragma unroll
      for (int q = 0; q < kTrajPer; ++q)
        if (in[q]) {
          list_j[pos[q]] = jj[q];
          list_w[pos[q]] = w[q];
          if (support[jj[q]]++ == 0) {
            const uint32_t t = atomicAdd(&n_touched, 1u);
            touched_ids[t] = jj[q];
          }
        }
    }
    // ordered sums: lane 0 the weights, lane k the k-th weighted components.
    // The whole workgroup stages a chunk of products in LDS, then each of
    // the dim + 1 lanes adds its row in vote order, so the sequential chains
    // wait on LDS reads only.
    __threadfence_block();
    __syncthreads();
    const uint32_t N = n_in;
    const int cols = dim + 1;
    double acc = 0;
    for (uint32_t c0 = 0; c0 < N; c0 += kSumChunk) {
      const uint32_t n = N - c0 < kSumChunk ? N - c0 : kSumChunk;
      {
        // two threads per entry, each with its entry's vote and weight and
        // every other component, so all of a thread's gathers are in flight
        const int e = tid % kSumChunk, half = tid / kSumChunk;
        if (static_cast<uint32_t>(e) < n) {
          const uint64_t j = static_cast<uint64_t>(list_j[c0 + e]);
          const double w = list_w[c0 + e];
          if (half == 0)
            stage[0][e] = w;
#pragma unroll
          for (int u = 0; u < 6; ++u) {
            const int k = 1 + half + 2 * u;
            if (k < cols)
              stage[k][e] = w * desc[static_cast<uint64_t>(k - 1) * V + j];
          }
        }
      }
      __syncthreads();
      if (tid < cols) {
        // 16 reads in flight, then 16 adds in order: one LDS round trip per
        // 16 entries instead of one per entry
        const double *row = stage[tid];
        uint32_t e = 0;
        for (; e + 16 <= n; e += 16) {
          double v[16];
#pragma unroll
          for (int u = 0; u < 16; ++u)
            v[u] = row[e + u];
#pragma unroll
          for (int u = 0; u < 16; ++u)
            acc += v[u];
        }
        for (; e < n; ++e)
          acc += row[e];
      }
      __syncthreads();
    }
    if (tid < cols)
      sums[tid] = acc;
    __syncthreads();
    if (tid == 0) {
      iters = it + 1;
      double weight = sums[0];
      if (weight <= 1e-14) {
        flag_exhausted = 1;
      } else {
        double movement = 0;
        double next[12];
        for (int k = 0; k < dim; ++k)
          next[k] = sums[k + 1] / weight;
        for (int k = 0; k < dim; ++k) {
          double d = mean[k] - next[k];
          movement += d * d;
        }
        for (int k = 0; k < dim; ++k)
          mean[k] = next[k];
        if (movement < 1e-6 * h2)
          flag_converged = 1;
      }
    }
    __syncthreads();
    if (flag_exhausted || flag_converged)
      break;
  }
  __syncthreads();
  const uint32_t T = n_touched;
  for (uint32_t e = tid; e < T; e += kTrajThreads) {
    const int32_t j = touched_ids[e];
    out_ids[e] = j;
    out_counts[e] = support[j];
    support[j] = 0;
  }
  if (tid == 0) {
    for (int k = 0; k < 12; ++k)
      out->mean[k] = k < dim ? mean[k] : 0.0;
    out->converged = flag_converged;
    out->exhausted = flag_exhausted;
    out->iterations = iters;
    out->touched = static_cast<int32_t>(T);
  }
}

// Fenwick tree over the unvisited votes: k-th unvisited in ascending order.
struct Fenwick {
  std::vector<int64_t> t;
  int64_t n = 0, total = 0, top = 1;
  explicit Fenwick(int64_t count) : t(count + 1, 0), n(count) {
    for (int64_t i = 1; i <= n; ++i) {
      t[i] += 1;
      int64_t j = i + (i & -i);
      if (j <= n)
        t[j] += t[i];
    }
    total = n;
    while (top * 2 <= n)
      top *= 2;
  }
  void remove(int64_t idx) {
    --total;
    for (int64_t i = idx + 1; i <= n; i += i & -i)
      t[i] -= 1;
  }
  int64_t kth(int64_t k) const {  // 0-based
    int64_t pos = 0;
    for (int64_t step = top; step; step >>= 1)
      if (pos + step <= n && t[pos + step] <= k) {
        pos += step;
        k -= t[pos];
      }
    return pos;  // index (0-based) of the k-th one
  }
};

struct Entry {
  int32_t vote;
  uint64_t count;
};

double ddist(const double *a, const double *b, int dim) {
  double s = 0;
  for (int i = 0; i < dim; ++i) {
    double d = a[i] - b[i];
    s += d * d;
  }
  return s;
}

}  // namespace

struct ClusterResult {
  std::vector<uint8_t> refl;
  std::vector<double> rot, trans, pn, po;
  std::vector<uint64_t> off;
  std::vector<int32_t> members;
  int32_t limited = 0;
};

// Host driver over device trajectories (clustering.cpp:32-154).
static void clustering_run(Session *s, uint64_t V, const uint8_t *kind_h, const double *desc_d,
                           const uint8_t *kind_d, const double *const rot_h[9], double box,
                           const double center[3], double bandwidth, int max_iterations,
                           int max_clusters, mvs_mt64 *rng, ClusterResult &res,
                           const double *const desc_h[12]) {
  const double h = bandwidth, h2 = h * h;
  uint32_t *support = scratch_as<uint32_t>(s, kSlotMs0, V ? V : 1);
  int32_t *list_j = scratch_as<int32_t>(s, kSlotMs1, V ? V : 1);
  double *list_w = scratch_as<double>(s, kSlotMs2, V ? V : 1);
  int32_t *touched = scratch_as<int32_t>(s, kSlotMs3, V ? V : 1);
  // pinned results of one trajectory: state, touched ids, their counts
  const uint64_t vcap = V ? V : 1;
  const size_t ids_at = (sizeof(TrajState) + 255) / 256 * 256;
  const size_t counts_at = ids_at + (vcap * sizeof(int32_t) + 255) / 256 * 256;
  char *pin = static_cast<char *>(host_stage(s, counts_at + vcap * sizeof(uint32_t)));
  char *pin_dev = nullptr;
  MVS_HIP(hipHostGetDevicePointer(reinterpret_cast<void **>(&pin_dev), pin, 0));
  const TrajState &hs = *reinterpret_cast<const TrajState *>(pin);
  const int32_t *pin_ids = reinterpret_cast<const int32_t *>(pin + ids_at);
  const uint32_t *pin_counts = reinterpret_cast<const uint32_t *>(pin + counts_at);
  int32_t *cand = scratch_as<int32_t>(s, kSlotMs6, V ? V : 1);
  const double reach = (1 + kAnchorRadius) * h;
  const double reach2 = reach * reach * (1 + 1e-6);
  const double keep2 = kAnchorRadius * kAnchorRadius * h2 * (1 - 1e-6);
  if (V)
    zero_device(s, support, V * sizeof(uint32_t));
  struct Center {
    double v[12];
    bool refl;
  };
  std::vector<Center> centers;
  std::vector<std::vector<Entry>> counters;
  std::vector<uint8_t> visited(V, 0);
  Fenwick fw(static_cast<int64_t>(V));
  uint64_t support_entries = 0;
  int32_t limited = 0;
  while (fw.total > 0) {
    uint64_t r = mt64_next(rng) % static_cast<uint64_t>(fw.total);
    int64_t seed = fw.kth(static_cast<int64_t>(r));
    bool refl = kind_h[seed] != 0;
    int dim = refl ? 4 : 12;
    MVS_LAUNCH(s, "ms_trajectory", k_trajectory, 1, kTrajThreads, 0, desc_d, kind_d, V,
               static_cast<int32_t>(seed), h2, reach2, keep2, max_iterations, support, cand,
               list_j, list_w, touched, reinterpret_cast<int32_t *>(pin_dev + ids_at),
               reinterpret_cast<uint32_t *>(pin_dev + counts_at),
               reinterpret_cast<TrajState *>(pin_dev));
    sync(s);
    const uint32_t T = static_cast<uint32_t>(hs.touched);
    if (T > V)
      raise(MVS_ERUNTIME, "mean-shift trajectory touched more votes than exist");
    // the kernel's stores into pinned host memory: state, ids, counts
    note_copy(false, sizeof(TrajState) + T * (sizeof(int32_t) + sizeof(uint32_t)));
    std::vector<Entry> sup(T);
    for (uint32_t e = 0; e < T; ++e)
      sup[e] = {pin_ids[e], pin_counts[e]};
    std::sort(sup.begin(), sup.end(), [](const Entry &a, const Entry &b) { return a.vote < b.vote; });
    for (const Entry &e : sup)
      if (!visited[e.vote]) {
        visited[e.vote] = 1;
        fw.remove(e.vote);
      }
    if (hs.exhausted && !visited[seed]) {
      visited[seed] = 1;
      fw.remove(seed);
    }
    if (!hs.converged)
      ++limited;
    int64_t merge = -1;
    for (size_t c = 0; c < centers.size(); ++c)
      if (centers[c].refl == refl && ddist(hs.mean, centers[c].v, dim) < h2 / 4) {
        merge = static_cast<int64_t>(c);
        break;
      }
    if (merge < 0) {
      Center c{};
      for (int k = 0; k < dim; ++k)
        c.v[k] = hs.mean[k];
      c.refl = refl;
      centers.push_back(c);
      support_entries += sup.size();
      counters.push_back(std::move(sup));
    } else {
      auto &cm = counters[merge];
      double old = static_cast<double>(cm.size()), now = static_cast<double>(sup.size());
      if (old + now > 0)
        for (int k = 0; k < dim; ++k)
          centers[merge].v[k] = (old * centers[merge].v[k] + now * hs.mean[k]) / (old + now);
      size_t before = cm.size();
      std::vector<Entry> merged;
      merged.reserve(cm.size() + sup.size());
      size_t i = 0, j = 0;
      while (i < cm.size() || j < sup.size()) {
        if (j >= sup.size() || (i < cm.size() && cm[i].vote < sup[j].vote))
          merged.push_back(cm[i++]);
        else if (i >= cm.size() || sup[j].vote < cm[i].vote)
          merged.push_back(sup[j++]);
        else {
          Entry e = cm[i++];
          e.count += sup[j++].count;
          merged.push_back(e);
        }
      }
      cm.swap(merged);
      support_entries += cm.size() - before;
    }
    if (support_entries > 5000000)
      raise(MVS_ERUNTIME, "Mean-shift support budget exceeded; lower --sample-ratio");
    if (centers.size() > 20000)
      raise(MVS_ERUNTIME, "Too many mean-shift modes; increase --bandwidth or lower --sample-ratio");
  }
  // winners: strictly larger count wins, earlier cluster keeps ties
  std::vector<uint64_t> best(V, 0);
  std::vector<int32_t> winner(V, -1);
  for (size_t c = 0; c < counters.size(); ++c)
    for (const Entry &e : counters[c])
      if (e.count > best[e.vote]) {
        best[e.vote] = e.count;
        winner[e.vote] = static_cast<int32_t>(c);
      }
  std::vector<std::vector<int32_t>> members(centers.size());
  for (uint64_t i = 0; i < V; ++i)
    if (winner[i] >= 0)
      members[winner[i]].push_back(static_cast<int32_t>(i));
  struct Out {
    bool refl;
    double rot[9], trans[3], pn[3], po;
    std::vector<int32_t> votes;
  };
  std::vector<Out> outs;
  const V3 origin = v3(center[0], center[1], center[2]);
  for (size_t c = 0; c < centers.size(); ++c) {
    if (members[c].empty())
      continue;
    Out o{};
    o.refl = centers[c].refl;
    o.votes = std::move(members[c]);
    const double *d = centers[c].v;
    if (o.refl) {
      V3 n = v3(d[0], d[1], d[2]);
      double length = norm(n);
      V3 pn = unit(n);
      double off = d[3] * box / length - dot(pn, origin);
      M3 r = reflection(pn);
      V3 t = pn * (-2 * off);
      std::memcpy(o.rot, r.m, sizeof o.rot);
      o.trans[0] = t.x, o.trans[1] = t.y, o.trans[2] = t.z;
      o.pn[0] = pn.x, o.pn[1] = pn.y, o.pn[2] = pn.z;
      o.po = off;
    } else {
      M3 m;
      for (int k = 0; k < 9; ++k)
        m.m[k] = d[k] * 2;
      V3 c0 = v3(m.m[0], m.m[3], m.m[6]), c1 = v3(m.m[1], m.m[4], m.m[7]);
      V3 a = unit(c0);
      V3 b = unit(c1 - a * dot(a, c1));
      if (norm(a) < 0.5 || norm(b) < 0.5) {
        int32_t v0 = o.votes[0];
        for (int k = 0; k < 9; ++k)
          m.m[k] = rot_h[k][v0];
      } else
        m = frame(a, b, cross(a, b));
      V3 t = v3(d[9], d[10], d[11]);
      V3 tr = (t * box + origin) - apply(m, origin);
      std::memcpy(o.rot, m.m, sizeof o.rot);
      o.trans[0] = tr.x, o.trans[1] = tr.y, o.trans[2] = tr.z;
    }
    outs.push_back(std::move(o));
  }
  std::stable_sort(outs.begin(), outs.end(),
                   [](const Out &a, const Out &b) { return a.votes.size() > b.votes.size(); });
  if (outs.size() > static_cast<size_t>(max_clusters))
    outs.resize(max_clusters);
  res = ClusterResult{};
  res.limited = limited;
  res.off.push_back(0);
  for (const Out &o : outs) {
    res.refl.push_back(o.refl);
    res.rot.insert(res.rot.end(), o.rot, o.rot + 9);
    res.trans.insert(res.trans.end(), o.trans, o.trans + 3);
    res.pn.insert(res.pn.end(), o.pn, o.pn + 3);
    res.po.push_back(o.po);
    res.members.insert(res.members.end(), o.votes.begin(), o.votes.end());
    res.off.push_back(res.members.size());
  }
  (void)desc_h;
}

}  // namespace mvs

using namespace mvs;

extern "C" {

int mvs_clustering(mvs_session *sp, uint64_t V, const uint8_t *kind, const double *const desc[12],
                   const double *const rotation[9], double box, const double center[3],
                   double bandwidth, int max_iterations, int max_clusters, mvs_mt64 *rng,
                   uint64_t *n_clusters, uint64_t *n_members, int32_t *limited) {
  return guarded([&] {
    Session *s = as_session(sp);
    if (!rng || !n_clusters || !n_members || !limited || !center || (V && (!kind || !desc)))
      raise(MVS_EINVAL, "null argument");
    if (V > static_cast<uint64_t>(INT32_MAX))
      raise(MVS_EOVERFLOW, "vote ids exceed int32 range");
    if (!(bandwidth > 0) || max_iterations < 1 || max_clusters < 1)
      raise(MVS_EINVAL, "invalid mean-shift parameters");
    DBuf<double> d(12 * (V ? V : 1));
    DBuf<uint8_t> k(V ? V : 1);
    for (int c = 0; c < 12; ++c)
      copy_h2d(s, d.ptr + c * V, desc[c], V * 8);
    copy_h2d(s, k.ptr, kind, V);
    ClusterResult res;
    clustering_run(s, V, kind, d.ptr, k.ptr, rotation, box, center, bandwidth, max_iterations,
                   max_clusters, rng, res, desc);
    s->cl_refl = std::move(res.refl);
    s->cl_rot = std::move(res.rot);
    s->cl_trans = std::move(res.trans);
    s->cl_pn = std::move(res.pn);
    s->cl_po = std::move(res.po);
    s->cl_off = std::move(res.off);
    s->cl_members = std::move(res.members);
    *n_clusters = s->cl_refl.size();
    *n_members = s->cl_members.size();
    *limited = res.limited;
  });
}

int mvs_clustering_fetch(mvs_session *sp, uint8_t *reflection, double *rotation9,
                         double *translation3, double *plane_normal3, double *plane_offset,
                         uint64_t *member_offsets, int32_t *members) {
  return guarded([&] {
    Session *s = as_session(sp);
    auto put = [](auto *dst, const auto &v) {
      if (dst && !v.empty())
        std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
    };
    put(reflection, s->cl_refl);
    put(rotation9, s->cl_rot);
    put(translation3, s->cl_trans);
    put(plane_normal3, s->cl_pn);
    put(plane_offset, s->cl_po);
    put(member_offsets, s->cl_off);
    put(members, s->cl_members);
  });
}

}  // extern "C"
