// symmetry::Backend over the libmvs C-ABI (include/mvs.h).
//
// Each method is one host-layer call of libmvs with the argument meaning,
// output order and error classes of the reference's CPU backend
// (src/cpu/backend.cpp, include/symmetry/pipeline.h:52-69):
//   curvature       -> mvs_curvature
//   select          -> mvs_sample_vertices (host, one draw per positive
//                      weight) + mvs_prune
//   pairing         -> mvs_pairing + mvs_pairing_fetch (FP64 predicate)
//   transformations -> mvs_transformations + mvs_transformations_fetch
//   clustering      -> mvs_clustering + mvs_clustering_fetch
//   verification    -> mvs_verification + mvs_verification_fetch
// A session (stream + scratch) lives for one call only, so no device state
// survives between calls, as in src/cuda/backend.cu:341-351.
#include "hip_backend.hpp"

#include <chrono>
#include <cstring>
#include <sstream>
#include <stdexcept>
#include <string>

#include "../../include/mvs.h"

namespace symmetry {
namespace {

static_assert(sizeof(Edge) == sizeof(mvs_edge), "edge record layout");
static_assert(sizeof(TransformKind) == 1, "vote kind is one byte");

thread_local std::vector<CudaMetric> g_metrics;

// Status -> the exception class the reference throws for the same failure.
void check(int rc) {
  if (rc == MVS_OK)
    return;
  std::string msg = mvs_last_error();
  switch (rc) {
  case MVS_EINVAL:
    throw std::invalid_argument(msg);
  case MVS_EOVERFLOW:
    throw std::overflow_error(msg);
  default:  // budgets, HIP errors, mean-shift budgets
    throw std::runtime_error(msg);
  }
}

// std::mt19937_64 <-> libstdc++ state layout (x[312], p) through the
// engine's textual state, which libstdc++ writes as the 312 words then p.
mvs_mt64 state_of(const std::mt19937_64 &g) {
  std::stringstream ss;
  ss << g;
  mvs_mt64 st{};
  for (auto &w : st.x)
    ss >> w;
  ss >> st.p;
  if (!ss)
    throw std::runtime_error("unexpected std::mt19937_64 state format");
  return st;
}
void restore(std::mt19937_64 &g, const mvs_mt64 &st) {
  std::stringstream ss;
  for (auto w : st.x)
    ss << w << ' ';
  ss << st.p;
  ss >> g;
  if (!ss)
    throw std::runtime_error("unexpected std::mt19937_64 state format");
}

mvs_points points(PointsView p) { return {p.size(), p.x.data, p.y.data, p.z.data}; }

mvs_features features(FeaturesView f) {
  mvs_features r{};
  r.count = f.size();
  const PointsView *v[3] = {&f.normal, &f.dmin, &f.dmax};
  for (int k = 0; k < 3; ++k) {
    r.lane[3 * k] = v[k]->x.data;
    r.lane[3 * k + 1] = v[k]->y.data;
    r.lane[3 * k + 2] = v[k]->z.data;
  }
  r.lane[9] = f.kmin.data;
  r.lane[10] = f.kmax.data;
  r.valid = f.valid.data;
  return r;
}

// One libmvs session per call; records a metric when profiling, from the
// library's own counters: bytes copied each way and the copies' device time
// (events around each copy), the peak of the device bytes the stage held, and
// whether every buffer was released once the session is gone
// (src/cuda/profile.cpp:61-93).
class Call {
public:
  Call(int device, bool profile, const char *stage) : profile_(profile), stage_(stage) {
    if (profile_) {
      mvs_thread_metrics_reset();
      mvs_memory_metrics m{};
      check(mvs_thread_metrics(&m));
      live0_ = m.live_bytes;
    }
    check(mvs_session_create(device, &s_));
    if (profile_)
      check(mvs_session_enable_timing(s_, 1));
    t0_ = std::chrono::steady_clock::now();
  }
  ~Call() {
    if (s_ && profile_) {
      CudaMetric m;
      m.stage = stage_;
      double ms = 0;
      uint64_t launches = 0;
      if (mvs_dev_timing(s_, nullptr, &ms, &launches) == MVS_OK) {
        m.custom_kernel_ms = ms;
        m.custom_launches = launches;
      }
      if (mvs_dev_timing(s_, "copy_h2d", &ms, nullptr) == MVS_OK)
        m.h2d_ms = ms;
      if (mvs_dev_timing(s_, "copy_d2h", &ms, nullptr) == MVS_OK)
        m.d2h_ms = ms;
      m.gpu_interval_ms = std::chrono::duration<double, std::milli>(
                              std::chrono::steady_clock::now() - t0_)
                              .count();
      mvs_session_destroy(s_);
      s_ = nullptr;
      mvs_memory_metrics c{};
      if (mvs_thread_metrics(&c) == MVS_OK) {
        m.h2d_bytes = c.h2d_bytes;
        m.d2h_bytes = c.d2h_bytes;
        m.peak_buffer_bytes = c.peak_bytes - live0_;
        m.all_buffers_released = c.live_bytes == live0_;
      }
      g_metrics.push_back(m);
    } else if (s_) {
      mvs_session_destroy(s_);
    }
  }
  Call(const Call &) = delete;
  Call &operator=(const Call &) = delete;
  mvs_session *get() const { return s_; }

private:
  mvs_session *s_ = nullptr;
  bool profile_;
  const char *stage_;
  uint64_t live0_ = 0;
  std::chrono::steady_clock::time_point t0_;
};

class HipBackend final : public Backend {
public:
  HipBackend(bool profile, int device) : profile_(profile), device_(device) {
    mvs_session *probe = nullptr;
    check(mvs_session_create(device_, &probe));  // MVS_ENODEV without a GPU
    mvs_session_destroy(probe);
  }
  const char *name() const noexcept override { return "hip"; }

  Features curvature(MeshView m, TopologyView t, const Options &o) const override {
    Call c(device_, profile_, "curvature");
    const Offset n = m.vertices.size();
    Features f;
    f.resize(n);
    mvs_points v = points(m.vertices);
    mvs_topology topo{};
    topo.normals = points(t.normals);
    topo.face_normals = points(t.face_normals);
    topo.neighbor_offsets = t.neighbors.offsets.data;
    topo.neighbor_indices = t.neighbors.indices.data;
    topo.edge_count = t.edges.size();
    topo.edges = reinterpret_cast<const mvs_edge *>(t.edges.data);
    mvs_features_out out{};
    Points *lanes[3] = {&f.normal, &f.dmin, &f.dmax};
    for (int k = 0; k < 3; ++k) {
      out.lane[3 * k] = lanes[k]->x.data();
      out.lane[3 * k + 1] = lanes[k]->y.data();
      out.lane[3 * k + 2] = lanes[k]->z.data();
    }
    out.lane[9] = f.kmin.data();
    out.lane[10] = f.kmax.data();
    out.valid = f.valid.data();
    check(mvs_curvature(c.get(), &v, &topo, o.smoothing, &out));
    return f;
  }

  Selection select(MeshView m, FeaturesView f, ArrayView<const Scalar> w, const Options &o,
                   std::mt19937_64 &random) const override {
    Call c(device_, profile_, "sample_prune");
    require_size(w.size(), m.vertices.size());
    Selection r;
    mvs_mt64 st = state_of(random);
    r.sample.resize(w.size());
    uint64_t count = 0;
    check(mvs_sample_vertices(w.size(), w.data, o.sample_ratio, &st, r.sample.data(), &count));
    r.sample.resize(count);
    restore(random, st);
    mvs_features fv = features(f);
    r.pruned.resize(count);
    uint64_t kept = 0;
    check(mvs_prune(c.get(), &fv, count, r.sample.data(), o.prune, r.pruned.data(), &kept));
    r.pruned.resize(kept);
    return r;
  }

  Pairs pairing(MeshView m, FeaturesView f, ArrayView<const Index> sel, ModelContext ctx,
                const Options &o) const override {
    Call c(device_, profile_, "pairing");
    mvs_points v = points(m.vertices);
    mvs_features fv = features(f);
    uint64_t P = 0;
    check(mvs_pairing(c.get(), &v, &fv, sel.size(), sel.data, ctx.box_size, o.pair_threshold,
                      o.max_pairs, MVS_FP64, &P));
    Pairs p;
    p.a.resize(P);
    p.b.resize(P);
    if (P)
      check(mvs_pairing_fetch(c.get(), p.a.data(), p.b.data()));
    return p;
  }

  Votes transformations(MeshView m, FeaturesView f, PairsView p, ModelContext ctx,
                        const Options &o) const override {
    Call c(device_, profile_, "transformation");
    require_size(p.b.size(), p.a.size());
    mvs_points v = points(m.vertices);
    mvs_features fv = features(f);
    const double center[3] = {ctx.center.x, ctx.center.y, ctx.center.z};
    uint64_t V = 0;
    check(mvs_transformations(c.get(), &v, &fv, p.size(), p.a.data, p.b.data, ctx.box_size,
                              center, o.max_votes, &V));
    Votes r;
    r.pairs.resize(V);
    r.kind.resize(V);
    r.rotation.resize(V);
    r.translation.resize(V);
    r.plane_normal.resize(V);
    r.plane_offset.resize(V);
    r.descriptor.resize(V);
    double *lanes[28];
    for (int k = 0; k < 9; ++k)
      lanes[k] = r.rotation.component[k].data();
    Points *vec[2] = {&r.translation, &r.plane_normal};
    for (int k = 0; k < 2; ++k) {
      lanes[9 + 3 * k] = vec[k]->x.data();
      lanes[10 + 3 * k] = vec[k]->y.data();
      lanes[11 + 3 * k] = vec[k]->z.data();
    }
    lanes[15] = r.plane_offset.data();
    for (int k = 0; k < 12; ++k)
      lanes[16 + k] = r.descriptor.component[k].data();
    if (V)
      check(mvs_transformations_fetch(c.get(), r.pairs.data(),
                                      reinterpret_cast<uint8_t *>(r.kind.data()), lanes));
    return r;
  }

  ClusterOutput clustering(VotesView vo, ModelContext ctx, const Options &o,
                           std::mt19937_64 &random) const override {
    Call c(device_, profile_, "mean_shift");
    const Offset V = vo.size();
    const double *desc[12], *rot[9];
    for (int k = 0; k < 12; ++k)
      desc[k] = vo.descriptor.component[k].data;
    for (int k = 0; k < 9; ++k)
      rot[k] = vo.rotation.component[k].data;
    const double center[3] = {ctx.center.x, ctx.center.y, ctx.center.z};
    mvs_mt64 st = state_of(random);
    uint64_t nc = 0, nm = 0;
    int32_t limited = 0;
    check(mvs_clustering(c.get(), V, reinterpret_cast<const uint8_t *>(vo.kind.data), desc, rot,
                         ctx.box_size, center, o.bandwidth, o.max_iterations, o.max_clusters, &st,
                         &nc, &nm, &limited));
    restore(random, st);
    std::vector<uint8_t> refl(nc);
    std::vector<double> r9(9 * nc), t3(3 * nc), n3(3 * nc), po(nc);
    std::vector<uint64_t> off(nc + 1);
    std::vector<int32_t> mem(nm);
    check(mvs_clustering_fetch(c.get(), refl.data(), r9.data(), t3.data(), n3.data(), po.data(),
                               off.data(), mem.data()));
    ClusterOutput out;
    out.limited_iterations = limited;
    for (uint64_t i = 0; i < nc; ++i) {
      ClusterRecord rec;
      rec.reflection = refl[i] != 0;
      std::memcpy(rec.rotation.data(), &r9[9 * i], 9 * sizeof(double));
      rec.translation = {t3[3 * i], t3[3 * i + 1], t3[3 * i + 2]};
      rec.plane_normal = {n3[3 * i], n3[3 * i + 1], n3[3 * i + 2]};
      rec.plane_offset = po[i];
      std::vector<Index> members(mem.begin() + off[i], mem.begin() + off[i + 1]);
      out.clusters.append(rec, members);
    }
    // unverified clusters carry empty sides and zero coverage / rms, as the
    // CPU backend leaves them (src/cpu/clustering.cpp:152)
    out.clusters.verification = Verification::empty(out.clusters.size());
    return out;
  }

  Verification verification(MeshView m, TopologyView t, ClustersView cl, ModelContext ctx,
                            const Options &o) const override {
    Call c(device_, profile_, "verification");
    const Offset n = m.vertices.size(), F = m.faces.size(), C = cl.size();
    std::vector<uint8_t> refl(C);
    std::vector<double> r9(9 * C), t3(3 * C), n3(3 * C), po(C);
    for (Offset i = 0; i < C; ++i) {
      const ClusterRecord &rec = cl.records[i];
      refl[i] = rec.reflection;
      std::memcpy(&r9[9 * i], rec.rotation.data(), 9 * sizeof(double));
      for (int k = 0; k < 3; ++k) {
        t3[3 * i + k] = rec.translation[k];
        n3[3 * i + k] = rec.plane_normal[k];
      }
      po[i] = rec.plane_offset;
    }
    const mvs_points verts = points(m.vertices), normals = points(t.normals);
    const uint64_t zero = 0;
    mvs_verification_counts k{};
    check(mvs_verification(c.get(), &verts, F, m.faces.component[0].data,
                           m.faces.component[1].data, m.faces.component[2].data, &normals,
                           t.weights.data, t.neighbors.offsets.size() ? t.neighbors.offsets.data
                                                                      : &zero,
                           t.neighbors.indices.data, C, refl.data(), r9.data(), t3.data(),
                           n3.data(), po.data(), o.verify_threshold, ctx.box_size, &k));
    Verification r;
    Csr *sides[2] = {&r.side_a, &r.side_b};
    GroupedCsr *patches[2] = {&r.patches_a, &r.patches_b};
    const uint64_t totals[2] = {k.side_a, k.side_b}, npatch[2] = {k.patches_a, k.patches_b};
    mvs_verification_side out[2];
    for (int s = 0; s < 2; ++s) {
      sides[s]->offsets.resize(C + 1);
      sides[s]->indices.resize(totals[s]);
      patches[s]->groups.resize(C + 1);
      patches[s]->rows.offsets.resize(npatch[s] + 1);
      patches[s]->rows.indices.resize(totals[s]);
      out[s] = {sides[s]->offsets.data(), sides[s]->indices.data(), patches[s]->groups.data(),
                patches[s]->rows.offsets.data(), patches[s]->rows.indices.data()};
    }
    r.coverage.resize(C);
    r.rms.resize(C);
    check(mvs_verification_fetch(c.get(), &out[0], &out[1], r.coverage.data(), r.rms.data()));
    return r;
  }

private:
  bool profile_;
  int device_;
};

}  // namespace

std::unique_ptr<Backend> make_hip_backend(bool profile, int device) {
  return std::make_unique<HipBackend>(profile, device);
}

std::vector<CudaMetric> take_hip_metrics() {
  std::vector<CudaMetric> r;
  r.swap(g_metrics);
  return r;
}

}  // namespace symmetry
