// detect() over the libmvs host layer (src/core/pipeline.cpp:5-52).
#include "detect.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <unordered_map>

#include "../../include/mvs.h"

namespace mvs_app {
namespace {

void check(int rc) {
  if (rc == MVS_OK)
    return;
  const std::string msg = mvs_last_error();
  if (rc == MVS_EINVAL)
    throw std::invalid_argument(msg);
  if (rc == MVS_EOVERFLOW)
    throw std::overflow_error(msg);
  throw std::runtime_error(msg);
}

struct Session {
  mvs_session *s = nullptr;
  explicit Session(int device) { check(mvs_session_create(device, &s)); }
  ~Session() {
    if (s)
      mvs_session_destroy(s);
  }
  Session(const Session &) = delete;
  Session &operator=(const Session &) = delete;
};

// Host topology of a prepared mesh (the outputs of mvs_prepare).
struct Topology {
  std::vector<int32_t> f[3];
  std::vector<double> n[3], fn[3], weights;
  std::vector<uint64_t> nbr_off, inc_off;
  std::vector<int32_t> nbr_idx, inc_idx;
  std::vector<mvs_edge> edges;
  double box = 0, center[3] = {0, 0, 0};
  int rejected = 0, nonmanifold = 0, conflicts = 0;
};

Topology prepare_mesh(const Mesh &m) {
  const uint64_t nv = m.vertices(), nf = m.faces();
  Topology t;
  for (int k = 0; k < 3; ++k) {
    t.f[k].resize(nf);
    t.n[k].resize(nv);
    t.fn[k].resize(nf);
  }
  t.weights.resize(nv);
  t.nbr_off.resize(nv + 1);
  t.inc_off.resize(nv + 1);
  t.nbr_idx.resize(6 * nf);
  t.inc_idx.resize(3 * nf);
  t.edges.resize(3 * nf);
  mvs_prepared o{};
  o.f0 = t.f[0].data(), o.f1 = t.f[1].data(), o.f2 = t.f[2].data();
  o.nx = t.n[0].data(), o.ny = t.n[1].data(), o.nz = t.n[2].data();
  o.fnx = t.fn[0].data(), o.fny = t.fn[1].data(), o.fnz = t.fn[2].data();
  o.weights = t.weights.data();
  o.nbr_off = t.nbr_off.data(), o.nbr_idx = t.nbr_idx.data();
  o.inc_off = t.inc_off.data(), o.inc_idx = t.inc_idx.data();
  o.edges = t.edges.data();
  check(mvs_prepare(static_cast<int64_t>(nv), m.x.data(), m.y.data(), m.z.data(),
                    static_cast<int64_t>(nf), m.f0.data(), m.f1.data(), m.f2.data(), &o));
  const size_t F = static_cast<size_t>(o.face_count);
  for (int k = 0; k < 3; ++k) {
    t.f[k].resize(F);
    t.fn[k].resize(F);
  }
  t.nbr_idx.resize(static_cast<size_t>(o.nbr_nnz));
  t.inc_idx.resize(static_cast<size_t>(o.inc_nnz));
  t.edges.resize(static_cast<size_t>(o.edge_count));
  t.box = o.box_size;
  for (int k = 0; k < 3; ++k)
    t.center[k] = o.center[k];
  t.rejected = o.rejected_faces;
  t.nonmanifold = o.nonmanifold_edges;
  t.conflicts = o.orientation_conflicts;
  return t;
}

// One level of uniform midpoint subdivision (src/core/prepare.cpp:59-84):
// edge midpoints are appended in the order the faces first reach them, and
// every face becomes its three corner triangles and the middle one.
void subdivide(Mesh &m) {
  if (m.faces() > 500000)
    throw std::runtime_error("Subdivision face budget exceeded");
  std::unordered_map<uint64_t, int32_t> mid;
  mid.reserve(3 * m.faces());
  auto midpoint = [&](int32_t a, int32_t b) {
    const uint64_t key = (static_cast<uint64_t>(static_cast<uint32_t>(std::min(a, b))) << 32) |
                         static_cast<uint32_t>(std::max(a, b));
    auto it = mid.find(key);
    if (it != mid.end())
      return it->second;
    const int32_t id = static_cast<int32_t>(m.x.size());
    // (p_a + p_b) / 2, the division as a product with 1 / 2
    m.x.push_back((m.x[a] + m.x[b]) * 0.5);
    m.y.push_back((m.y[a] + m.y[b]) * 0.5);
    m.z.push_back((m.z[a] + m.z[b]) * 0.5);
    mid.emplace(key, id);
    return id;
  };
  Mesh out;
  const size_t nf = m.faces();
  for (auto *v : {&out.f0, &out.f1, &out.f2})
    v->reserve(4 * nf);
  auto put = [&](int32_t a, int32_t b, int32_t c) {
    out.f0.push_back(a);
    out.f1.push_back(b);
    out.f2.push_back(c);
  };
  for (size_t i = 0; i < nf; ++i) {
    const int32_t p = m.f0[i], q = m.f1[i], r = m.f2[i];
    const int32_t a = midpoint(p, q);
    const int32_t b = midpoint(q, r);
    const int32_t c = midpoint(r, p);
    put(p, a, c);
    put(a, q, b);
    put(c, b, r);
    put(a, b, c);
  }
  m.f0.swap(out.f0);
  m.f1.swap(out.f1);
  m.f2.swap(out.f2);
}

double ms_since(std::chrono::steady_clock::time_point &clock) {
  const auto now = std::chrono::steady_clock::now();
  const double ms = std::chrono::duration<double, std::milli>(now - clock).count();
  clock = now;
  return ms;
}

}  // namespace

void validate_options(const Options &o) {
  const uint64_t cap = static_cast<uint64_t>(std::numeric_limits<int32_t>::max());
  if (o.max_pairs > cap || o.max_votes > cap)
    throw std::invalid_argument("Pair/vote budgets exceed 32-bit ID capacity");
  if (!(o.sample_ratio > 0 && o.sample_ratio <= 1) || !(o.prune > 0 && o.prune < 1) ||
      !(o.pair_threshold > 0 && o.pair_threshold <= 1) ||
      !(o.bandwidth > 0 && o.bandwidth <= 2) ||
      !(o.verify_threshold > 0 && o.verify_threshold < 0.5) || o.smoothing < 0 ||
      o.smoothing > 1000 || o.subdivide < 0 || o.subdivide > 3 || o.max_iterations < 1 ||
      o.max_iterations > 10000 || o.max_clusters < 1 || o.max_clusters > 1000 ||
      o.max_pairs < 1 || o.max_votes < 1)
    throw std::runtime_error("Invalid parameter range; see --help");
}

Result detect(Mesh mesh, const Options &options, int device, const Progress &progress) {
  auto clock = std::chrono::steady_clock::now();
  Result r;
  r.options = options;
  auto timed = [&](const char *stage) {
    const double ms = ms_since(clock);
    r.timings.emplace_back(stage, ms);
    if (progress)
      progress(stage, ms);
  };

  // prepare (src/core/prepare.cpp:102-149): libmvs validates the mesh,
  // rejects degenerate faces and builds the topology; subdivision runs on the
  // surviving faces and the topology is rebuilt on the subdivided mesh.  That
  // second pass applies the degenerate-face test again, which the reference
  // does not: a child of a face whose area lies within 4x the test's bound
  // (1e-14 box^2) would be dropped here and kept there.
  validate_options(options);
  if (mesh.vertices() > static_cast<size_t>(std::numeric_limits<int32_t>::max()) ||
      mesh.faces() > static_cast<size_t>(std::numeric_limits<int32_t>::max()))
    throw std::runtime_error("Mesh exceeds 32-bit index capacity");
  Topology topo = prepare_mesh(mesh);
  const int rejected = topo.rejected;
  mesh.f0 = topo.f[0];
  mesh.f1 = topo.f[1];
  mesh.f2 = topo.f[2];
  if (options.subdivide > 0) {
    for (int i = 0; i < options.subdivide; ++i)
      subdivide(mesh);
    topo = prepare_mesh(mesh);
    mesh.f0 = topo.f[0];
    mesh.f1 = topo.f[1];
    mesh.f2 = topo.f[2];
  }
  r.rejected_faces = rejected;
  r.nonmanifold_edges = topo.nonmanifold;
  r.orientation_conflicts = topo.conflicts;
  r.box_size = topo.box;
  for (int k = 0; k < 3; ++k)
    r.center[k] = topo.center[k];
  r.mesh = std::move(mesh);
  timed("prepare");

  const Mesh &m = r.mesh;
  const uint64_t nv = m.vertices(), nf = m.faces();
  Session s(device);
  const mvs_points verts{nv, m.x.data(), m.y.data(), m.z.data()};
  const mvs_points normals{nv, topo.n[0].data(), topo.n[1].data(), topo.n[2].data()};
  const mvs_topology tv{normals,
                        {nf, topo.fn[0].data(), topo.fn[1].data(), topo.fn[2].data()},
                        topo.nbr_off.data(),
                        topo.nbr_idx.data(),
                        topo.edges.size(),
                        topo.edges.data()};

  // curvature
  mvs_features_out fo{};
  for (int k = 0; k < 11; ++k) {
    r.feat[k].resize(nv);
    fo.lane[k] = r.feat[k].data();
  }
  r.valid.resize(nv);
  fo.valid = r.valid.data();
  check(mvs_curvature(s.s, &verts, &tv, options.smoothing, &fo));
  timed("curvature");
  mvs_features fv{};
  fv.count = nv;
  for (int k = 0; k < 11; ++k)
    fv.lane[k] = r.feat[k].data();
  fv.valid = r.valid.data();

  // select: weighted sampling then the anisotropy prune; one random engine
  // seeded once serves sampling and mean-shift (pipeline.cpp:25, 31, 41)
  mvs_mt64 rng;
  mvs_mt64_seed(&rng, options.seed);
  r.sample.resize(nv);
  uint64_t ns = 0;
  check(mvs_sample_vertices(nv, topo.weights.data(), options.sample_ratio, &rng,
                            r.sample.data(), &ns));
  r.sample.resize(ns);
  r.pruned.resize(ns);
  uint64_t np = 0;
  check(mvs_prune(s.s, &fv, ns, r.sample.data(), options.prune, r.pruned.data(), &np));
  r.pruned.resize(np);
  timed("sample_prune");

  // pairing (FP64 predicate: the reference's pair set and order)
  uint64_t P = 0;
  check(mvs_pairing(s.s, &verts, &fv, np, r.pruned.data(), r.box_size, options.pair_threshold,
                    options.max_pairs, MVS_FP64, &P));
  r.pair_a.resize(P);
  r.pair_b.resize(P);
  if (P)
    check(mvs_pairing_fetch(s.s, r.pair_a.data(), r.pair_b.data()));
  timed("pairing");

  // transformation votes
  uint64_t V = 0;
  check(mvs_transformations(s.s, &verts, &fv, P, r.pair_a.data(), r.pair_b.data(), r.box_size,
                            r.center, options.max_votes, &V));
  r.vote_pair.resize(V);
  r.vote_kind.resize(V);
  double *lanes[28];
  for (int k = 0; k < 28; ++k) {
    r.vote_lanes[k].resize(V);
    lanes[k] = r.vote_lanes[k].data();
  }
  if (V)
    check(mvs_transformations_fetch(s.s, r.vote_pair.data(), r.vote_kind.data(), lanes));
  timed("transformation");

  // mean-shift clustering
  const double *desc[12], *rot[9];
  for (int k = 0; k < 12; ++k)
    desc[k] = r.vote_lanes[16 + k].data();
  for (int k = 0; k < 9; ++k)
    rot[k] = r.vote_lanes[k].data();
  uint64_t nc = 0, nm = 0;
  int32_t limited = 0;
  check(mvs_clustering(s.s, V, r.vote_kind.data(), desc, rot, r.box_size, r.center,
                       options.bandwidth, options.max_iterations, options.max_clusters, &rng, &nc,
                       &nm, &limited));
  r.limited_iterations = limited;
  r.c_reflection.resize(nc);
  r.c_rotation.resize(9 * nc);
  r.c_translation.resize(3 * nc);
  r.c_plane_normal.resize(3 * nc);
  r.c_plane_offset.resize(nc);
  r.c_member_offsets.resize(nc + 1);
  r.c_members.resize(nm);
  check(mvs_clustering_fetch(s.s, r.c_reflection.data(), r.c_rotation.data(),
                             r.c_translation.data(), r.c_plane_normal.data(),
                             r.c_plane_offset.data(), r.c_member_offsets.data(),
                             r.c_members.data()));
  timed("mean_shift");

  // verification; unverified clusters keep empty sides and zero coverage
  // and rms (src/cpu/clustering.cpp:152)
  r.coverage.assign(nc, 0.0);
  r.rms.assign(nc, 0.0);
  for (auto &sd : r.side) {
    sd.offsets.assign(nc + 1, 0);
    sd.groups.assign(nc + 1, 0);
  }
  if (options.patching && nc) {
    mvs_verification_counts k{};
    check(mvs_verification(s.s, &verts, nf, m.f0.data(), m.f1.data(), m.f2.data(), &normals,
                           topo.weights.data(), topo.nbr_off.data(), topo.nbr_idx.data(), nc,
                           r.c_reflection.data(), r.c_rotation.data(), r.c_translation.data(),
                           r.c_plane_normal.data(), r.c_plane_offset.data(),
                           options.verify_threshold, r.box_size, &k));
    const uint64_t totals[2] = {k.side_a, k.side_b}, patches[2] = {k.patches_a, k.patches_b};
    mvs_verification_side out[2];
    for (int sd = 0; sd < 2; ++sd) {
      Side &x = r.side[sd];
      x.ids.resize(totals[sd]);
      x.patch_offsets.resize(patches[sd] + 1);
      x.patch_members.resize(totals[sd]);
      out[sd] = {x.offsets.data(), x.ids.data(), x.groups.data(), x.patch_offsets.data(),
                 x.patch_members.data()};
    }
    check(mvs_verification_fetch(s.s, &out[0], &out[1], r.coverage.data(), r.rms.data()));
  }
  timed("verification");
  return r;
}

}  // namespace mvs_app
