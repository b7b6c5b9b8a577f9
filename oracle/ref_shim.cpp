// Test-infrastructure shim: exposes the reference CPU backend (compiled from
// the reference's own sources by oracle/Makefile into oracle/_ref/) through the
// same C signatures as oracle.h, prefixed ref_.  Used only to pin the oracle
// restatement and as the CPU baseline; never linked into the product.
#include "oracle.h"
#include "symmetry/math.h"
#include "symmetry/pipeline.h"
#include "symmetry/sampling.h"

#include <cstring>
#include <sstream>
#include <stdexcept>
#include <string>

using namespace symmetry;

namespace {
thread_local std::string g_error;
int fail(int code, const std::exception &e) {
  g_error = e.what();
  return code;
}
template <class F> int guarded(F &&f) {
  try {
    f();
    return OR_OK;
  } catch (const std::invalid_argument &e) {
    return fail(OR_EINVAL, e);
  } catch (const std::overflow_error &e) {
    return fail(OR_EOVERFLOW, e);
  } catch (const std::runtime_error &e) {
    std::string m = e.what();
    return fail(m.find("budget") != std::string::npos ? OR_EBUDGET : OR_ERUNTIME, e);
  } catch (const std::exception &e) {
    return fail(OR_ERUNTIME, e);
  }
}
std::mt19937_64 to_engine(const or_mt64 *g) {
  std::stringstream s;
  for (int i = 0; i < 312; ++i)
    s << g->x[i] << ' ';
  s << g->p;
  std::mt19937_64 e;
  s >> e;
  return e;
}
void from_engine(const std::mt19937_64 &e, or_mt64 *g) {
  std::stringstream s;
  s << e;
  for (int i = 0; i < 312; ++i)
    s >> g->x[i];
  s >> g->p;
}
template <class T> ArrayView<const T> av(const T *p, int64_t n) {
  return {n ? p : nullptr, static_cast<Offset>(n)};
}
PointsView pv(const double *x, const double *y, const double *z, int64_t n) {
  return {av(x, n), av(y, n), av(z, n)};
}
} // namespace

extern "C" {
const char *ref_last_error(void) { return g_error.c_str(); }

void ref_mt64_seed(or_mt64 *g, uint64_t seed) { from_engine(std::mt19937_64(seed), g); }
uint64_t ref_mt64_next(or_mt64 *g) {
  auto e = to_engine(g);
  uint64_t v = e();
  from_engine(e, g);
  return v;
}

int ref_prepare(int64_t nv, const double *x, const double *y, const double *z,
                int64_t nf, const int32_t *f0, const int32_t *f1, const int32_t *f2,
                or_prepared *o) {
  return guarded([&] {
    Mesh m;
    m.vertices.x.assign(x, x + nv);
    m.vertices.y.assign(y, y + nv);
    m.vertices.z.assign(z, z + nv);
    m.faces.component[0].assign(f0, f0 + nf);
    m.faces.component[1].assign(f1, f1 + nf);
    m.faces.component[2].assign(f2, f2 + nf);
    Options opt;
    Prepared p = prepare(std::move(m), opt);
    int64_t F = static_cast<int64_t>(p.mesh.faces.size());
    o->face_count = F;
    for (int64_t i = 0; i < F; ++i) {
      o->f0[i] = p.mesh.faces.component[0][i];
      o->f1[i] = p.mesh.faces.component[1][i];
      o->f2[i] = p.mesh.faces.component[2][i];
      o->fnx[i] = p.topology.face_normals.x[i];
      o->fny[i] = p.topology.face_normals.y[i];
      o->fnz[i] = p.topology.face_normals.z[i];
    }
    for (int64_t i = 0; i < nv; ++i) {
      o->nx[i] = p.topology.normals.x[i];
      o->ny[i] = p.topology.normals.y[i];
      o->nz[i] = p.topology.normals.z[i];
      o->weights[i] = p.topology.weights[i];
    }
    const auto &nb = p.topology.neighbors, &inc = p.topology.incident_faces;
    for (size_t i = 0; i < nb.offsets.size(); ++i)
      o->nbr_off[i] = nb.offsets[i];
    std::memcpy(o->nbr_idx, nb.indices.data(), nb.indices.size() * sizeof(int32_t));
    o->nbr_nnz = static_cast<int64_t>(nb.indices.size());
    for (size_t i = 0; i < inc.offsets.size(); ++i)
      o->inc_off[i] = inc.offsets[i];
    std::memcpy(o->inc_idx, inc.indices.data(), inc.indices.size() * sizeof(int32_t));
    o->inc_nnz = static_cast<int64_t>(inc.indices.size());
    static_assert(sizeof(Edge) == sizeof(or_edge), "edge layout");
    std::memcpy(o->edges, p.topology.edges.data(), p.topology.edges.size() * sizeof(Edge));
    o->edge_count = static_cast<int64_t>(p.topology.edges.size());
    o->box_size = p.context.box_size;
    o->center[0] = p.context.center.x;
    o->center[1] = p.context.center.y;
    o->center[2] = p.context.center.z;
    o->rejected_faces = p.diagnostics.rejected_faces;
    o->nonmanifold_edges = p.diagnostics.nonmanifold_edges;
    o->orientation_conflicts = p.diagnostics.orientation_conflicts;
  });
}

int ref_curvature(int64_t nv, const double *x, const double *y, const double *z,
                  const double *nx, const double *ny, const double *nz,
                  const double *fnx, const double *fny, const double *fnz,
                  const uint64_t *nbr_off, const int32_t *nbr_idx, int64_t ne,
                  const or_edge *edges, int smoothing, double *feat, uint8_t *valid) {
  return guarded([&] {
    MeshView m{pv(x, y, z, nv), {}};
    TopologyView t{};
    t.normals = pv(nx, ny, nz, nv);
    t.face_normals = pv(fnx, fny, fnz, 1 << 30); // indexed by face id only
    t.neighbors = {av(nbr_off, nv + 1), av(nbr_idx, static_cast<int64_t>(nbr_off[nv]))};
    t.edges = av(reinterpret_cast<const Edge *>(edges), ne);
    Options o;
    o.smoothing = smoothing;
    auto cpu = make_cpu_backend();
    Features f = cpu->curvature(m, t, o);
    const std::vector<double> *lanes[11] = {&f.normal.x, &f.normal.y, &f.normal.z,
                                            &f.dmin.x,   &f.dmin.y,   &f.dmin.z,
                                            &f.dmax.x,   &f.dmax.y,   &f.dmax.z,
                                            &f.kmin,     &f.kmax};
    for (int k = 0; k < 11; ++k)
      std::memcpy(feat + k * nv, lanes[k]->data(), nv * sizeof(double));
    std::memcpy(valid, f.valid.data(), nv);
  });
}

// Backend::select (sampling + prune) in one call.
int ref_select(int64_t nv, const double *weights, const double *kmin,
               const double *kmax, const uint8_t *valid, double ratio, double prune,
               or_mt64 *rng, int32_t *sample, int64_t *nsample, int32_t *pruned,
               int64_t *npruned) {
  return guarded([&] {
    std::vector<double> zeros(nv, 0.0);
    MeshView m{pv(zeros.data(), zeros.data(), zeros.data(), nv), {}};
    PointsView zp = pv(zeros.data(), zeros.data(), zeros.data(), nv);
    FeaturesView f{zp, zp, zp, av(kmin, nv), av(kmax, nv), av(valid, nv)};
    Options o;
    o.sample_ratio = ratio;
    o.prune = prune;
    auto e = to_engine(rng);
    auto cpu = make_cpu_backend();
    Selection s = cpu->select(m, f, av(weights, nv), o, e);
    from_engine(e, rng);
    std::memcpy(sample, s.sample.data(), s.sample.size() * sizeof(int32_t));
    std::memcpy(pruned, s.pruned.data(), s.pruned.size() * sizeof(int32_t));
    *nsample = static_cast<int64_t>(s.sample.size());
    *npruned = static_cast<int64_t>(s.pruned.size());
  });
}

int ref_pairing(int64_t nv, const double *x, const double *y, const double *z,
                const double *kmin, const double *kmax, int64_t ns,
                const int32_t *selected, double box, double threshold,
                uint64_t max_pairs, int64_t cap, int32_t *out_a, int32_t *out_b,
                int64_t *count) {
  return guarded([&] {
    std::vector<double> zeros(nv, 0.0);
    std::vector<uint8_t> ones(nv, 1);
    PointsView zp = pv(zeros.data(), zeros.data(), zeros.data(), nv);
    MeshView m{pv(x, y, z, nv), {}};
    FeaturesView f{zp, zp, zp, av(kmin, nv), av(kmax, nv), av(ones.data(), nv)};
    Options o;
    o.pair_threshold = threshold;
    o.max_pairs = max_pairs;
    auto cpu = make_cpu_backend();
    Pairs p = cpu->pairing(m, f, av(selected, ns), ModelContext{box, {}}, o);
    if (static_cast<int64_t>(p.size()) > cap)
      throw std::overflow_error("pair output capacity exceeded");
    std::memcpy(out_a, p.a.data(), p.size() * sizeof(int32_t));
    std::memcpy(out_b, p.b.data(), p.size() * sizeof(int32_t));
    *count = static_cast<int64_t>(p.size());
  });
}

int ref_transformations(int64_t nv, const double *x, const double *y, const double *z,
                        const double *feat, int64_t np, const int32_t *pa,
                        const int32_t *pb, double box, const double *center,
                        uint64_t max_votes, int64_t cap, int32_t *pair, uint8_t *kind,
                        double *lanes, int64_t *count) {
  return guarded([&] {
    std::vector<uint8_t> ones(nv, 1);
    auto lane = [&](int k) { return feat + k * nv; };
    FeaturesView f{pv(lane(0), lane(1), lane(2), nv), pv(lane(3), lane(4), lane(5), nv),
                   pv(lane(6), lane(7), lane(8), nv), av(lane(9), nv),
                   av(lane(10), nv),                   av(ones.data(), nv)};
    MeshView m{pv(x, y, z, nv), {}};
    PairsView p{av(pa, np), av(pb, np)};
    Options o;
    o.max_votes = max_votes;
    auto cpu = make_cpu_backend();
    Votes v = cpu->transformations(m, f, p, ModelContext{box, {center[0], center[1], center[2]}}, o);
    int64_t V = static_cast<int64_t>(v.size());
    if (V > cap)
      throw std::overflow_error("vote output capacity exceeded");
    for (int64_t i = 0; i < V; ++i) {
      pair[i] = v.pairs[i];
      kind[i] = static_cast<uint8_t>(v.kind[i]);
    }
    const std::vector<double> *src[28];
    for (int k = 0; k < 9; ++k)
      src[k] = &v.rotation.component[k];
    src[9] = &v.translation.x;
    src[10] = &v.translation.y;
    src[11] = &v.translation.z;
    src[12] = &v.plane_normal.x;
    src[13] = &v.plane_normal.y;
    src[14] = &v.plane_normal.z;
    src[15] = &v.plane_offset;
    for (int k = 0; k < 12; ++k)
      src[16 + k] = &v.descriptor.component[k];
    for (int k = 0; k < 28; ++k)
      std::memcpy(lanes + k * cap, src[k]->data(), V * sizeof(double));
    *count = V;
  });
}

int ref_clustering(int64_t nv, const uint8_t *kind, const double *desc,
                   const double *rotation, double box, const double *center,
                   double bandwidth, int max_iterations, int max_clusters, or_mt64 *rng,
                   int32_t *n_clusters, uint8_t *c_refl, double *c_rot, double *c_trans,
                   double *c_pn, double *c_po, uint64_t *member_off, int32_t *members,
                   int32_t *limited) {
  return guarded([&] {
    std::vector<double> zeros(nv, 0.0);
    std::vector<Index> ids(nv);
    for (int64_t i = 0; i < nv; ++i)
      ids[i] = static_cast<Index>(i);
    VotesView v{};
    v.pairs = av(ids.data(), nv);
    v.kind = av(reinterpret_cast<const TransformKind *>(kind), nv);
    for (int k = 0; k < 9; ++k)
      v.rotation.component[k] = av(rotation + k * nv, nv);
    v.translation = pv(zeros.data(), zeros.data(), zeros.data(), nv);
    v.plane_normal = v.translation;
    v.plane_offset = av(zeros.data(), nv);
    for (int k = 0; k < 12; ++k)
      v.descriptor.component[k] = av(desc + k * nv, nv);
    Options o;
    o.bandwidth = bandwidth;
    o.max_iterations = max_iterations;
    o.max_clusters = max_clusters;
    auto e = to_engine(rng);
    auto cpu = make_cpu_backend();
    ClusterOutput r = cpu->clustering(v, ModelContext{box, {center[0], center[1], center[2]}}, o, e);
    from_engine(e, rng);
    int64_t C = static_cast<int64_t>(r.clusters.size());
    *n_clusters = static_cast<int32_t>(C);
    *limited = r.limited_iterations;
    for (int64_t c = 0; c < C; ++c) {
      const ClusterRecord &rec = r.clusters.records[c];
      c_refl[c] = rec.reflection;
      for (int k = 0; k < 9; ++k)
        c_rot[9 * c + k] = rec.rotation[k];
      for (int k = 0; k < 3; ++k) {
        c_trans[3 * c + k] = rec.translation[k];
        c_pn[3 * c + k] = rec.plane_normal[k];
      }
      c_po[c] = rec.plane_offset;
    }
    for (size_t i = 0; i < r.clusters.votes.offsets.size(); ++i)
      member_off[i] = r.clusters.votes.offsets[i];
    std::memcpy(members, r.clusters.votes.indices.data(),
                r.clusters.votes.indices.size() * sizeof(int32_t));
  });
}
// Backend::verification of the CPU backend (src/cpu/verification.cpp:50-110).
// Sides: offsets (nc+1) and indices (capacity nc*nv each); patches: groups
// (nc+1), row offsets (capacity nc*nv+1) and members (capacity nc*nv);
// counts[4] = side_a, side_b, patches_a, patches_b totals.
int ref_verification(int64_t nv, const double *x, const double *y, const double *z,
                     int64_t nf, const int32_t *f0, const int32_t *f1, const int32_t *f2,
                     const double *nx, const double *ny, const double *nz,
                     const double *weights, const uint64_t *nbr_off, const int32_t *nbr_idx,
                     int64_t nc, const uint8_t *refl, const double *rot, const double *trans,
                     const double *pn, const double *po, double verify_threshold, double box,
                     uint64_t *a_off, int32_t *a_idx, uint64_t *b_off, int32_t *b_idx,
                     uint64_t *pa_groups, uint64_t *pa_off, int32_t *pa_idx,
                     uint64_t *pb_groups, uint64_t *pb_off, int32_t *pb_idx,
                     double *coverage, double *rms, uint64_t *counts) {
  return guarded([&] {
    MeshView m{pv(x, y, z, nv), {}};
    m.faces.component[0] = av(f0, nf);
    m.faces.component[1] = av(f1, nf);
    m.faces.component[2] = av(f2, nf);
    TopologyView t{};
    t.normals = pv(nx, ny, nz, nv);
    t.weights = av(weights, nv);
    t.neighbors = {av(nbr_off, nv + 1), av(nbr_idx, static_cast<int64_t>(nbr_off[nv]))};
    std::vector<ClusterRecord> recs(nc);
    for (int64_t c = 0; c < nc; ++c) {
      recs[c].reflection = refl[c] != 0;
      for (int k = 0; k < 9; ++k)
        recs[c].rotation[k] = rot[9 * c + k];
      recs[c].translation = {trans[3 * c], trans[3 * c + 1], trans[3 * c + 2]};
      recs[c].plane_normal = {pn[3 * c], pn[3 * c + 1], pn[3 * c + 2]};
      recs[c].plane_offset = po[c];
    }
    std::vector<Offset> voff(nc + 1, 0);
    ClustersView cv{av(recs.data(), nc), {av(voff.data(), nc + 1), {}}};
    Options o;
    o.verify_threshold = verify_threshold;
    auto cpu = make_cpu_backend();
    Verification r = cpu->verification(m, t, cv, ModelContext{box, {}}, o);
    auto put = [](auto *dst, const auto &v) {
      if (!v.empty())
        std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
    };
    put(a_off, r.side_a.offsets);
    put(a_idx, r.side_a.indices);
    put(b_off, r.side_b.offsets);
    put(b_idx, r.side_b.indices);
    put(pa_groups, r.patches_a.groups);
    put(pa_off, r.patches_a.rows.offsets);
    put(pa_idx, r.patches_a.rows.indices);
    put(pb_groups, r.patches_b.groups);
    put(pb_off, r.patches_b.rows.offsets);
    put(pb_idx, r.patches_b.rows.indices);
    put(coverage, r.coverage);
    put(rms, r.rms);
    counts[0] = r.side_a.indices.size();
    counts[1] = r.side_b.indices.size();
    counts[2] = r.patches_a.rows.size();
    counts[3] = r.patches_b.rows.size();
  });
}

// The reference's exact squared point-triangle distance (math.cpp:72-112).
double ref_point_triangle_distance2(const double *p, const double *a, const double *b,
                                    const double *c) {
  return point_triangle_distance2(Vec{p[0], p[1], p[2]}, Vec{a[0], a[1], a[2]},
                                  Vec{b[0], b[1], b[2]}, Vec{c[0], c[1], c[2]});
}

} // extern "C"
