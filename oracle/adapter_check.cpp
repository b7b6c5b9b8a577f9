// TEST INFRASTRUCTURE ONLY.  Runs the reference's own detect()
// (src/core/pipeline.cpp:5-52) on one mesh twice: with its CPU backend
// (src/cpu/backend.cpp) and with the HIP adapter over libmvs
// (mitra_algo_amd/adapter/hip_backend.cpp).  Then it compares every stage with
// the reference's own test bars:
//   features 1e-10 (tests/cuda_curvature.cpp:9), sample/pruned/pairs exact
//   (tests/cuda_pairing.cpp, cuda_backend.cpp:159), votes 2e-12 relative
//   (tests/cuda_backend.cpp:11-15), clusters: members exact, records 1e-10
//   (tests/cuda_clustering.cpp), RNG state equal after select and clustering
//   (tests/cuda_backend.cpp:159, cuda_clustering.cpp:30).
// Prints one JSON object; exit status 0 when everything matches.
// Built into oracle/_ref/ by oracle/Makefile, with the reference's sources
// compiled where they lie.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>

#include "../mitra_algo_amd/adapter/hip_backend.hpp"
#include "symmetry/io.h"

using namespace symmetry;

namespace {
struct Diff {
  double worst = 0;
  bool ok = true;
  void close(double a, double b, double tol) {
    double d = std::abs(a - b) / (1 + std::abs(a));
    if (!(std::isfinite(a) && std::isfinite(b)) || !(d <= tol))
      ok = false;
    if (d > worst || std::isnan(d))
      worst = d;
  }
};
}  // namespace

static int run(int argc, char **argv);

int main(int argc, char **argv) {
  try {
    return run(argc, argv);
  } catch (const std::exception &e) {
    std::printf("{\"ok\": false, \"error\": \"%s\"}\n", e.what());
    return 1;
  }
}

static int run(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: adapter_check mesh.off [pair_threshold sample_ratio seed]\n");
    return 2;
  }
  Options o;
  if (argc > 2)
    o.pair_threshold = std::atof(argv[2]);
  if (argc > 3)
    o.sample_ratio = std::atof(argv[3]);
  if (argc > 4)
    o.seed = std::strtoull(argv[4], nullptr, 10);
  Mesh mesh = read_mesh(argv[1]);
  auto cpu = make_cpu_backend();
  auto hip = make_hip_backend(true);
  Result a = detect(mesh, o, *cpu);
  Result b = detect(mesh, o, *hip);
  auto metrics = take_hip_metrics();

  Diff feat, vote, clus;
  bool valid_eq = a.features.valid == b.features.valid;
  auto fa = a.features.as_view(), fb = b.features.as_view();
  for (Offset i = 0; i < fa.size(); ++i) {
    Feature x = fa[i], y = fb[i];
    feat.close(x.kmin, y.kmin, 1e-10);
    feat.close(x.kmax, y.kmax, 1e-10);
    for (int k = 0; k < 3; ++k) {
      feat.close(x.normal[k], y.normal[k], 1e-10);
      feat.close(x.dmin[k], y.dmin[k], 1e-10);
      feat.close(x.dmax[k], y.dmax[k], 1e-10);
    }
  }
  bool sample_eq = a.sample == b.sample, pruned_eq = a.pruned == b.pruned;
  bool pairs_eq = a.pairs.a == b.pairs.a && a.pairs.b == b.pairs.b;
  bool votes_eq = a.votes.pairs == b.votes.pairs && a.votes.kind == b.votes.kind;
  if (votes_eq) {
    auto va = a.votes.as_view(), vb = b.votes.as_view();
    for (Offset i = 0; i < va.size(); ++i) {
      Vote x = va[i], y = vb[i];
      for (int k = 0; k < 9; ++k)
        vote.close(x.rotation[k], y.rotation[k], 2e-12);
      for (int k = 0; k < 3; ++k) {
        vote.close(x.translation[k], y.translation[k], 2e-12);
        vote.close(x.plane_normal[k], y.plane_normal[k], 2e-12);
      }
      vote.close(x.plane_offset, y.plane_offset, 2e-12);
      for (Offset k = 0; k < x.descriptor.size(); ++k)
        vote.close(x.descriptor[k], y.descriptor[k], 2e-12);
    }
  }
  bool clusters_eq = a.clusters.size() == b.clusters.size() &&
                     a.clusters.votes.offsets == b.clusters.votes.offsets &&
                     a.clusters.votes.indices == b.clusters.votes.indices &&
                     a.limited_iterations == b.limited_iterations;
  if (clusters_eq)
    for (Offset i = 0; i < a.clusters.size(); ++i) {
      const ClusterRecord &x = a.clusters.records[i], &y = b.clusters.records[i];
      clusters_eq = clusters_eq && x.reflection == y.reflection;
      for (int k = 0; k < 9; ++k)
        clus.close(x.rotation[k], y.rotation[k], 1e-10);
      for (int k = 0; k < 3; ++k) {
        clus.close(x.translation[k], y.translation[k], 1e-10);
        clus.close(x.plane_normal[k], y.plane_normal[k], 1e-10);
      }
      clus.close(x.plane_offset, y.plane_offset, 1e-10);
    }

  // RNG contract: identical engine state after select and after clustering
  bool rng_eq = true;
  {
    std::mt19937_64 ga(o.seed), gb(o.seed);
    auto prepared = prepare(a.mesh, o);
    auto t = prepared.topology.as_view();
    auto sa = cpu->select(a.mesh.as_view(), fa, t.weights, o, ga);
    auto sb = hip->select(a.mesh.as_view(), fa, t.weights, o, gb);
    rng_eq = rng_eq && ga == gb && sa.sample == sb.sample && sa.pruned == sb.pruned;
    auto ca = cpu->clustering(a.votes.as_view(), prepared.context, o, ga);
    auto cb = hip->clustering(a.votes.as_view(), prepared.context, o, gb);
    rng_eq = rng_eq && ga == gb && ca.clusters.size() == cb.clusters.size();
  }

  // verification of the CPU run's clusters through both backends: sets,
  // patches and coverage identical, rms identical
  bool verify_eq = true;
  {
    Prepared p = prepare(read_mesh(argv[1]), o);
    Verification va = cpu->verification(p.mesh.as_view(), p.topology.as_view(),
                                        a.clusters.as_view(), p.context, o);
    Verification vb = hip->verification(p.mesh.as_view(), p.topology.as_view(),
                                        a.clusters.as_view(), p.context, o);
    verify_eq = va.side_a.offsets == vb.side_a.offsets && va.side_a.indices == vb.side_a.indices &&
                va.side_b.offsets == vb.side_b.offsets && va.side_b.indices == vb.side_b.indices &&
                va.patches_a.groups == vb.patches_a.groups &&
                va.patches_a.rows.offsets == vb.patches_a.rows.offsets &&
                va.patches_a.rows.indices == vb.patches_a.rows.indices &&
                va.patches_b.groups == vb.patches_b.groups &&
                va.patches_b.rows.offsets == vb.patches_b.rows.offsets &&
                va.patches_b.rows.indices == vb.patches_b.rows.indices &&
                va.coverage == vb.coverage && va.rms == vb.rms &&
                a.clusters.verification.side_a.indices == va.side_a.indices;
  }

  bool ok = valid_eq && feat.ok && sample_eq && pruned_eq && pairs_eq && votes_eq && vote.ok &&
            clusters_eq && clus.ok && rng_eq && verify_eq;
  std::ostringstream js;
  js.precision(17);
  js << std::boolalpha;
  js << "{\"ok\": " << (ok ? "true" : "false") << ", \"backend\": \"" << hip->name()
     << "\", \"vertices\": " << mesh.vertices.size()
     << ", \"samples\": " << b.sample.size() << ", \"pruned\": " << b.pruned.size()
     << ", \"pairs\": " << b.pairs.size() << ", \"votes\": " << b.votes.size()
     << ", \"clusters\": " << b.clusters.size() << ", \"valid_equal\": " << valid_eq
     << ", \"feature_max_rel\": " << feat.worst << ", \"sample_equal\": " << sample_eq
     << ", \"pruned_equal\": " << pruned_eq << ", \"pairs_equal\": " << pairs_eq
     << ", \"votes_order_equal\": " << votes_eq << ", \"vote_max_rel\": " << vote.worst
     << ", \"clusters_equal\": " << clusters_eq << ", \"cluster_max_rel\": " << clus.worst
     << ", \"rng_equal\": " << rng_eq << ", \"verification_equal\": " << verify_eq
     << ", \"side_a\": " << b.clusters.verification.side_a.indices.size()
     << ", \"metrics\": [";
  for (size_t i = 0; i < metrics.size(); ++i)
    js << (i ? ", " : "") << "{\"stage\": \"" << metrics[i].stage
       << "\", \"interval_ms\": " << metrics[i].gpu_interval_ms
       << ", \"kernel_ms\": " << metrics[i].custom_kernel_ms
       << ", \"launches\": " << metrics[i].custom_launches
       << ", \"h2d_bytes\": " << metrics[i].h2d_bytes
       << ", \"d2h_bytes\": " << metrics[i].d2h_bytes
       << ", \"h2d_ms\": " << metrics[i].h2d_ms << ", \"d2h_ms\": " << metrics[i].d2h_ms
       << ", \"peak_buffer_bytes\": " << metrics[i].peak_buffer_bytes
       << ", \"released\": " << (metrics[i].all_buffers_released ? "true" : "false") << "}";
  js << "]}";
  std::printf("%s\n", js.str().c_str());
  return ok ? 0 : 1;
}
