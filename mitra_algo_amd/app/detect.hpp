// The detector's pipeline over libmvs (src/core/pipeline.cpp:5-52):
// prepare -> curvature -> sample + prune -> pairing -> transformation votes
// -> mean-shift -> verification, with the reference's options, checks,
// random-number consumption and stage timings.
#pragma once
#include <cstdint>
#include <functional>
#include <string>
#include <utility>
#include <vector>

#include "mesh_io.hpp"

namespace mvs_app {

// Options with the reference's defaults (include/symmetry/options.h:5-11).
struct Options {
  double sample_ratio = .5, prune = .85, pair_threshold = .15, bandwidth = .08,
         verify_threshold = .025;
  int smoothing = 3, subdivide = 0, max_iterations = 100, max_clusters = 20;
  uint64_t seed = 42, max_pairs = 200000, max_votes = 600000;
  bool patching = true;
};

// validate_options (src/core/prepare.cpp:85-100): the same ranges and messages.
void validate_options(const Options &o);

// One verified side of every cluster: CSR over clusters, patches grouped.
struct Side {
  std::vector<uint64_t> offsets{0};      // clusters + 1
  std::vector<int32_t> ids;
  std::vector<uint64_t> groups{0};       // clusters + 1, into patch rows
  std::vector<uint64_t> patch_offsets{0};
  std::vector<int32_t> patch_members;
};

struct Result {
  Mesh mesh;  // after degenerate-face rejection and subdivision
  Options options;
  double box_size = 0, center[3] = {0, 0, 0};
  int rejected_faces = 0, nonmanifold_edges = 0, orientation_conflicts = 0,
      limited_iterations = 0;
  // features: lanes normal xyz, dmin xyz, dmax xyz, kmin, kmax; valid
  std::vector<double> feat[11];
  std::vector<uint8_t> valid;
  std::vector<int32_t> sample, pruned;
  std::vector<int32_t> pair_a, pair_b;
  // votes: pair id, kind (0 rigid, 1 reflection), 28 lanes (rotation 9,
  // translation 3, plane normal 3, plane offset, descriptor 12)
  std::vector<int32_t> vote_pair;
  std::vector<uint8_t> vote_kind;
  std::vector<double> vote_lanes[28];
  // clusters: records and member votes
  std::vector<uint8_t> c_reflection;
  std::vector<double> c_rotation, c_translation, c_plane_normal, c_plane_offset;
  std::vector<uint64_t> c_member_offsets{0};
  std::vector<int32_t> c_members;
  Side side[2];
  std::vector<double> coverage, rms;
  std::vector<std::pair<std::string, double>> timings;
  size_t clusters() const { return c_reflection.size(); }
};

using Progress = std::function<void(const std::string &, double)>;

// detect() on device `device`; throws std::invalid_argument,
// std::overflow_error or std::runtime_error with the reference's messages.
Result detect(Mesh mesh, const Options &options, int device, const Progress &progress = {});

}  // namespace mvs_app
