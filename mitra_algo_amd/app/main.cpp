// mvs_cli: the detector's command line on libmvs, with the options, checks,
// messages, exit codes and output directory of the reference's CLI
// (src/app/main.cpp:35-160).  The GPU backend is the only one here:
// --backend hip (default), with cuda accepted as its alias so scripts
// written for the reference's GPU backend run unchanged.
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <iostream>
#include <limits>
#include <stdexcept>
#include <string>

#include "detect.hpp"
#include "export.hpp"
#include "mesh_io.hpp"

namespace {

void help() {
  std::cout << R"(Partial symmetry detector (rigid motion + plane reflection, scale=1)
Usage: mvs_cli MODEL.off|MODEL.obj --output DIRECTORY [options]
  --backend hip|cuda    default hip (MI355X / gfx950); cuda is an alias
  --device N            GPU index, default 0
  --sample-ratio X      (0,1], default 0.5, area-weighted vertex sampling
  --prune X             (0,1), default 0.85, larger retains more vertices
  --pair-threshold X    (0,1], default 0.15, curvature tolerance
  --bandwidth X         (0,2], default 0.08, normalized mean-shift radius
  --verify-threshold X  (0,0.5), default 0.025, fraction of mesh box diagonal
  --smoothing N         0..1000, default 3, tensor smoothing iterations
  --subdivide N         0..3, default 0, uniform midpoint subdivision
  --seed N              default 42, reproducible sampling and mean shift
  --max-iterations N    1..10000, default 100 per mean-shift trajectory
  --max-clusters N      1..1000, default 20 exported/verified hypotheses
  --max-pairs N         default 200000, fail explicitly at resource limit
  --max-votes N         default 600000, fail explicitly at resource limit
  --no-patching         export hypotheses without geometric verification
  --help
Requires consistently oriented manifold triangles. Boundaries are supported.
Polygon faces are fan-triangulated (use convex polygons or triangles).
Output directory must be new or empty.
)";
}

double parse_number(const std::string &flag, const std::string &value) {
  size_t used = 0;
  const double x = std::stod(value, &used);
  if (used != value.size() || !std::isfinite(x))
    throw std::runtime_error("Invalid numeric value for " + flag);
  return x;
}

int parse_int(const std::string &flag, const std::string &value) {
  const double x = parse_number(flag, value);
  if (x < 0 || x > std::numeric_limits<int>::max() || std::floor(x) != x)
    throw std::runtime_error("Invalid integer for " + flag);
  return static_cast<int>(x);
}

unsigned long long parse_unsigned(const std::string &flag, const std::string &value) {
  if (value.empty() || value[0] == '-')
    throw std::runtime_error("Invalid unsigned value for " + flag);
  size_t used = 0;
  const unsigned long long x = std::stoull(value, &used);
  if (used != value.size())
    throw std::runtime_error("Invalid unsigned value for " + flag);
  return x;
}

}  // namespace

int main(int argc, char **argv) {
  try {
    if (argc == 1) {
      help();
      return 2;
    }
    mvs_app::Options o;
    std::string backend = "hip", input, out;
    int device = 0;
    for (int i = 1; i < argc; ++i) {
      const std::string a = argv[i];
      if (a == "--help") {
        help();
        return 0;
      }
      if (a == "--no-patching") {
        o.patching = false;
        continue;
      }
      if (a.rfind("--", 0) != 0) {
        if (!input.empty())
          throw std::runtime_error("Only one input mesh is allowed");
        input = a;
        continue;
      }
      if (i + 1 >= argc)
        throw std::runtime_error("Missing value for " + a);
      const std::string v = argv[++i];
      if (a == "--output")
        out = v;
      else if (a == "--backend")
        backend = v;
      else if (a == "--device")
        device = parse_int(a, v);
      else if (a == "--sample-ratio")
        o.sample_ratio = parse_number(a, v);
      else if (a == "--prune")
        o.prune = parse_number(a, v);
      else if (a == "--pair-threshold")
        o.pair_threshold = parse_number(a, v);
      else if (a == "--bandwidth")
        o.bandwidth = parse_number(a, v);
      else if (a == "--verify-threshold")
        o.verify_threshold = parse_number(a, v);
      else if (a == "--smoothing")
        o.smoothing = parse_int(a, v);
      else if (a == "--subdivide")
        o.subdivide = parse_int(a, v);
      else if (a == "--seed")
        o.seed = parse_unsigned(a, v);
      else if (a == "--max-iterations")
        o.max_iterations = parse_int(a, v);
      else if (a == "--max-clusters")
        o.max_clusters = parse_int(a, v);
      else if (a == "--max-pairs")
        o.max_pairs = parse_unsigned(a, v);
      else if (a == "--max-votes")
        o.max_votes = parse_unsigned(a, v);
      else
        throw std::runtime_error("Unknown option: " + a);
    }
    if (input.empty() || out.empty())
      throw std::runtime_error("Input and --output are required");
    if (std::filesystem::exists(out) &&
        (!std::filesystem::is_directory(out) || !std::filesystem::is_empty(out)))
      throw std::runtime_error("Output directory must be new or empty "
                               "(existing results are preserved)");
    if (backend == "cpu")
      throw std::runtime_error("CPU backend unavailable: this build runs on the GPU "
                               "(--backend hip)");
    if (backend != "hip" && backend != "cuda")
      throw std::runtime_error("Unknown backend: " + backend);
    std::cout << "backend=hip" << std::endl;
    mvs_app::Result r =
        mvs_app::detect(mvs_app::read_mesh(input), o, device,
                        [](const std::string &stage, double ms) {
                          std::cout << stage << ": " << ms << " ms" << std::endl;
                        });
    mvs_app::export_result(r, out);
    std::cout << "vertices=" << r.mesh.vertices() << " sample=" << r.sample.size()
              << " pruned=" << r.pruned.size() << " pairs=" << r.pair_a.size()
              << " votes=" << r.vote_pair.size() << " clusters=" << r.clusters() << "\n";
    for (size_t c = 0; c < r.clusters(); ++c)
      std::cout << c << " " << (r.c_reflection[c] ? "reflection" : "rigid")
                << " support=" << r.c_member_offsets[c + 1] - r.c_member_offsets[c]
                << " coverage=" << r.coverage[c] << " rms=" << r.rms[c] << "\n";
    if (r.limited_iterations)
      std::cout << "Warning: " << r.limited_iterations
                << " mean-shift trajectories reached the iteration limit. "
                   "Increase --max-iterations if needed.\n";
    std::cout << "Results: " << std::filesystem::path(out) << "\n";
    return 0;
  } catch (const std::exception &e) {
    std::cerr << "Error: " << e.what() << '\n';
    return 1;
  }
}
