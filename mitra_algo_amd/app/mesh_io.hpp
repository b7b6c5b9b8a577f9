// Mesh files of the detector: OFF and OBJ input, with the reference's
// acceptance rules and messages (src/io/mesh.cpp:10-98).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace mvs_app {

struct Mesh {
  std::vector<double> x, y, z;
  std::vector<int32_t> f0, f1, f2;
  size_t vertices() const { return x.size(); }
  size_t faces() const { return f0.size(); }
};

// Reads an .off or .obj file (extension compared case-blind).  Polygons are
// fan-triangulated from their first corner; OBJ indices are 1-based, negative
// ones count back from the last vertex read, and "v/t/n" keeps the vertex.
// Throws std::runtime_error with the reference's messages.
Mesh read_mesh(const std::string &path);

}  // namespace mvs_app
