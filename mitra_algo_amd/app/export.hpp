// Writes a detection result as the reference's output directory
// (src/io/export.cpp:37-189).
#pragma once
#include <string>

#include "detect.hpp"

namespace mvs_app {

// Creates `directory` (and parents) and writes features.csv, pairs.csv,
// votes.csv, descriptors.csv, mesh.off and result.json.
void export_result(const Result &r, const std::string &directory);

}  // namespace mvs_app
