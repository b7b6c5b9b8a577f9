// Result files (src/io/export.cpp:37-189): features.csv, pairs.csv,
// votes.csv, descriptors.csv, mesh.off and result.json (schema version 1),
// numbers at 17 significant digits, in the reference's column order and
// spelling so that its readers and viewers take them unchanged.
#include "export.hpp"

#include <filesystem>
#include <fstream>
#include <iomanip>
#include <stdexcept>

namespace mvs_app {
namespace {

std::ofstream open_out(const std::filesystem::path &p) {
  std::ofstream out(p);
  if (!out)
    throw std::runtime_error("Cannot write " + p.string());
  out << std::setprecision(17);
  return out;
}

template <class T> void json_list(std::ostream &s, const T *a, size_t n) {
  s << '[';
  for (size_t i = 0; i < n; ++i) {
    if (i)
      s << ',';
    s << a[i];
  }
  s << ']';
}

void json_vec(std::ostream &s, const double *v) {
  s << '[' << v[0] << ',' << v[1] << ',' << v[2] << ']';
}

// Patches of cluster c on one side: a list of id lists.
void json_patches(std::ostream &s, const Side &sd, size_t c) {
  s << '[';
  for (uint64_t g = sd.groups[c]; g < sd.groups[c + 1]; ++g) {
    if (g != sd.groups[c])
      s << ',';
    const uint64_t b = sd.patch_offsets[g], e = sd.patch_offsets[g + 1];
    json_list(s, sd.patch_members.data() + b, e - b);
  }
  s << ']';
}

}  // namespace

void export_result(const Result &r, const std::string &directory) {
  const std::filesystem::path dir(directory);
  std::filesystem::create_directories(dir);
  const Mesh &m = r.mesh;
  const size_t nv = m.vertices(), nf = m.faces(), P = r.pair_a.size(), V = r.vote_pair.size(),
               C = r.clusters();

  auto f = open_out(dir / "features.csv");
  f << "vertex,x,y,z,kmin,kmax,nx,ny,nz,dminx,dminy,dminz,dmaxx,dmaxy,dmaxz,"
       "valid,sampled,pruned\n";
  std::vector<bool> sampled(nv), pruned(nv);
  for (int32_t i : r.sample)
    sampled[i] = true;
  for (int32_t i : r.pruned)
    pruned[i] = true;
  for (size_t i = 0; i < r.valid.size(); ++i) {
    f << i << ',' << m.x[i] << ',' << m.y[i] << ',' << m.z[i] << ',' << r.feat[9][i] << ','
      << r.feat[10][i];
    for (int k = 0; k < 9; ++k)
      f << ',' << r.feat[k][i];
    f << ',' << (r.valid[i] != 0) << ',' << sampled[i] << ',' << pruned[i] << '\n';
  }

  auto p = open_out(dir / "pairs.csv");
  p << "pair,a,b\n";
  for (size_t i = 0; i < P; ++i)
    p << i << ',' << r.pair_a[i] << ',' << r.pair_b[i] << '\n';

  auto v = open_out(dir / "votes.csv");
  v << "vote,pair,type,r00,r01,r02,r10,r11,r12,r20,r21,r22,tx,ty,tz,nx,ny,nz,d,"
       "cluster\n";
  std::vector<int> cluster_of(V, -1);
  for (size_t c = 0; c < C; ++c)
    for (uint64_t j = r.c_member_offsets[c]; j < r.c_member_offsets[c + 1]; ++j)
      cluster_of[r.c_members[j]] = static_cast<int>(c);
  for (size_t i = 0; i < V; ++i) {
    v << i << ',' << r.vote_pair[i] << ',' << (r.vote_kind[i] ? "reflection" : "rigid");
    for (int k = 0; k < 16; ++k)
      v << ',' << r.vote_lanes[k][i];
    v << ',' << cluster_of[i] << '\n';
  }

  // rigid votes carry 12 descriptor components, reflections 4
  auto d = open_out(dir / "descriptors.csv");
  d << "vote,dimension";
  for (int k = 0; k < 12; ++k)
    d << ",component" << k;
  d << '\n';
  for (size_t i = 0; i < V; ++i) {
    const int dim = r.vote_kind[i] ? 4 : 12;
    d << i << ',' << dim;
    for (int k = 0; k < 12; ++k) {
      d << ',';
      if (k < dim)
        d << r.vote_lanes[16 + k][i];
    }
    d << '\n';
  }

  auto o = open_out(dir / "mesh.off");
  o << "OFF\n" << nv << ' ' << nf << " 0\n";
  for (size_t i = 0; i < nv; ++i)
    o << m.x[i] << ' ' << m.y[i] << ' ' << m.z[i] << '\n';
  for (size_t i = 0; i < nf; ++i)
    o << "3 " << m.f0[i] << ' ' << m.f1[i] << ' ' << m.f2[i] << '\n';

  const Options &op = r.options;
  auto s = open_out(dir / "result.json");
  s << "{\"schema_version\":1,\"mode\":\"rigid_and_plane_reflection\",\"box_size\":"
    << r.box_size << ",\"parameters\":{\"seed\":" << op.seed
    << ",\"sample_ratio\":" << op.sample_ratio << ",\"prune\":" << op.prune
    << ",\"pair_threshold\":" << op.pair_threshold << ",\"bandwidth\":" << op.bandwidth
    << ",\"verify_threshold\":" << op.verify_threshold << ",\"smoothing\":" << op.smoothing
    << ",\"subdivide\":" << op.subdivide << ",\"max_iterations\":" << op.max_iterations
    << ",\"max_clusters\":" << op.max_clusters << ",\"max_pairs\":" << op.max_pairs
    << ",\"max_votes\":" << op.max_votes
    << ",\"patching\":" << (op.patching ? "true" : "false") << "},\"counts\":{\"vertices\":"
    << nv << ",\"faces\":" << nf << ",\"pairs\":" << P << ",\"votes\":" << V
    << ",\"clusters\":" << C << "},\"diagnostics\":{\"rejected_faces\":" << r.rejected_faces
    << ",\"nonmanifold_edges\":" << r.nonmanifold_edges
    << ",\"orientation_conflicts\":" << r.orientation_conflicts
    << ",\"limited_iterations\":" << r.limited_iterations << "},\"timings_ms\":{";
  for (size_t i = 0; i < r.timings.size(); ++i) {
    if (i)
      s << ',';
    s << '"' << r.timings[i].first << "\":" << r.timings[i].second;
  }
  s << "},\"vertices\":[";
  for (size_t i = 0; i < nv; ++i) {
    if (i)
      s << ',';
    const double x[3] = {m.x[i], m.y[i], m.z[i]};
    json_vec(s, x);
  }
  s << "],\"faces\":[";
  for (size_t i = 0; i < nf; ++i) {
    if (i)
      s << ',';
    s << '[' << m.f0[i] << ',' << m.f1[i] << ',' << m.f2[i] << ']';
  }
  s << "],\"sample\":";
  json_list(s, r.sample.data(), r.sample.size());
  s << ",\"pruned\":";
  json_list(s, r.pruned.data(), r.pruned.size());
  s << ",\"pairs\":[";
  for (size_t i = 0; i < P; ++i) {
    if (i)
      s << ',';
    s << '[' << r.pair_a[i] << ',' << r.pair_b[i] << ']';
  }
  s << "],\"clusters\":[";
  for (size_t c = 0; c < C; ++c) {
    if (c)
      s << ',';
    s << "{\"type\":\"" << (r.c_reflection[c] ? "reflection" : "rigid") << "\",\"rotation\":";
    json_list(s, r.c_rotation.data() + 9 * c, 9);
    s << ",\"translation\":";
    json_vec(s, r.c_translation.data() + 3 * c);
    s << ",\"plane_normal\":";
    json_vec(s, r.c_plane_normal.data() + 3 * c);
    s << ",\"plane_offset\":" << r.c_plane_offset[c] << ",\"votes\":";
    const uint64_t b = r.c_member_offsets[c], e = r.c_member_offsets[c + 1];
    json_list(s, r.c_members.data() + b, e - b);
    s << ",\"pair_ids\":[";
    for (uint64_t j = b; j < e; ++j) {
      if (j != b)
        s << ',';
      s << r.vote_pair[r.c_members[j]];
    }
    s << "],\"coverage\":" << r.coverage[c] << ",\"rms\":" << r.rms[c] << ",\"side_a\":";
    const Side &A = r.side[0], &B = r.side[1];
    json_list(s, A.ids.data() + A.offsets[c], A.offsets[c + 1] - A.offsets[c]);
    s << ",\"side_b\":";
    json_list(s, B.ids.data() + B.offsets[c], B.offsets[c + 1] - B.offsets[c]);
    s << ",\"patches_a\":";
    json_patches(s, A, c);
    s << ",\"patches_b\":";
    json_patches(s, B, c);
    s << '}';
  }
  s << "]}\n";
  if (!f || !p || !v || !d || !o || !s)
    throw std::runtime_error("Output write failed");
}

}  // namespace mvs_app
