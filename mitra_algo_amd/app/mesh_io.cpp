// OFF / OBJ reader (src/io/mesh.cpp:10-98): the same header limits, polygon
// rules, index conventions, checks and messages.
#include "mesh_io.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace mvs_app {
namespace {

void add_fan(Mesh &m, const std::vector<long long> &poly) {
  for (size_t i = 1; i + 1 < poly.size(); ++i) {
    m.f0.push_back(static_cast<int32_t>(poly[0]));
    m.f1.push_back(static_cast<int32_t>(poly[i]));
    m.f2.push_back(static_cast<int32_t>(poly[i + 1]));
  }
}

std::string lower_extension(const std::string &path) {
  std::string e = std::filesystem::path(path).extension().string();
  for (char &c : e)
    c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
  return e;
}

void read_off(std::istream &in, Mesh &m) {
  // '#' starts a comment anywhere on a line, so the counts may share the
  // magic word's line
  std::string line, text;
  while (std::getline(in, line))
    text += line.substr(0, line.find('#')) + '\n';
  std::istringstream s(text);
  std::string magic;
  long long n = 0, f = 0, e = 0;
  s >> magic >> n >> f >> e;
  if (!s || magic != "OFF" || n < 3 || f < 1 || n > 10000000 || f > 20000000)
    throw std::runtime_error("Invalid OFF header");
  m.x.resize(n);
  m.y.resize(n);
  m.z.resize(n);
  for (long long i = 0; i < n; ++i)
    if (!(s >> m.x[i] >> m.y[i] >> m.z[i]))
      throw std::runtime_error("Truncated OFF vertices");
  std::vector<long long> poly;
  for (long long i = 0; i < f; ++i) {
    int count = 0;
    if (!(s >> count) || count < 3 || count > 1000000)
      throw std::runtime_error("Invalid OFF polygon");
    poly.assign(count, 0);
    for (auto &v : poly) {
      int id;
      if (!(s >> id))
        throw std::runtime_error("Truncated OFF face");
      v = id;
    }
    add_fan(m, poly);
  }
}

void read_obj(std::istream &in, Mesh &m) {
  std::string line;
  std::vector<long long> poly;
  while (std::getline(in, line)) {
    std::istringstream s(line);
    std::string tag;
    s >> tag;
    if (tag == "v") {
      double a, b, c;
      if (!(s >> a >> b >> c))
        throw std::runtime_error("Invalid OBJ vertex");
      m.x.push_back(a);
      m.y.push_back(b);
      m.z.push_back(c);
    } else if (tag == "f") {
      poly.clear();
      std::string token;
      while (s >> token) {
        if (token[0] == '#')
          break;
        token = token.substr(0, token.find('/'));
        size_t used = 0;
        const long long i = std::stoll(token, &used);
        if (used != token.size() || i == 0)
          throw std::runtime_error("Invalid OBJ index");
        const long long nv = static_cast<long long>(m.x.size());
        const long long id = i > 0 ? i - 1 : nv + i;
        if (id < 0 || id >= nv)
          throw std::runtime_error("OBJ index out of range");
        poly.push_back(id);
      }
      if (poly.size() < 3)
        throw std::runtime_error("Invalid OBJ face");
      add_fan(m, poly);
    }
  }
}

}  // namespace

Mesh read_mesh(const std::string &path) {
  std::ifstream in(path);
  if (!in)
    throw std::runtime_error("Cannot open mesh: " + path);
  Mesh m;
  const std::string ext = lower_extension(path);
  if (ext == ".off")
    read_off(in, m);
  else if (ext == ".obj")
    read_obj(in, m);
  else
    throw std::runtime_error("Supported mesh formats: OFF and OBJ");
  if (m.x.empty() || m.f0.empty())
    throw std::runtime_error("Mesh is empty");
  for (size_t i = 0; i < m.x.size(); ++i)
    if (!std::isfinite(m.x[i]) || !std::isfinite(m.y[i]) || !std::isfinite(m.z[i]))
      throw std::runtime_error("Non-finite vertex");
  const long long nv = static_cast<long long>(m.x.size());
  for (size_t i = 0; i < m.f0.size(); ++i)
    for (int32_t v : {m.f0[i], m.f1[i], m.f2[i]})
      if (v < 0 || v >= nv)
        throw std::runtime_error("Face index out of range");
  return m;
}

}  // namespace mvs_app
