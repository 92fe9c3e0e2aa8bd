// mesh_io.hpp -- triangle mesh readers (OBJ, OFF, STL ascii/binary, PLY
// ascii/binary) for the offset3d_hip CLI.  They replace geogram's mesh_load
// on the reference's input path (src/vor3d/Dexelize.cpp:247): polygons are
// fan-triangulated, vertices kept as given.
#pragma once

#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace vo_mesh {

struct Mesh {
  std::vector<double> v;   // 3 per vertex
  std::vector<int32_t> f;  // 3 per triangle
  int64_t nv() const { return int64_t(v.size() / 3); }
  int64_t nf() const { return int64_t(f.size() / 3); }
};

inline void fan(Mesh& m, const std::vector<int64_t>& poly) {
  for (size_t i = 1; i + 1 < poly.size(); ++i) {
    m.f.push_back(int32_t(poly[0]));
    m.f.push_back(int32_t(poly[i]));
    m.f.push_back(int32_t(poly[i + 1]));
  }
}

inline std::string lower_ext(const std::string& path) {
  const size_t d = path.find_last_of('.');
  std::string e = d == std::string::npos ? "" : path.substr(d);
  for (auto& c : e) c = char(std::tolower(c));
  return e;
}

inline Mesh load_obj(const std::string& path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("cannot open " + path);
  Mesh m;
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream s(line);
    std::string t;
    if (!(s >> t)) continue;
    if (t == "v") {
      double x, y, z;
      s >> x >> y >> z;
      m.v.insert(m.v.end(), {x, y, z});
    } else if (t == "f") {
      std::vector<int64_t> poly;
      std::string w;
      while (s >> w) {
        const int64_t i = std::stoll(w.substr(0, w.find('/')));
        poly.push_back(i > 0 ? i - 1 : m.nv() + i);
      }
      fan(m, poly);
    }
  }
  return m;
}

inline Mesh load_off(const std::string& path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("cannot open " + path);
  std::stringstream all;
  std::string line;
  while (std::getline(in, line)) all << line.substr(0, line.find('#')) << ' ';
  std::string head;
  all >> head;
  if (head.size() < 3 || head.compare(head.size() - 3, 3, "OFF") != 0) throw std::runtime_error("not an OFF file");
  int64_t nv, nf, ne;
  all >> nv >> nf >> ne;
  Mesh m;
  m.v.resize(size_t(3 * nv));
  for (auto& x : m.v) all >> x;
  for (int64_t i = 0; i < nf; ++i) {
    int64_t n = -1;
    all >> n;
    if (n < 0 || n > (int64_t(1) << 20)) throw std::runtime_error("bad face in OFF file");
    std::vector<int64_t> poly(static_cast<size_t>(n));
    for (auto& p : poly) all >> p;
    fan(m, poly);
  }
  if (!all && !all.eof()) throw std::runtime_error("truncated OFF file");
  return m;
}

inline Mesh load_stl(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  if (!in) throw std::runtime_error("cannot open " + path);
  std::string data((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  Mesh m;
  if (data.size() >= 84) {
    uint32_t n;
    std::memcpy(&n, data.data() + 80, 4);
    if (84 + 50ull * n == data.size()) {
      for (uint32_t t = 0; t < n; ++t) {
        const char* r = data.data() + 84 + 50ull * t + 12;
        for (int k = 0; k < 9; ++k) {
          float x;
          std::memcpy(&x, r + 4 * k, 4);
          m.v.push_back(double(x));
        }
        for (int k = 0; k < 3; ++k) m.f.push_back(int32_t(3 * t + k));
      }
      return m;
    }
  }
  std::istringstream s(data);
  std::string t;
  while (s >> t) {
    if (t == "vertex") {
      double x, y, z;
      s >> x >> y >> z;
      m.v.insert(m.v.end(), {x, y, z});
    }
  }
  for (int64_t i = 0; i < m.nv(); ++i) m.f.push_back(int32_t(i));
  m.f.resize(m.f.size() / 3 * 3);
  return m;
}

namespace ply {
inline size_t type_size(const std::string& t) {
  if (t == "char" || t == "int8" || t == "uchar" || t == "uint8") return 1;
  if (t == "short" || t == "int16" || t == "ushort" || t == "uint16") return 2;
  if (t == "int" || t == "int32" || t == "uint" || t == "uint32" || t == "float" || t == "float32") return 4;
  if (t == "double" || t == "float64") return 8;
  throw std::runtime_error("PLY type " + t);
}
inline double read_bin(const char* p, const std::string& t, bool swap) {
  char b[8];
  const size_t n = type_size(t);
  for (size_t i = 0; i < n; ++i) b[i] = swap ? p[n - 1 - i] : p[i];
  if (t == "char" || t == "int8") return double(*reinterpret_cast<int8_t*>(b));
  if (t == "uchar" || t == "uint8") return double(*reinterpret_cast<uint8_t*>(b));
  if (t == "short" || t == "int16") { int16_t v; std::memcpy(&v, b, 2); return v; }
  if (t == "ushort" || t == "uint16") { uint16_t v; std::memcpy(&v, b, 2); return v; }
  if (t == "int" || t == "int32") { int32_t v; std::memcpy(&v, b, 4); return v; }
  if (t == "uint" || t == "uint32") { uint32_t v; std::memcpy(&v, b, 4); return v; }
  if (t == "float" || t == "float32") { float v; std::memcpy(&v, b, 4); return v; }
  double v;
  std::memcpy(&v, b, 8);
  return v;
}
struct Prop {
  std::string type, count_type, item_type, name;
  bool list = false;
};
struct Element {
  std::string name;
  int64_t count = 0;
  std::vector<Prop> props;
};
}  // namespace ply

inline Mesh load_ply(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  if (!in) throw std::runtime_error("cannot open " + path);
  std::string data((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  const size_t h = data.find("end_header");
  if (data.compare(0, 3, "ply") != 0 || h == std::string::npos) throw std::runtime_error("not a PLY file");
  const size_t body = data.find('\n', h) + 1;
  std::istringstream hs(data.substr(0, body));
  std::string line, fmt;
  std::vector<ply::Element> els;
  while (std::getline(hs, line)) {
    std::istringstream s(line);
    std::string t;
    s >> t;
    if (t == "format") {
      s >> fmt;
    } else if (t == "element") {
      ply::Element e;
      s >> e.name >> e.count;
      els.push_back(e);
    } else if (t == "property" && !els.empty()) {
      ply::Prop p;
      s >> p.type;
      if (p.type == "list") {
        p.list = true;
        s >> p.count_type >> p.item_type >> p.name;
      } else {
        s >> p.name;
      }
      els.back().props.push_back(p);
    }
  }
  const bool ascii = fmt == "ascii";
  const bool swap = fmt == "binary_big_endian";
  if (!ascii && fmt != "binary_little_endian" && !swap) throw std::runtime_error("PLY format " + fmt);
  Mesh m;
  std::istringstream as(ascii ? data.substr(body) : std::string());
  size_t k = body;
  auto scalar = [&](const std::string& type) -> double {
    if (ascii) {
      double v;
      as >> v;
      return v;
    }
    if (k + ply::type_size(type) > data.size()) throw std::runtime_error("truncated PLY file");
    const double v = ply::read_bin(data.data() + k, type, swap);
    k += ply::type_size(type);
    return v;
  };
  for (const auto& e : els) {
    for (int64_t i = 0; i < e.count; ++i) {
      double xyz[3] = {0, 0, 0};
      for (const auto& p : e.props) {
        if (p.list) {
          const int64_t n = int64_t(scalar(p.count_type));
          if (n < 0 || n > (int64_t(1) << 20)) throw std::runtime_error("bad face in PLY file");
          std::vector<int64_t> poly(static_cast<size_t>(n));
          for (auto& q : poly) q = int64_t(scalar(p.item_type));
          if (e.name == "face") fan(m, poly);
        } else {
          const double v = scalar(p.type);
          if (p.name == "x") xyz[0] = v;
          if (p.name == "y") xyz[1] = v;
          if (p.name == "z") xyz[2] = v;
        }
      }
      if (e.name == "vertex") m.v.insert(m.v.end(), xyz, xyz + 3);
    }
  }
  return m;
}

inline bool is_mesh_path(const std::string& path) {
  const std::string e = lower_ext(path);
  return e == ".obj" || e == ".off" || e == ".stl" || e == ".ply";
}

inline Mesh load(const std::string& path) {
  const std::string e = lower_ext(path);
  Mesh m;
  if (e == ".obj") m = load_obj(path);
  else if (e == ".off") m = load_off(path);
  else if (e == ".stl") m = load_stl(path);
  else if (e == ".ply") m = load_ply(path);
  else throw std::runtime_error("unsupported mesh format " + e);
  for (int32_t i : m.f)
    if (i < 0 || i >= m.nv()) throw std::runtime_error("triangle index out of range in " + path);
  return m;
}

}  // namespace vo_mesh
