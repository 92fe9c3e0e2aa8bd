// offset3d.cpp -- the offset3d CLI (app/cli3d/offset3d.cpp) on the GPU path.
//
//   offset3d_hip input [-o output] [-j json] [-d dexel_size] [-n num_dexels]
//                [-p padding] [-r radius] [-m hip|hip64] [-x op] [-u] [-f] [-t n]
//
// Same options and defaults as the reference (offset3d.cpp:20-51): the input
// is a mesh (.obj .off .stl .ply, rasterized on the GPU as create_dexels does,
// Dexelize.cpp:231-274) or a dexel volume in CompressedVolume's text format;
// -x applies noop / dilation / erosion / closing / opening (offset3d.cpp:116-136)
// with radius -r in dexels (-u: in mm, :73-75).  -m hip stores endpoints as
// FP32 on the device, hip64 as FP64 (predicates are FP64 in both).  Output:
// .obj -> one hexahedron (8 vertices, 6 quads) per segment, as dexel_dump
// (Dexelize.cpp:312-351); .xyz -> the two endpoints of every segment at the
// dexel centre (points_dump, :287-310); anything else -> the volume in text
// form.  -j writes the reference's json keys.  -t is accepted and ignored.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/voroffset_hip.h"
#include "mesh_io.hpp"

namespace {

struct Volume {
  double origin[3] = {0, 0, 0}, extent[3] = {0, 0, 0};
  double spacing = 1;
  int padding = 0;
  int nx = 0, ny = 0;
  std::vector<int64_t> off{0};
  std::vector<double> z;
  int64_t segments() const {
    int64_t s = 0;
    for (size_t c = 0; c + 1 < off.size(); ++c) s += (off[c + 1] - off[c]) / 2;
    return s;
  }
};

[[noreturn]] void die(const std::string& m) {
  std::cerr << "offset3d_hip: " << m << std::endl;
  std::exit(1);
}

Volume load_volume(const std::string& path) {
  std::ifstream in(path);
  if (!in) die("cannot open " + path);
  Volume v;
  in >> v.origin[0] >> v.origin[1] >> v.origin[2] >> v.extent[0] >> v.extent[1] >> v.extent[2] >> v.nx >> v.ny >>
      v.padding >> v.spacing;
  if (!in || v.nx < 0 || v.ny < 0) die("bad volume header in " + path);
  v.off.assign(size_t(v.nx) * v.ny + 1, 0);
  for (int64_t c = 0; c < int64_t(v.nx) * v.ny; ++c) {
    int64_t n;
    in >> n;
    if (!in || n < 0) die("bad column in " + path);
    for (int64_t i = 0; i < n; ++i) {
      double t;
      in >> t;
      v.z.push_back(t);
    }
    v.off[c + 1] = int64_t(v.z.size());
  }
  return v;
}

void save_volume(const std::string& path, const Volume& v) {
  std::ofstream o(path);
  o << std::setprecision(17);
  o << v.origin[0] << ' ' << v.origin[1] << ' ' << v.origin[2] << '\n'
    << v.extent[0] << ' ' << v.extent[1] << ' ' << v.extent[2] << '\n'
    << v.nx << ' ' << v.ny << '\n' << v.padding << '\n' << v.spacing << '\n';
  for (size_t c = 0; c + 1 < v.off.size(); ++c) {
    o << (v.off[c + 1] - v.off[c]);
    for (int64_t i = v.off[c]; i < v.off[c + 1]; ++i) o << ' ' << v.z[i];
    o << '\n';
  }
}

// dexel_dump (Dexelize.cpp:312-351): the hexahedra come from the device
// (vo_dexel_mesh); this writes them as OBJ quads.
void dump_obj(vo_context* ctx, const std::string& path, const Volume& v) {
  vo_volume hv{v.nx, v.ny, v.off.data(), v.z.empty() ? nullptr : v.z.data(), 0.0, 0.0};
  vo_mesh_result* r = nullptr;
  if (vo_dexel_mesh(ctx, &hv, v.origin[0], v.origin[1], v.spacing, &r) != VO_OK)
    die(std::string("dexel mesh: ") + vo_last_error(ctx));
  int64_t nv = 0, nq = 0;
  vo_mesh_shape(r, &nv, &nq);
  std::vector<double> vert(size_t(3 * nv));
  std::vector<int64_t> quad(size_t(4 * nq));
  vo_mesh_copy(r, vert.data(), quad.data());
  vo_mesh_free(r);
  std::ofstream o(path);
  o << std::setprecision(17);
  for (int64_t i = 0; i < nv; ++i) o << "v " << vert[3 * i] << ' ' << vert[3 * i + 1] << ' ' << vert[3 * i + 2] << '\n';
  for (int64_t i = 0; i < nq; ++i)
    o << "f " << quad[4 * i] + 1 << ' ' << quad[4 * i + 1] + 1 << ' ' << quad[4 * i + 2] + 1 << ' ' << quad[4 * i + 3] + 1
      << '\n';
}

void dump_xyz(const std::string& path, const Volume& v) {
  std::ofstream o(path);
  o << std::setprecision(17);
  for (int y = 0; y < v.ny; ++y)
    for (int x = 0; x < v.nx; ++x) {
      const int64_t c = x + int64_t(v.nx) * y;
      const double cx = v.origin[0] + (x + 0.5) * v.spacing, cy = v.origin[1] + (y + 0.5) * v.spacing;
      for (int64_t i = v.off[c]; i + 1 < v.off[c + 1]; i += 2) {
        o << cx << ' ' << cy << ' ' << v.z[i] * v.spacing << '\n';
        o << cx << ' ' << cy << ' ' << v.z[i + 1] * v.spacing << '\n';
      }
    }
}

void take(vo_result* r, Volume& v) {
  int64_t count = 0;
  vo_result_shape(r, &v.nx, &v.ny, &count);
  v.off.assign(size_t(v.nx) * v.ny + 1, 0);
  v.z.assign(size_t(count), 0.0);
  vo_result_copy(r, v.off.data(), v.z.empty() ? nullptr : v.z.data());
  vo_result_free(r);
}

// vor3d::create_dexels (Dexelize.cpp:231-274) with the CompressedVolume
// constructor's geometry (CompressedVolume.cpp:11-23).
Volume dexelize(vo_context* ctx, const vo_mesh::Mesh& m, double& voxel, int padding, int num_dexels) {
  if (m.nv() == 0) die("empty mesh");
  double lo[3], hi[3];
  for (int d = 0; d < 3; ++d) lo[d] = hi[d] = m.v[d];
  for (int64_t i = 0; i < m.nv(); ++i)
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], m.v[3 * i + d]);
      hi[d] = std::max(hi[d], m.v[3 * i + d]);
    }
  Volume v;
  for (int d = 0; d < 3; ++d) v.extent[d] = hi[d] - lo[d];
  if (num_dexels > 0) voxel = std::max(v.extent[0], std::max(v.extent[1], v.extent[2])) / num_dexels;
  if (!(voxel > 0)) die("dexel size must be > 0");
  v.spacing = voxel;
  v.padding = padding;
  for (int d = 0; d < 3; ++d) v.origin[d] = lo[d] - padding * voxel;
  v.nx = int(std::ceil(v.extent[0] / voxel) + 2 * padding);
  v.ny = int(std::ceil(v.extent[1] / voxel) + 2 * padding);
  vo_result* r = nullptr;
  if (vo_rasterize(ctx, m.v.data(), m.nv(), m.f.data(), m.nf(), v.nx, v.ny, v.origin[0], v.origin[1], voxel, &r) !=
      VO_OK)
    die(std::string("rasterize: ") + vo_last_error(ctx));
  take(r, v);
  return v;
}

bool exists(const std::string& p) { return bool(std::ifstream(p)); }

std::string json_escape(const std::string& s) {
  std::string o;
  for (char c : s) {
    if (c == '"' || c == '\\') o += '\\';
    o += c;
  }
  return o;
}

}  // namespace

int main(int argc, char** argv) {
  struct {
    std::string input, output_mesh = "output.obj", output_json, method = "hip", operation = "dilation";
    double radius = 8, dexels_size = 1;
    int padding = 0, num_dexels = 256;
    bool force = false, radius_in_mm = false;
  } a;
  std::vector<std::string> args(argv + 1, argv + argc);
  auto value = [&](size_t& i) -> std::string {
    if (i + 1 >= args.size()) die("missing value for " + args[i]);
    return args[++i];
  };
  for (size_t i = 0; i < args.size(); ++i) {
    const std::string& s = args[i];
    if (s == "-h" || s == "--help") {
      std::cout << "usage: offset3d_hip input [-o output] [-j json] [-d dexel_size] [-n num_dexels] [-p padding]\n"
                   "                    [-r radius] [-m hip|hip64] [-x noop|dilation|erosion|closing|opening]\n"
                   "                    [-u] [-f] [-t threads]\n";
      return 0;
    } else if (s == "-i" || s == "--input") a.input = value(i);
    else if (s == "-o" || s == "--output") a.output_mesh = value(i);
    else if (s == "-j" || s == "--json") a.output_json = value(i);
    else if (s == "-d" || s == "--dexels_size") a.dexels_size = std::atof(value(i).c_str());
    else if (s == "-n" || s == "--num_dexels") a.num_dexels = std::atoi(value(i).c_str());
    else if (s == "-p" || s == "--padding") a.padding = std::atoi(value(i).c_str());
    else if (s == "-t" || s == "--num_thread") value(i);
    else if (s == "-r" || s == "--radius") a.radius = std::atof(value(i).c_str());
    else if (s == "-m" || s == "--method") a.method = value(i);
    else if (s == "-x" || s == "--apply") a.operation = value(i);
    else if (s == "-f" || s == "--force") a.force = true;
    else if (s == "-u" || s == "--radius_in_mm") a.radius_in_mm = true;
    else if (!s.empty() && s[0] == '-') die("unknown option " + s);
    else a.input = s;
  }
  if (a.input.empty() || !exists(a.input)) die("input file required (and must exist)");
  int precision;
  if (a.method == "hip") precision = VO_FP32;
  else if (a.method == "hip64") precision = VO_FP64;
  else die("invalid method: " + a.method + " (hip, hip64)");
  int op;
  if (a.operation == "noop") op = VO_OP_NOOP;
  else if (a.operation == "dilation") op = VO_OP_DILATION;
  else if (a.operation == "erosion") op = VO_OP_EROSION;
  else if (a.operation == "closing") op = VO_OP_CLOSING;
  else if (a.operation == "opening") op = VO_OP_OPENING;
  else die("invalid operation: " + a.operation);

  vo_context* ctx = nullptr;
  if (vo_create(&ctx, 0) != VO_OK) die("no HIP device");
  const auto t0 = std::chrono::steady_clock::now();
  Volume in;
  try {
    if (vo_mesh::is_mesh_path(a.input)) {
      in = dexelize(ctx, vo_mesh::load(a.input), a.dexels_size, a.padding, a.num_dexels);
    } else {
      in = load_volume(a.input);
      a.dexels_size = in.spacing;
    }
  } catch (const std::exception& e) {
    die(e.what());
  }
  if (a.radius_in_mm) a.radius /= in.spacing;
  std::cout << "Grid size: " << in.nx << " " << in.ny << "\nSpacing (in mm): " << in.spacing
            << "\nOrigin (in mm): " << in.origin[0] << " " << in.origin[1] << " " << in.origin[2]
            << "\nExtent (in mm): " << in.extent[0] << " " << in.extent[1] << " " << in.extent[2]
            << "\nRadius (in #dexels): " << a.radius << std::endl;

  Volume out = in;
  double t1 = 0, t2 = 0;
  if (op != VO_OP_NOOP) {
    const double z_lo = in.origin[2] / in.spacing;
    const double z_hi = z_lo + 2 * in.padding + in.extent[2] / in.spacing;
    vo_volume v{in.nx, in.ny, in.off.data(), in.z.empty() ? nullptr : in.z.data(), z_lo, z_hi};
    vo_result* r = nullptr;
    if (vo_morph(ctx, op, precision, a.radius, &v, &r, &t1, &t2) != VO_OK)
      die(std::string("morphology: ") + vo_last_error(ctx));
    take(r, out);
  }
  const double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

  if (!a.output_mesh.empty()) {
    if (exists(a.output_mesh) && !a.force) {
      std::cout << "Output mesh already exists. Please use -f to force overwriting." << std::endl;
    } else {
      const std::string e = vo_mesh::lower_ext(a.output_mesh);
      if (e == ".obj") dump_obj(ctx, a.output_mesh, out);
      else if (e == ".xyz") dump_xyz(a.output_mesh, out);
      else save_volume(a.output_mesh, out);
    }
  }
  if (!a.output_json.empty()) {
    if (exists(a.output_json) && !a.force) {
      std::cout << "Output json already exists. Please use -f to force overwriting." << std::endl;
    } else {
      std::ofstream j(a.output_json);
      j << std::setprecision(17) << "{\n"
        << "    \"method\": \"" << a.method << "\",\n"
        << "    \"model_name\": \"" << json_escape(a.input) << "\",\n"
        << "    \"voxel_size\": " << in.spacing << ",\n"
        << "    \"padding\": " << in.padding << ",\n"
        << "    \"num_dexels\": " << a.num_dexels << ",\n"
        << "    \"radius\": " << a.radius << ",\n"
        << "    \"grid_size\": [" << in.nx << ", " << in.ny << "],\n"
        << "    \"num_segments\": " << in.segments() << ",\n"
        << "    \"operation\": \"" << a.operation << "\",\n"
        << "    \"num_segments_out\": " << out.segments() << ",\n"
        << "    \"time\": " << total * 1e3 << ",\n"
        << "    \"time_first_pass\": " << t1 << ",\n"
        << "    \"time_second_pass\": " << t2 << "\n}\n";
    }
  }
  vo_destroy(ctx);
  return 0;
}
