// voroffset_hip.hpp -- header-only C++ wrapper of the C ABI in voroffset_hip.h.
//
// Mirrors the calling convention of vor3d::VoronoiMorpho (src/vor3d/Voronoi.h:13-44):
// a method object whose dilation / erosion take a volume and a radius in
// dexels and report time_1 / time_2 in milliseconds, and which throws
// std::runtime_error on failure (as vor_assert does, src/vor3d/Common.cpp:8-18).
// Volumes are passed as CSR: offsets (nx*ny + 1, column x + nx*y) and sorted
// endpoints (the per-column vectors of vor3d::CompressedVolume, concatenated).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "voroffset_hip.h"

namespace vo_hip {

struct Csr {
  int32_t nx = 0, ny = 0;
  std::vector<int64_t> offsets;  // nx*ny + 1
  std::vector<double> z;         // endpoints, dexel units of absolute z
};

class Morpho {
 public:
  // precision: VO_FP32 (endpoints stored as float on the device) or VO_FP64
  explicit Morpho(int precision = VO_FP32, int device = 0) : precision_(precision) {
    if (vo_create(&ctx_, device) != VO_OK) throw std::runtime_error("vo_create failed on device " + std::to_string(device));
  }
  ~Morpho() { vo_destroy(ctx_); }
  Morpho(const Morpho&) = delete;
  Morpho& operator=(const Morpho&) = delete;

  // op: VO_OP_DILATION, VO_OP_EROSION, VO_OP_CLOSING, VO_OP_OPENING (or NOOP).
  // z_lo, z_hi: the volume's z-range, origin_z / spacing and
  // z_lo + 2 * padding + extent_z / spacing (used by erosion).
  Csr apply(int op, const Csr& in, double radius, double z_lo, double z_hi, double& time_1, double& time_2) {
    if (in.offsets.size() != size_t(int64_t(in.nx) * in.ny + 1))
      throw std::runtime_error("vo_hip: offsets must hold nx*ny + 1 entries");
    vo_volume v{in.nx, in.ny, in.offsets.data(), in.z.empty() ? nullptr : in.z.data(), z_lo, z_hi};
    vo_result* r = nullptr;
    int rc = vo_morph(ctx_, op, precision_, radius, &v, &r, &time_1, &time_2);
    if (rc != VO_OK) throw std::runtime_error(std::string("vo_morph: ") + vo_last_error(ctx_));
    Csr out;
    int64_t count = 0;
    vo_result_shape(r, &out.nx, &out.ny, &count);
    out.offsets.resize(size_t(int64_t(out.nx) * out.ny + 1));
    out.z.resize(size_t(count));
    rc = vo_result_copy(r, out.offsets.data(), out.z.empty() ? nullptr : out.z.data());
    vo_result_free(r);
    if (rc != VO_OK) throw std::runtime_error("vo_result_copy failed");
    return out;
  }
  Csr dilation(const Csr& in, double radius, double z_lo, double z_hi, double& t1, double& t2) {
    return apply(VO_OP_DILATION, in, radius, z_lo, z_hi, t1, t2);
  }
  Csr erosion(const Csr& in, double radius, double z_lo, double z_hi, double& t1, double& t2) {
    return apply(VO_OP_EROSION, in, radius, z_lo, z_hi, t1, t2);
  }

  vo_context* context() { return ctx_; }

 private:
  vo_context* ctx_ = nullptr;
  int precision_;
};

}  // namespace vo_hip
