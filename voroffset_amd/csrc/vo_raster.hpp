// vo_raster.hpp -- mesh -> dexels on the GPU (gfx950).
//
// Replaces the ray casting of vor3d::create_dexels (Dexelize.cpp:166-225):
// for every dexel centre (origin_x + (x + 0.5) * spacing, origin_y + (y + 0.5)
// * spacing) (CompressedVolumeBase.cpp:5-11), the z / spacing of every
// triangle whose xy-projection contains it, sorted ascending; pairs of
// consecutive values are the dexel's segments (even-odd fill).  Containment is
// the SOS test of point_in_triangle_2d (Dexelize.cpp:56-92), and a triangle
// whose projection has zero inexact orientation contributes nothing
// (Dexelize.cpp:107-121,151-158).  The arithmetic follows the reference's
// operation order with FMA contraction off, so every decision and every z
// equals the CPU's bit for bit.
//
// Steps: (1) count, per 8x8-dexel tile, the triangles whose bounding box may
// contain one of its dexel centres; (2) scan; (3) fill the tile lists;
// (4) one wavefront per tile, one lane per dexel: the tile's triangles are
// staged 64 at a time in LDS and every lane tests its centre against them,
// counting hits; (5) scan the counts into the output offsets; (6) the same
// sweep writes the hits, and each lane sorts its column.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vo_wave.hpp"

namespace vo {

constexpr int kRasterTile = 8;  // dexels per tile side (64 = one wavefront)

struct RasterGrid {
  int nx, ny;
  double ox, oy, spacing;
  int tx, ty;  // tiles per row / column
};

// Tile range of triangle f: dexel centres that may lie in its closed xy
// bounding box, widened by one dexel (pruning only; the SOS test decides).
__device__ __forceinline__ bool tri_tiles(const double* V, const int32_t* F, int64_t f, const RasterGrid& g,
                                          int& ax, int& bx, int& ay, int& by) {
#pragma clang fp contract(off)
  double lo[2] = {INFINITY, INFINITY}, hi[2] = {-INFINITY, -INFINITY};
  for (int k = 0; k < 3; ++k) {
    const int64_t v = F[3 * f + k];
    for (int d = 0; d < 2; ++d) {
      lo[d] = fmin(lo[d], V[3 * v + d]);
      hi[d] = fmax(hi[d], V[3 * v + d]);
    }
  }
  const double o[2] = {g.ox, g.oy};
  const int n[2] = {g.nx, g.ny};
  int il[2], ih[2];
  for (int d = 0; d < 2; ++d) {
    const double a = floor((lo[d] - o[d]) / g.spacing - 0.5) - 1.0;
    const double b = ceil((hi[d] - o[d]) / g.spacing - 0.5) + 1.0;
    il[d] = int(fmin(fmax(a, 0.0), double(n[d])));         // clamped before the conversion
    ih[d] = int(fmax(fmin(b, double(n[d] - 1)), -1.0));
  }
  if (il[0] > ih[0] || il[1] > ih[1]) return false;
  ax = il[0] / kRasterTile;
  bx = ih[0] / kRasterTile;
  ay = il[1] / kRasterTile;
  by = ih[1] / kRasterTile;
  return true;
}

// phase 0: count per tile; phase 1: fill (cursor starts at the tile offsets)
__global__ void raster_bin_kernel(const double* V, const int32_t* F, int64_t nf, RasterGrid g, int phase,
                                  int64_t* tile_cnt, int64_t* tile_cur, int64_t* tile_tris) {
  const int64_t f = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  int ax, bx, ay, by;
  if (!tri_tiles(V, F, f, g, ax, bx, ay, by)) return;
  for (int b = ay; b <= by; ++b)
    for (int a = ax; a <= bx; ++a) {
      const int64_t t = int64_t(b) * g.tx + a;
      if (phase == 0)
        atomicAdd(reinterpret_cast<unsigned long long*>(tile_cnt + t), 1ull);
      else
        tile_tris[atomicAdd(reinterpret_cast<unsigned long long*>(tile_cur + t), 1ull)] = f;
    }
}

__device__ __forceinline__ int sos_orientation(double x1, double y1, double x2, double y2, double& area) {
#pragma clang fp contract(off)
  area = y1 * x2 - x1 * y2;
  if (area > 0) return 1;
  if (area < 0) return -1;
  if (y2 > y1) return 1;
  if (y2 < y1) return -1;
  if (x1 > x2) return 1;
  if (x1 < x2) return -1;
  return 0;
}

// The hit of the vertical ray through (qx, qy) with triangle p (9 doubles:
// x1 y1 z1 x2 y2 z2 x3 y3 z3); returns false when it misses.
__device__ __forceinline__ bool ray_hit(const double* p, double qx, double qy, double& z) {
#pragma clang fp contract(off)
  const double x1 = p[0] - qx, y1 = p[1] - qy;
  const double x2 = p[3] - qx, y2 = p[4] - qy;
  const double x3 = p[6] - qx, y3 = p[7] - qy;
  double a, b, c;
  const int sa = sos_orientation(x2, y2, x3, y3, a);
  if (sa == 0) return false;
  if (sos_orientation(x3, y3, x1, y1, b) != sa) return false;
  if (sos_orientation(x1, y1, x2, y2, c) != sa) return false;
  const double sum = a + b + c;
  a /= sum;
  b /= sum;
  c /= sum;
  z = a * p[2] + b * p[5] + c * p[8];
  const double a11 = p[3] - p[0], a12 = p[4] - p[1];
  const double a21 = p[6] - p[0], a22 = p[7] - p[1];
  const double delta = a11 * a22 - a12 * a21;
  return delta != 0;
}

// phase 0: per-dexel hit counts into len[col]; phase 1: hits into
// out[off[col] ..], each column then sorted.  One wavefront per tile.
__global__ __launch_bounds__(kWave) void raster_cast_kernel(const double* V, const int32_t* F, RasterGrid g,
                                                            const int64_t* tile_off, const int64_t* tile_tris,
                                                            int phase, int64_t* len, const int64_t* off,
                                                            double* out) {
#pragma clang fp contract(off)
  __shared__ double tri[kWave * 9];
  const int L = lane_id();
  const int64_t t = blockIdx.x;
  const int x = int(t % g.tx) * kRasterTile + (L % kRasterTile);
  const int y = int(t / g.tx) * kRasterTile + (L / kRasterTile);
  const bool active = x < g.nx && y < g.ny;
  const double qx = (x + 0.5) * g.spacing + g.ox, qy = (y + 0.5) * g.spacing + g.oy;
  const int64_t col = int64_t(x) + int64_t(g.nx) * y;
  const int64_t t0 = tile_off[t], t1 = tile_off[t + 1];
  int64_t n = 0;
  const int64_t base = (phase == 1 && active) ? off[col] : 0;
  for (int64_t c0 = t0; c0 < t1; c0 += kWave) {
    const int cnt = int(min<int64_t>(kWave, t1 - c0));
    if (L < cnt) {
      const int64_t f = tile_tris[c0 + L];
      for (int k = 0; k < 3; ++k) {
        const int64_t v = F[3 * f + k];
        tri[L * 9 + 3 * k] = V[3 * v];
        tri[L * 9 + 3 * k + 1] = V[3 * v + 1];
        tri[L * 9 + 3 * k + 2] = V[3 * v + 2];
      }
    }
    wave_sync();
    if (active) {
      for (int i = 0; i < cnt; ++i) {
        double z;
        if (ray_hit(tri + 9 * i, qx, qy, z)) {
          if (phase == 1) out[base + n] = z / g.spacing;
          ++n;
        }
      }
    }
    wave_sync();
  }
  if (!active) return;
  if (phase == 0) {
    len[col] = n;
    return;
  }
  // insertion sort of the column (columns hold few hits)
  double* c = out + base;
  for (int64_t i = 1; i < n; ++i) {
    const double v = c[i];
    int64_t j = i - 1;
    while (j >= 0 && c[j] > v) {
      c[j + 1] = c[j];
      --j;
    }
    c[j + 1] = v;
  }
}

}  // namespace vo
