// vo_parity.hpp -- parity metrics on the device (gfx950).
//
//   * measure_kernel: per column the summed length of its segments and their
//     number (pairs; an odd tail is ignored), for CompressedVolume::get_volume
//     and numSegments (CompressedVolume.cpp:61-83).
//   * xor_kernel: VoronoiMorpho::calculateXor per column
//     (Voronoi.cpp:91-111; calculate_ray_xor and removepoint,
//     MorphologyOperators.cpp:319-374): A \ B and B \ A obtained as
//     complements of unions of complements, united, and segments shorter than
//     1e-10 dropped.  The composition is evaluated literally, so the result is
//     the reference's endpoint for endpoint: complements are index views over
//     a sorted endpoint array, the two inner unions go to per-column scratch,
//     the outer union streams into the output.
// One thread per column; phase 0 counts, phase 1 writes at scanned offsets.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vo {

constexpr double kMinSegSize = 1e-10;  // MorphologyOperators.cpp:4-5

// A sorted endpoint array as is.
template <class T>
struct PlainView {
  const T* p;
  int64_t size;
  __device__ double at(int64_t i) const { return double(p[i]); }
};

// negate_ray (MorphologyOperators.cpp:230-259) of a sorted endpoint array
// within [lo, hi]: lo is prepended unless the first endpoint equals it, hi
// appended unless the last one equals it (each equal endpoint is dropped).
template <class T>
struct NegView {
  const T* p;
  int64_t n, size;
  double lo, hi;
  int pre, s0, t0;
  __device__ NegView(const T* p_, int64_t n_, double lo_, double hi_) : p(p_), n(n_), lo(lo_), hi(hi_) {
    if (n == 0) {
      pre = 1;
      s0 = t0 = 0;
      size = 2;
      return;
    }
    const bool s = double(p[0]) == lo, t = double(p[n - 1]) == hi;
    s0 = s ? 1 : 0;
    t0 = t ? 1 : 0;
    pre = s ? 0 : 1;
    size = pre + (n - s0 - t0) + (t ? 0 : 1);
  }
  __device__ double at(int64_t i) const {
    if (n == 0) return i == 0 ? lo : hi;
    if (i < pre) return lo;
    const int64_t j = i - pre + s0;
    return j < n - t0 ? double(p[j]) : hi;
  }
};

// Sink of a union into an array (push_merge: touching or overlapping
// intervals fuse; `pair` appends without merging).
struct ArraySink {
  double* buf;
  int64_t n = 0;
  __device__ void merge(double a, double b) {
    if (n == 0 || buf[n - 1] < a) {
      buf[n] = a;
      buf[n + 1] = b;
      n += 2;
    } else if (b > buf[n - 1]) {
      buf[n - 1] = b;
    }
  }
  __device__ void pair(double a, double b) {
    buf[n] = a;
    buf[n + 1] = b;
    n += 2;
  }
};

// Sink of the outer union: fuses like ArraySink, then drops finished
// segments shorter than kMinSegSize (removepoint) while counting or writing.
template <class T>
struct FinalSink {
  T* dst;  // null: count only
  int64_t pos = 0;
  double cs = 0, ce = 0;
  bool has = false;
  __device__ void flush() {
    if (has && !(ce - cs < kMinSegSize)) {
      if (dst) {
        dst[pos] = T(cs);
        dst[pos + 1] = T(ce);
      }
      pos += 2;
    }
    has = false;
  }
  __device__ void merge(double a, double b) {
    if (!has || ce < a) {
      flush();
      cs = a;
      ce = b;
      has = true;
    } else if (b > ce) {
      ce = b;
    }
  }
  __device__ void pair(double a, double b) {
    flush();
    cs = a;
    ce = b;
    has = true;
  }
};

// unionSegs (MorphologyOperators.cpp:117-158): merge of two sorted interval
// lists, the second first on equal starts; an empty side returns the other
// side unchanged.
template <class VA, class VB, class Sink>
__device__ void union_into(const VA& a, const VB& b, Sink& out) {
  if (a.size == 0) {
    for (int64_t i = 0; i + 1 < b.size; i += 2) out.pair(b.at(i), b.at(i + 1));
    return;
  }
  if (b.size == 0) {
    for (int64_t i = 0; i + 1 < a.size; i += 2) out.pair(a.at(i), a.at(i + 1));
    return;
  }
  int64_t ia = 0, ib = 0;
  while (ia + 1 < a.size || ib + 1 < b.size) {
    const bool take_b = ib + 1 < b.size && (ia + 1 >= a.size || b.at(ib) <= a.at(ia));
    if (take_b) {
      out.merge(b.at(ib), b.at(ib + 1));
      ib += 2;
    } else {
      out.merge(a.at(ia), a.at(ia + 1));
      ia += 2;
    }
  }
}

template <class T>
__global__ void measure_kernel(int64_t ncols, const int64_t* off, const T* z, double* length, int64_t* segments) {
  const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= ncols) return;
  const int64_t s = off[c], e = off[c + 1];
  double l = 0;
  for (int64_t i = s; i + 1 < e; i += 2) l += double(z[i + 1]) - double(z[i]);
  length[c] = l;
  segments[c] = (e - s) / 2;
}

// Per column: scratch of 2 * (n_a + n_b + 2) doubles at
// 2 * ((off_a[c] - off_a[0]) + (off_b[c] - off_b[0]) + 2c); n1/n2 keep the
// sizes of the two inner unions between the phases.
template <class T>
__global__ void xor_kernel(int64_t ncols, const int64_t* off_a, const T* za, const int64_t* off_b, const T* zb,
                           double lo, double hi, int phase, double* scratch, int64_t* n1, int64_t* n2,
                           int64_t* len, const int64_t* out_off, T* out) {
  const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= ncols) return;
  const int64_t sa = off_a[c], na = off_a[c + 1] - sa;
  const int64_t sb = off_b[c], nb = off_b[c + 1] - sb;
  double* t1 = scratch + 2 * ((sa - off_a[0]) + (sb - off_b[0]) + 2 * c);
  double* t2 = t1 + (na + nb + 2);
  int64_t m1, m2;
  if (phase == 0) {
    ArraySink s1{t1}, s2{t2};
    union_into(NegView<T>(za + sa, na, lo, hi), PlainView<T>{zb + sb, nb}, s1);
    union_into(NegView<T>(zb + sb, nb, lo, hi), PlainView<T>{za + sa, na}, s2);
    m1 = n1[c] = s1.n;
    m2 = n2[c] = s2.n;
  } else {
    m1 = n1[c];
    m2 = n2[c];
  }
  FinalSink<T> f{phase == 0 ? nullptr : out + out_off[c]};
  union_into(NegView<double>(t1, m1, lo, hi), NegView<double>(t2, m2, lo, hi), f);
  f.flush();
  if (phase == 0) len[c] = f.pos;
}

// dexel_dump (Dexelize.cpp:312-351): one hexahedron per segment of column
// (x, y): min = (origin + (x, y) * s, z_i * s), max = (origin + (x + 1,
// y + 1) * s, z_{i+1} * s), corners min + b * (max - min) in the reference's
// corner order (0,0,1)
// (1,0,1) (0,1,1) (1,1,1) (0,0,0) (1,0,0) (0,1,0) (1,1,0), and its six faces
// as outward quads.  seg_off[c] = segments before column c.
__constant__ int kHexQuads[6][4] = {{4, 6, 7, 5}, {0, 1, 3, 2}, {4, 5, 1, 0}, {6, 2, 3, 7}, {4, 0, 2, 6}, {5, 7, 3, 1}};

template <class T>
__global__ void dexel_mesh_kernel(int nx, int ny, const int64_t* off, const T* z, double ox, double oy,
                                  double z_shift, double s, const int64_t* seg_off, double* vertices,
                                  int64_t* quads) {
#pragma clang fp contract(off)
  const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= int64_t(nx) * ny) return;
  const int x = int(c % nx), y = int(c / nx);
  const double x0 = ox + x * s, x1 = ox + (x + 1) * s;
  const double y0 = oy + y * s, y1 = oy + (y + 1) * s;
  int64_t seg = seg_off[c];
  for (int64_t i = off[c]; i + 1 < off[c + 1]; i += 2, ++seg) {
    const double z0 = (double(z[i]) + z_shift) * s, z1 = (double(z[i + 1]) + z_shift) * s;
    // corner = min + b * (max - min), b in {0, 1}, as the reference forms it
    const double xe = x0 + (x1 - x0), ye = y0 + (y1 - y0), ze = z0 + (z1 - z0);
    double* v = vertices + seg * 24;
    for (int k = 0; k < 8; ++k) {
      v[3 * k] = (k & 1) ? xe : x0;
      v[3 * k + 1] = (k & 2) ? ye : y0;
      v[3 * k + 2] = (k < 4) ? ze : z0;
    }
    int64_t* q = quads + seg * 24;
    for (int f = 0; f < 6; ++f)
      for (int k = 0; k < 4; ++k) q[4 * f + k] = seg * 8 + kHexQuads[f][k];
  }
}

__global__ void pair_count_kernel(int64_t ncols, const int64_t* off, int64_t* pairs) {
  const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c < ncols) pairs[c] = (off[c + 1] - off[c]) / 2;
}

}  // namespace vo
