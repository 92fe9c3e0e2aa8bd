// vo_wave.hpp -- wavefront-level building blocks of the dexel sweeps (gfx950).
//
// Every sweep kernel runs one 64-lane wavefront per workgroup, so
// __syncthreads() is a wavefront barrier and all lists live in that
// workgroup's LDS.  Lists are z-sorted arrays; lane l handles elements
// l, l+64, l+128, ...  Compaction uses ballot + mbcnt, ordering uses ranks
// found by binary search in LDS, and all geometric predicates are evaluated in
// FP64 from the stored endpoint values (FP32 or FP64) and integer line
// offsets, so the decisions do not depend on the storage precision.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vo {

constexpr int kWave = 64;
constexpr int kGapCode = INT32_MIN;  // "no owner" code of an elementary interval
constexpr int kGatherTile = kWave;   // output columns per tile of the gather form of pass 2

__device__ __forceinline__ int lane_id() { return threadIdx.x; }
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
__device__ __forceinline__ int count_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
}
__device__ __forceinline__ int popc(uint64_t m) { return __popcll(m); }
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

// Workgroup b of a launch runs on XCD b % 8 (the dispatcher deals workgroups
// to the eight XCDs round-robin).  xcd_tile maps b to a logical tile so that
// each XCD receives one contiguous range of tiles: neighbouring tiles, which
// read and write the same cache lines, then meet in the same L2.  A
// bijection on [0, n).
constexpr int kXcds = 8;
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t n) {
  const int64_t q = n / kXcds, r = n % kXcds;
  const int64_t xcd = b % kXcds, k = b / kXcds;
  return xcd * q + (xcd < r ? xcd : r) + k;
}
// Batched launches, whose tiles differ in cost from volume to volume: the
// tiles go to the XCDs in chunks of kXcdChunk consecutive tiles, dealt
// round-robin, so that every XCD gets a share of every stretch of the batch
// (contiguous eighths left the XCD holding the costliest volumes working
// alone at the end: 39.5 ms against 35.8 ms for 256 uneven volumes) while
// neighbouring tiles still meet in one L2.  A bijection on [0, n): the tiles
// past the last whole round of chunks keep b.
constexpr int64_t kXcdChunk = 64;
__device__ __forceinline__ int64_t xcd_chunk_tile(int64_t b, int64_t n) {
  const int64_t full = n / (kXcds * kXcdChunk) * (kXcds * kXcdChunk);
  if (b >= full) return b;
  const int64_t xcd = b % kXcds, k = b / kXcds;
  return ((k / kXcdChunk) * kXcds + xcd) * kXcdChunk + k % kXcdChunk;
}
// Persistent form of a batched launch (one wavefront per workgroup): the
// grid holds about as many workgroups as the CUs keep resident, and each
// takes tiles until none is left.  Queue q (q = 0 .. 7, counters cleared
// before the launch) hands out the tiles xcd_chunk_tile deals to the
// workgroups b = q, q + 8, ... of a one-tile-per-workgroup grid, i.e. the
// tiles XCD q would run; a workgroup takes from its own XCD's queue first
// and then from the others in turn, so an XCD that runs out of work early
// takes over tiles of the others.  Every tile is taken exactly once, and
// every wavefront leaves after one failed take per queue.
template <class F>
__device__ void take_tiles(unsigned* queue, int64_t n, F&& body) {
  const int home = int(blockIdx.x % kXcds);
  for (int s = 0; s < kXcds; ++s) {
    const int q = (home + s) % kXcds;
    const int64_t nq = q < n ? (n - 1 - q) / kXcds + 1 : 0;  // tiles b = q, q + 8, ... below n
    for (;;) {
      unsigned k = 0;
      if (lane_id() == 0) k = atomicAdd(&queue[q], 1u);
      // a scalar: the tile, its job and the job's arguments then stay in
      // scalar registers, as they do when the tile comes from blockIdx
      k = __builtin_amdgcn_readfirstlane(k);
      if (int64_t(k) >= nq) break;
      body(xcd_chunk_tile(int64_t(k) * kXcds + q, n));
      __syncthreads();  // the next tile's body reuses the workgroup's LDS
    }
  }
}
__device__ __forceinline__ void wave_sync() { __syncthreads(); }

// Status bits written by the kernels (host turns them into error codes).
enum : int {
  kStatusArenaOverflow = 1,   // an LDS arena of a sweep form exceeded its capacity (shorter chunks fit)
  kStatusP1List = 2,          // pass 1: live piece list
  kStatusP1Union = 4,         // pass 1: forward/backward union
  kStatusP2Interiors = 8,     // pass 2: interior list
  kStatusP2Disks = 16,        // pass 2: disk list
  kStatusP2Slice = 32,        // pass 2: intervals of one line slice
  kStatusGatherOverflow = 64, // pass 2 gather form, batched: a tile's lists outgrew LDS
  kStatusSlotOverflow = 128,  // pass 1: a column's triples exceeded its slots (caps)
  kStatusP2Arena = 256,       // pass 2 sweep form: a column's output exceeded its arena
  kStatusListOverflow = kStatusP1List | kStatusP1Union | kStatusP2Interiors | kStatusP2Disks | kStatusP2Slice,
};

// Input-check bits (a word of their own, written by check_input_kernel).
enum : int {
  kInputMalformed = 1,   // decreasing offsets, odd column, unsorted or non-finite endpoint
  kInputOutOfRange = 2,  // an endpoint outside the range the operation needs
};

__device__ __forceinline__ void raise_status(int* status, int bits) {
  if (lane_id() == 0) atomicOr(status, bits);
}

// Number of entries of the non-decreasing sequence get(0..len-1) that are
// < v (inclusive=false) or <= v (inclusive=true).
template <class Get>
__device__ __forceinline__ int rank_in(const Get& get, int len, double v, bool inclusive) {
  int lo = 0, hi = len;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const double x = get(mid);
    const bool before = inclusive ? (x <= v) : (x < v);
    if (before)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// A z-sorted list of disjoint (possibly touching) pieces [lo, hi] with an
// integer owner key, held in LDS.
template <class T>
struct PieceRef {
  T* lo;
  T* hi;
  int* key;
};

// Scratch for merge_max: merged endpoint sequence and compacted intervals.
template <class T>
struct MergeScratch {
  T* ev;        // merged endpoint values      [cap]
  int* ecount;  // (#A endpoints <= here) | (#B endpoints <= here) << 16
  T* tlo;       // non-transparent elementary intervals [cap]
  T* thi;
  int* tcode;
  int cap;
};

// Painter merge of two z-sorted disjoint piece lists: every z is owned by the
// covering piece with the larger key (A wins ties).  Zero-length pieces
// (points) survive when they win.  Adjacent output with equal keys is fused.
// Writes at most `cap_o` pieces to `o` and returns the count, or -1 on
// overflow of either the scratch or the output.
template <class T>
__device__ int merge_max(PieceRef<T> a, int na, PieceRef<T> b, int nb, PieceRef<T> o, int cap_o,
                         MergeScratch<T> s) {
  const int L = lane_id();
  const int ne = 2 * (na + nb);
  if (ne > s.cap || ne >= (1 << 15)) return -1;
  if (ne == 0) return 0;
  auto ea = [&](int u) -> double { return (u & 1) ? double(a.hi[u >> 1]) : double(a.lo[u >> 1]); };
  auto eb = [&](int w) -> double { return (w & 1) ? double(b.hi[w >> 1]) : double(b.lo[w >> 1]); };
  // 1. merged endpoint order (A before B on equal values)
  for (int u = L; u < 2 * na; u += kWave) {
    const double v = ea(u);
    const int r = rank_in(eb, 2 * nb, v, false);
    s.ev[u + r] = (u & 1) ? a.hi[u >> 1] : a.lo[u >> 1];
    s.ecount[u + r] = (u + 1) | (r << 16);
  }
  for (int w = L; w < 2 * nb; w += kWave) {
    const double v = eb(w);
    const int r = rank_in(ea, 2 * na, v, true);
    s.ev[w + r] = (w & 1) ? b.hi[w >> 1] : b.lo[w >> 1];
    s.ecount[w + r] = r | ((w + 1) << 16);
  }
  wave_sync();
  // 2. owners of the elementary intervals; drop zero-length ones unless a
  //    winning point owns them
  int nt = 0;
  for (int base = 0; base < ne - 1; base += kWave) {
    const int p = base + L;
    bool keep = false;
    int code = kGapCode;
    T lo = T(0), hi = T(0);
    if (p < ne - 1) {
      lo = s.ev[p];
      hi = s.ev[p + 1];
      const int c = s.ecount[p];
      const int ca = c & 0xffff, cb = c >> 16;
      const bool inA = ca & 1, inB = cb & 1;
      const int ia = (ca - 1) >> 1, ib = (cb - 1) >> 1;
      bool point = false;
      if (inB && (!inA || b.key[ib] > a.key[ia])) {
        code = b.key[ib];
        point = !(b.lo[ib] < b.hi[ib]);
      } else if (inA) {
        code = a.key[ia];
        point = !(a.lo[ia] < a.hi[ia]);
      }
      const bool zero = !(lo < hi);
      keep = !zero || (code != kGapCode && point);
    }
    const uint64_t m = ballot(keep);
    const int idx = nt + count_below(m);
    if (keep && idx < s.cap) {
      s.tlo[idx] = lo;
      s.thi[idx] = hi;
      s.tcode[idx] = code;
    }
    nt += popc(m);
  }
  if (nt > s.cap) return -1;
  wave_sync();
  // 3. runs of equal owner -> output pieces
  int ns = 0, nf = 0;
  for (int base = 0; base < nt; base += kWave) {
    const int t = base + L;
    bool start = false, fin = false;
    int code = kGapCode;
    if (t < nt) {
      code = s.tcode[t];
      const int prev = t > 0 ? s.tcode[t - 1] : kGapCode;
      const int next = t + 1 < nt ? s.tcode[t + 1] : kGapCode;
      start = code != kGapCode && prev != code;
      fin = code != kGapCode && next != code;
    }
    const uint64_t ms = ballot(start), mf = ballot(fin);
    const int is = ns + count_below(ms), jf = nf + count_below(mf);
    if (start && is < cap_o) {
      o.lo[is] = s.tlo[t];
      o.key[is] = code;
    }
    if (fin && jf < cap_o) o.hi[jf] = s.thi[t];
    ns += popc(ms);
    nf += popc(mf);
  }
  wave_sync();
  return ns > cap_o ? -1 : ns;
}

// Keeps the elements for which keep(i) holds, compacting src -> dst (which
// must not alias).  Returns the new count.
template <class Keep, class Move>
__device__ __forceinline__ int compact(int n, const Keep& keep, const Move& move) {
  const int L = lane_id();
  int out = 0;
  for (int base = 0; base < n; base += kWave) {
    const int i = base + L;
    const bool k = i < n && keep(i);
    const uint64_t m = ballot(k);
    if (k) move(i, out + count_below(m));
    out += popc(m);
  }
  wave_sync();
  return out;
}

// Crossing of the distance functions of two z-disjoint vertical sites on the
// sweep line.  Site l = [al, bl] lies below site u = [au, bu] (bl <= au);
// wl, wu are their squared offsets from the line.  f(z) = w + dist(z, site)^2.
// f_l - f_u is non-decreasing in z; returns the z where it changes sign
// (+-inf when one site wins everywhere).
__device__ __forceinline__ double site_crossing(double al, double bl, double wl, double au, double bu,
                                                double wu) {
  const double g = au - bl;  // gap between the sites, >= 0
  const double dw = wl - wu;
  const double f_bl = dw - g * g;  // f_l - f_u at z = bl
  const double f_au = dw + g * g;  // f_l - f_u at z = au
  if (f_bl >= 0) {
    const double d = au - al;
    if (dw - d * d <= 0) return au - sqrt(dw);  // inside [al, bl]
    if (d > 0) return 0.5 * (al + au) - dw / (2.0 * d);  // below al
    return -INFINITY;
  }
  if (f_au <= 0) {
    const double d = bu - bl;
    if (dw + d * d >= 0) return bl + sqrt(-dw);  // inside [au, bu]
    if (d > 0) return 0.5 * (bl + bu) - dw / (2.0 * d);  // above bu
    return INFINITY;
  }
  return 0.5 * (bl + au) - dw / (2.0 * g);  // inside the gap
}

// Exact integer reach of a site whose offset across the sweep is dy:
// the largest t >= 0 with t^2 + dy^2 <= R2, or -1 when dy^2 > R2.
__device__ __forceinline__ int reach_of(int dy, double R2) {
  const double v = R2 - double(dy) * double(dy);
  if (v < 0) return -1;
  int t = int(sqrt(v));
  while (double(t + 1) * double(t + 1) <= v) ++t;
  while (t > 0 && double(t) * double(t) > v) --t;
  return t;
}

// Inclusive wavefront scans.
__device__ __forceinline__ double wave_scan_max(double v) {
  const int L = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const double o = __shfl_up(v, d);
    if (L >= d) v = fmax(v, o);
  }
  return v;
}

__device__ __forceinline__ double wave_rscan_min(double v) {  // suffix min
  const int L = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const double o = __shfl_down(v, d);
    if (L + d < kWave) v = fmin(v, o);
  }
  return v;
}

}  // namespace vo
