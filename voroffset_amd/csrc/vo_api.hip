// vo_api.hip -- host side of the C ABI declared in include/voroffset_hip.h.
//
// Orchestration of one dilation (VoronoiMorphoVorPower::dilation,
// src/vor3d/VoronoiVorPower.cpp:24-96) on the device:
//   pass 1 kernel  -> per-column triples (lo, hi, dy) in a chunked arena
//   pass 2 kernel  -> per-column output intervals in a chunked arena
//   exclusive scan of the output lengths + gather -> CSR result
// Erosion (Voronoi.cpp:8-17) adds a ring-complement kernel before and a
// crop-complement kernel after; closing / opening chain the two on the
// device (offset3d.cpp:124-133).  A workspace pool in the context keeps every
// buffer alive between calls, so repeated calls of the same size allocate
// nothing.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <thread>
#include <vector>

#include "../../include/voroffset_hip.h"
#include "vo_parity.hpp"
#include "vo_raster.hpp"
#include "vo_sweeps.hpp"

namespace {

// Kernel shapes.  The LDS footprint of each instantiation is checked below.
constexpr int kP1Cap = 128, kP1NCap = 64;
constexpr int kP2SCap = 128, kP2PCap = 256, kP2NCap = 64, kP2OCap = 256;
// Most workspace the gather's interval lists and output arena may take when
// their capacity grows for tiles with many output intervals per column.
constexpr size_t kGatherRoomBytes = size_t(4) << 30;
// gather form: intervals per output column list.  Large radii fuse the
// dilated pieces into few intervals, so they get the short list (and more
// waves per CU); a tile that overflows is recomputed by the sweep form.
// Tuning table: per radius bucket, the z-sweep event capacity per slab and
// the gather list capacity (defaults below; vo_autotune measures them).
constexpr int kTuneBuckets = 5;
inline int tune_bucket(int rint) { return rint < 8 ? 0 : rint < 16 ? 1 : rint < 32 ? 2 : rint < 64 ? 3 : 4; }
constexpr int kEventChoices[] = {512, 768, 1024, 1536};
struct Tuning {
  int events[kTuneBuckets] = {768, 768, 768, 768, 768};
  int ocap[kTuneBuckets] = {24, 16, 16, 16, 16};
};
// pass-1 z-sweep form: LDS events per z-slab (denser windows take more slabs)
// and histogram bins; small enough for ~5 waves per SIMD
constexpr int kZECap = 768, kZNB = 512, kZWMax = 64 + 2 * 128, kZWMax8 = 64 + 2 * 256;

template <class T>
struct Shapes;
// (p1_cap_big: the painter's piece lists when the default ones overflow,
// as long as its LDS fits a CU's 160 KB)
template <>
struct Shapes<float> {
  static constexpr int p1_cap_big = 1024;
  static constexpr int p1_arena = 2048, p1_cmax = 64;
  static constexpr int p2_farena = 512, p2_cmax = 64;
};
template <>
struct Shapes<double> {
  static constexpr int p1_cap_big = 640;
  static constexpr int p1_arena = 1024, p1_cmax = 32;
  static constexpr int p2_farena = 256, p2_cmax = 32;
};

template <class T>
using P1S = vo::Pass1Smem<T, kP1Cap, kP1NCap, Shapes<T>::p1_arena, Shapes<T>::p1_cmax>;
template <class T>
using P2S = vo::Pass2Smem<T, kP2SCap, kP2PCap, kP2NCap, kP2OCap, Shapes<T>::p2_farena, Shapes<T>::p2_cmax>;
static_assert(sizeof(P1S<float>) <= 65536, "pass-1 LDS (fp32)");
static_assert(sizeof(P1S<double>) <= 65536, "pass-1 LDS (fp64)");
template <class T>
using P1P = vo::Pass1PSmem<T, kP1Cap, Shapes<T>::p1_arena, Shapes<T>::p1_cmax>;
static_assert(sizeof(P1P<float>) <= 65536, "pass-1 painter LDS (fp32)");
static_assert(sizeof(P1P<double>) <= 65536, "pass-1 painter LDS (fp64)");
template <class T>
using P1PBig = vo::Pass1PSmem<T, Shapes<T>::p1_cap_big, Shapes<T>::p1_arena, Shapes<T>::p1_cmax>;
static_assert(sizeof(P1PBig<float>) <= 160 * 1024, "pass-1 painter LDS, large lists (fp32)");
static_assert(sizeof(P1PBig<double>) <= 160 * 1024, "pass-1 painter LDS, large lists (fp64)");
static_assert(sizeof(P2S<float>) <= 65536, "pass-2 LDS (fp32)");
static_assert(sizeof(P2S<double>) <= 65536, "pass-2 LDS (fp64)");
static_assert(sizeof(vo::Pass1ZSmem<double, kZECap, kZNB, kZWMax>) <= 65536, "pass-1 z-sweep LDS");

struct Buffer {
  void* ptr = nullptr;
  size_t bytes = 0;
};

enum Slot {
  kM_LO, kM_HI, kM_DY, kM_CNT,  // pass-1 output
  kO_Z, kO_START, kO_LEN,                  // pass-2 output
  kSTATUS, kSCAN_TMP, kTILES,
  kV0_OFF, kV0_Z, kV1_OFF, kV1_Z, kV2_OFF, kV2_Z, kV3_OFF, kV3_Z,  // device volumes
  kIN_OFF, kIN_Z,                                                   // staged host input
  kR_V, kR_F, kR_TCNT, kR_TOFF, kR_TCUR, kR_TTRIS, kR_LEN, kR_OFF, kR_Z,  // rasterizer
  kB_TAB0, kB_TAB1, kB_TAB2, kB_STATUS, kB_FIN_OFF, kB_FIN_AUX,         // batched form
  kB_QUEUE,                                                 // batched form: work-queue counters
  kG_LISTS, kG_HTAB,                                         // gather interval lists, H table
  kM_PAIRS, kM_SEGOFF, kM_VERT, kM_QUAD,                     // dexel mesh writer
  kM_SLO, kM_SHI, kM_CHI,                                    // pass-1 block and column summaries
  kX_SCR, kX_N1, kX_N2, kX_LEN, kX_OFF, kX_Z, kX_LENGTH, kX_SEGS, kX_SUM,  // parity metrics
  kV_WORD, kV_TAB, kV_BASE,                // device-side input check
  kB_OUT,                                  // + (stage * 3 + kind) * 2 (+1 for z): stage outputs
  kSLOTS = kB_OUT + 4 * 3 * 2
};

struct DevVol {
  int nx = 0, ny = 0;
  const int64_t* off = nullptr;
  const void* z = nullptr;
  int64_t count = 0;
};

}  // namespace

// The no-alloc entry point's workspace: one array per GPU, held until the
// process ends.  Host code never uses its name as an address (that is a
// host-side stand-in): vo_create looks up each device's copy with
// hipGetSymbolAddress once, with that device current, and passes that
// address on.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// Bump allocation over a fixed device buffer: the workspace of the entry
// point that allocates nothing (vo_morph_device_noalloc).  Each slot keeps
// the region it was given for the call; `fixed` slots are the caller's
// output buffers, never moved.
struct ScratchArena {
  bool on = false;
  char* base = nullptr;
  size_t cap = 0, used = 0;
  size_t short_by = 0;  // bytes an output buffer needed when it was too small
  Buffer slot[kSLOTS];
  bool fixed[kSLOTS] = {};
};

// The last batch program run on a context: its key (tables, launch shapes,
// buffers) and, from its second identical call on, its captured graph.
struct BatchGraph {
  std::vector<char> key;
  hipGraphExec_t exec = nullptr;
};

struct vo_context {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  Buffer buf[kSLOTS];
  int64_t* host_words = nullptr;  // pinned: status + counts; [3] = input-check words
  std::string error;
  bool check_input = false;  // an input check was enqueued by this call: every synchronisation looks at it
  std::string inject;        // VO_INJECT_HIP_ERROR: a HIP call whose label contains it reports an error
  int64_t last_outputs = 0;
  int p1_caps = 96;         // pass-1 slots per column (grows on overflow; a multiple of vo::kSlotBlock)
  int p2_cap_per_col = 16;  // pass-2 arena endpoints per column (grows on overflow)
  bool force_sweep = false;  // VO_PASS2=sweep: always use the sweep form of pass 2
  int zsweep_events = 0;     // VO_ZSWEEP_EVENTS: force a z-sweep event capacity (512, 768, 1024, 1536)
  Tuning tuning;             // per radius bucket (vo_autotune, vo_tuning_load)
  int pass1_form = 0;        // VO_PASS1: 0 = z-sweep (default), 1 = "paint", 2 = "cull"
  int axis_forced = -1;      // VO_AXIS: -1 = chosen per volume, 0 = pass 1 along y, 1 = along x
  void* scratch = nullptr;   // this device's no-alloc scratch (device address, from hipGetSymbolAddress)
  ScratchArena arena;        // on during vo_morph_device_noalloc
  int64_t* batch_host = nullptr;  // pinned: a batch program's counts and statuses
  size_t batch_host_n = 0;
  void* staging = nullptr;        // pinned: host volumes on their way to and from the device
  size_t staging_bytes = 0;
  BatchGraph bgraph;              // the last batch program, captured on its second call
  bool use_graphs = true;         // VO_BATCH_GRAPH=0: never capture
  bool batch_queue = false;       // VO_BATCH_QUEUE=1: persistent work-queue grids for the batched passes
  int num_cus = 0;                // compute units of the device
  std::map<std::pair<const void*, size_t>, int> resident;  // workgroups per CU, by kernel and dynamic LDS
  int last_fallback_tiles = 0;
  // the radius whose gather table the H-table slot holds, and where (a call
  // with the same radius skips the upload)
  double htab_R = -1;
  const void* htab_at = nullptr;
  int64_t spec_out = 0;  // endpoints the next dilation's output is first sized for
  // an input check left to the pass 1 of the operation's first dilation
  // (word == null: none pending); see defer_input_check
  struct {
    int* word = nullptr;
    int range = 0;
    double lo = 0, hi = 0;
  } fused;
  std::vector<void*> batch_extra;  // outputs of batched jobs redone alone (freed by the next batch)
  // host copies of the input-check tables (every call ends synchronised, so
  // they outlive their uploads)
  std::vector<char> check_tab;
  std::vector<int64_t> check_base;
  std::vector<int> check_jobs;  // the caller's job index of each checked job
  int last_batch_alone = 0;        // jobs of the last batch redone by the single-volume path
};

struct vo_result {
  int32_t nx = 0, ny = 0;
  std::vector<int64_t> off;
  std::vector<double> z;
};

struct vo_triples {
  int32_t nx = 0, ny = 0;
  std::vector<int64_t> off;  // CSR over triples, column x + nx * y
  std::vector<double> y1, y2, r;
};

struct vo_mesh_result {
  std::vector<double> vertices;  // 3 per vertex
  std::vector<int64_t> quads;    // 4 per quad
};

namespace {

struct Failure {
  int code;
};

void fail(vo_context* ctx, int code, const std::string& msg) {
  ctx->error = msg;
  throw Failure{code};
}

void check(vo_context* ctx, hipError_t e, const char* what) {
  if (e != hipSuccess) fail(ctx, VO_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
  // fault hook for the error paths of callers (tests): the call itself ran
  if (!ctx->inject.empty() && std::strstr(what, ctx->inject.c_str()))
    fail(ctx, VO_ERR_DEVICE, std::string(what) + ": injected HIP error (VO_INJECT_HIP_ERROR)");
}

// The result of this call's input check (copied to host_words[3] right after
// the check kernel, so any later synchronisation has it).
void input_check(vo_context* ctx) {
  if (!ctx->check_input) return;
  int w[2];
  std::memcpy(w, ctx->host_words + 3, sizeof(w));
  if (w[0] == 0) return;
  ctx->check_input = false;
  const int j = w[1] >= 0 && w[1] < int(ctx->check_jobs.size()) ? ctx->check_jobs[w[1]] : w[1];
  const std::string job = " (job " + std::to_string(j) + ")";
  if (w[0] & vo::kInputMalformed)
    fail(ctx, VO_ERR_ARGUMENT, "input volume: decreasing offsets, odd column, unsorted or non-finite endpoint" + job);
  fail(ctx, VO_ERR_RANGE, "input volume: endpoint outside the z-range the operation needs "
                          "([z_lo - 1, z_hi + 1]; for a closing also after dilation by the radius)" + job);
}

void sync(vo_context* ctx) {
  check(ctx, hipStreamSynchronize(ctx->stream), "sync");
  input_check(ctx);
}

void* ws(vo_context* ctx, int slot, size_t bytes) {
  if (ctx->arena.on) {  // the no-alloc entry point: regions of the scratch
    ScratchArena& a = ctx->arena;
    Buffer& b = a.slot[slot];
    if (bytes == 0) bytes = 16;
    if (b.bytes >= bytes) return b.ptr;
    if (a.fixed[slot]) a.short_by = bytes;
    if (a.fixed[slot])
      fail(ctx, VO_ERR_CAPACITY, "no-alloc: the output buffer holds " + std::to_string(b.bytes) + " bytes, " +
                                     std::to_string(bytes) + " are needed");
    const size_t at = (a.used + 255) & ~size_t(255);
    if (at + bytes > a.cap)
      fail(ctx, VO_ERR_CAPACITY, "no-alloc: the " + std::to_string(a.cap >> 20) +
                                     " MiB scratch is too small for this volume (use vo_morph_device)");
    b.ptr = a.base + at;
    b.bytes = bytes;
    a.used = at + bytes;
    return b.ptr;
  }
  Buffer& b = ctx->buf[slot];
  if (bytes == 0) bytes = 16;
  if (b.bytes < bytes) {
    if (b.ptr) {
      check(ctx, hipStreamSynchronize(ctx->stream), "sync");  // the old buffer may still be in use
      check(ctx, hipFree(b.ptr), "hipFree");
    }
    b.ptr = nullptr;
    const size_t want = bytes + bytes / 4;
    if (hipMalloc(&b.ptr, want) != hipSuccess) {
      b.ptr = nullptr;
      b.bytes = 0;
      (void)hipGetLastError();
      fail(ctx, VO_ERR_DEVICE, "hipMalloc of " + std::to_string(want) + " bytes (workspace slot " +
                                   std::to_string(slot) + "): out of memory");
    }
    b.bytes = want;
    check(ctx, hipSuccess, "hipMalloc");  // (the fault hook's label)
  }
  return b.ptr;
}

// The buffer ws() last gave for `slot` in this mode.
void* slot_ptr(vo_context* ctx, int slot) { return ctx->arena.on ? ctx->arena.slot[slot].ptr : ctx->buf[slot].ptr; }

unsigned blocks_for(int64_t n, int threads) { return unsigned((n + threads - 1) / threads); }

// Exclusive scan of len[0..n] (len[n] must be 0) into off[0..n].
void exclusive_scan(vo_context* ctx, const int64_t* len, int64_t* off, int64_t n) {
  size_t tmp = 0;
  check(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, len, off, int(n), ctx->stream), "scan size");
  void* t = ws(ctx, kSCAN_TMP, tmp);
  check(ctx, hipcub::DeviceScan::ExclusiveSum(t, tmp, len, off, int(n), ctx->stream), "scan");
}

int64_t read_word(vo_context* ctx, const int64_t* dptr) {
  check(ctx, hipMemcpyAsync(ctx->host_words, dptr, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream),
        "copy word");
  sync(ctx);
  return ctx->host_words[0];
}

std::string status_text(int s) {
  std::string t;
  if (s & vo::kStatusP1List) t += " pass-1 piece list;";
  if (s & vo::kStatusP1Union) t += " pass-1 forward/backward union;";
  if (s & vo::kStatusP2Interiors) t += " pass-2 interior list;";
  if (s & vo::kStatusP2Disks) t += " pass-2 disk list;";
  if (s & vo::kStatusP2Slice) t += " pass-2 line slice;";
  if (s & vo::kStatusSlotOverflow) t += " pass-1 slots of a column;";
  if (s & vo::kStatusP2Arena) t += " pass-2 output arena of a column;";
  return "a sweep exceeded its LDS capacity:" + t;
}

// Appends the gather form's table for radius R: rint + 1 rows (|d|) of
// htab_width(rint) entries, H(|d|, dy) = sqrt(R^2 - d^2 - dy^2) for
// dy = 0 .. rint, NaN where that is negative, then the NaN entry that slots
// without a triple read.
void append_htab(double R, std::vector<double>& out) {
  const int rint = int(std::floor(R));
  const double R2 = R * R;
  for (int ad = 0; ad <= rint; ++ad) {
    for (int dy = 0; dy <= rint; ++dy) {
      const double v = R2 - double(dy) * double(dy) - double(ad) * double(ad);
      out.push_back(v >= 0 ? std::sqrt(v) : std::nan(""));
    }
    out.push_back(std::nan(""));
  }
}

// The tables of `radii`, concatenated, uploaded to `slot`.  `host` keeps the
// values alive until the upload has run (the caller synchronises before it
// goes out of scope).  Returns the device base.
// A table already in place for the same single radius (the slot is written
// by nothing else and is reallocated only to grow, that is for another
// radius) is not uploaded again; the no-alloc scratch is shared by other
// calls, so there the table is always uploaded.
const double* upload_htab(vo_context* ctx, int slot, const std::vector<double>& radii, std::vector<double>& host) {
  host.clear();
  for (double R : radii) append_htab(R, host);
  double* d = static_cast<double*>(ws(ctx, slot, host.size() * sizeof(double)));
  const bool cached = !ctx->arena.on && radii.size() == 1 && radii[0] == ctx->htab_R && d == ctx->htab_at;
  if (!cached) {
    ctx->htab_R = -1;  // (until the upload is in the stream)
    check(ctx, hipMemcpyAsync(d, host.data(), host.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream),
          "upload H table");
    if (!ctx->arena.on && radii.size() == 1) {
      ctx->htab_R = radii[0];
      ctx->htab_at = d;
    }
  }
  return d;
}

// The z-sweep form of pass 1 with 2R + 2 coverage bits per lane (one 64-bit
// word up to R = 31, two up to 63, four up to 127, eight up to 255).
constexpr int kZSweepMaxR = 255;
template <class T, int ECAP, int NB>
void launch_zsweep(vo_context* ctx, int rint, int64_t grid, const vo::Pass1ZArgs<T>& az) {
  const dim3 g(static_cast<unsigned>(grid)), b(static_cast<unsigned>(vo::kWave));
  if (rint <= 31)
    hipLaunchKernelGGL((vo::pass1_zsweep_kernel<T, ECAP, NB, kZWMax, 1>), g, b, 0, ctx->stream, az);
  else if (rint <= 63)
    hipLaunchKernelGGL((vo::pass1_zsweep_kernel<T, ECAP, NB, kZWMax, 2>), g, b, 0, ctx->stream, az);
  else if (rint <= 127)
    hipLaunchKernelGGL((vo::pass1_zsweep_kernel<T, ECAP, NB, kZWMax, 4>), g, b, 0, ctx->stream, az);
  else
    hipLaunchKernelGGL((vo::pass1_zsweep_kernel<T, ECAP, NB, kZWMax8, 8>), g, b, 0, ctx->stream, az);
}

template <class T>
void launch_zsweep_events(vo_context* ctx, int events, int rint, int64_t grid, const vo::Pass1ZArgs<T>& az) {
  switch (events) {
    case 512: launch_zsweep<T, 512, 256>(ctx, rint, grid, az); break;
    case 1024: launch_zsweep<T, 1024, 512>(ctx, rint, grid, az); break;
    case 1536: launch_zsweep<T, 1536, 512>(ctx, rint, grid, az); break;
    default: launch_zsweep<T, 768, 512>(ctx, rint, grid, az); break;
  }
}

// Grid of a batched launch over n tiles: one workgroup per tile, or, with a
// work queue, as many workgroups as the device keeps resident (at most n).
template <class K>
int64_t batch_grid(vo_context* ctx, K kernel, size_t lds, int64_t n, const unsigned* queue) {
  if (!queue) return n;
  const auto key = std::make_pair(reinterpret_cast<const void*>(kernel), lds);
  auto it = ctx->resident.find(key);
  if (it == ctx->resident.end()) {
    int per_cu = 0;
    check(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, vo::kWave, lds), "occupancy");
    it = ctx->resident.emplace(key, std::max(per_cu, 1)).first;
  }
  return std::max<int64_t>(1, std::min<int64_t>(n, int64_t(it->second) * std::max(ctx->num_cus, 1)));
}

// Batched z-sweep with the same event capacities, over the n global tiles
// from t0 on (one radius class of the batch); queue: 8 cleared counters for
// the persistent form, or null.
template <class T, int ECAP, int NB, int NW>
void launch_zsweep_batch_nw(vo_context* ctx, int64_t n, const vo::Pass1ZArgs<T>* jobs, const int64_t* tile_base,
                            int njobs, int64_t t0, unsigned* queue) {
  auto* k = queue ? vo::pass1_zsweep_batch_kernel<T, ECAP, NB, kZWMax, NW, true>
                  : vo::pass1_zsweep_batch_kernel<T, ECAP, NB, kZWMax, NW, false>;
  const dim3 g(static_cast<unsigned>(batch_grid(ctx, k, 0, n, queue))), b(static_cast<unsigned>(vo::kWave));
  hipLaunchKernelGGL(k, g, b, 0, ctx->stream, jobs, tile_base, njobs, t0, n, queue);
}

template <class T, int ECAP, int NB>
void launch_zsweep_batch(vo_context* ctx, int max_rint, int64_t n, const vo::Pass1ZArgs<T>* jobs,
                         const int64_t* tile_base, int njobs, int64_t t0, unsigned* queue) {
  if (max_rint <= 31)
    launch_zsweep_batch_nw<T, ECAP, NB, 1>(ctx, n, jobs, tile_base, njobs, t0, queue);
  else if (max_rint <= 63)
    launch_zsweep_batch_nw<T, ECAP, NB, 2>(ctx, n, jobs, tile_base, njobs, t0, queue);
  else
    launch_zsweep_batch_nw<T, ECAP, NB, 4>(ctx, n, jobs, tile_base, njobs, t0, queue);
}

template <class T>
void launch_zsweep_batch_events(vo_context* ctx, int events, int max_rint, int64_t n, const vo::Pass1ZArgs<T>* jobs,
                                const int64_t* tile_base, int njobs, int64_t t0, unsigned* queue) {
  switch (events) {
    case 512: launch_zsweep_batch<T, 512, 256>(ctx, max_rint, n, jobs, tile_base, njobs, t0, queue); break;
    case 1024: launch_zsweep_batch<T, 1024, 512>(ctx, max_rint, n, jobs, tile_base, njobs, t0, queue); break;
    case 1536: launch_zsweep_batch<T, 1536, 512>(ctx, max_rint, n, jobs, tile_base, njobs, t0, queue); break;
    default: launch_zsweep_batch<T, 768, 512>(ctx, max_rint, n, jobs, tile_base, njobs, t0, queue); break;
  }
}

// Radius class of a batched job: the pass-1 lane state (one coverage word up
// to R = 31, two up to 63, four up to 127) and, up to R = 31, the H table in
// LDS for the gather.
int radius_class(int rint) { return rint <= 31 ? 0 : (rint <= 63 ? 1 : 2); }

int events_for(const vo_context* ctx, int rint) {
  return ctx->zsweep_events ? ctx->zsweep_events : ctx->tuning.events[tune_bucket(rint)];
}

// Small radii fuse little, so their lists keep at least 24 intervals.
int ocap_for(const vo_context* ctx, int rint) {
  const int o = ctx->tuning.ocap[tune_bucket(rint)];
  return rint < 6 ? std::max(24, o) : o;
}

// The logical grid the passes work on.  Pass 1 sweeps along the ny
// positions of each of the nx lines, pass 2 across the lines; the CSR index
// of (line x, position y) is x * sx + y * sy.  Along y (the reference's
// axis, Voronoi2D.cpp sweeps each x-line) the logical grid is the physical
// one; along x it is transposed, and a pass-1 window of consecutive
// positions is then one contiguous range of the CSR.  Along x is used
// unless the rows are so short that 64-column tiles along them fill
// noticeably worse than tiles along y (VO_AXIS=x|y forces one).
struct Axis {
  int nx, ny;
  int64_t sx, sy;
};

Axis axis_for(const vo_context* ctx, int nx, int ny) {
  auto fill = [](int n) { return n <= 0 ? 1.0 : double(n) / (64.0 * ((n + 63) / 64)); };
  const bool along_x = ctx->axis_forced >= 0 ? ctx->axis_forced == 1 : fill(nx) >= 0.75 * fill(ny);
  if (along_x) return Axis{ny, nx, int64_t(nx), 1};
  return Axis{nx, ny, 1, int64_t(nx)};
}

template <class T>
void enqueue_input_check(vo_context* ctx, const std::vector<vo::InputCheck<T>>& jobs,
                         const std::vector<int64_t>& ncols, const std::vector<int>& job_ids);

// One dilation: `in` -> CSR in (out_off_slot, out_z_slot).
template <class T>
DevVol dilate(vo_context* ctx, const DevVol& in, double R, int out_off_slot, int out_z_slot, double* t1,
              double* t2) {
  // the passes work on a logical grid of nx lines of ny positions: pass 1
  // sweeps along the positions of each line, pass 2 across the lines
  const Axis ax = axis_for(ctx, in.nx, in.ny);
  const int nx = ax.nx, ny = ax.ny;
  const int64_t ncols = int64_t(nx) * ny;
  if (!(R >= 0) || !std::isfinite(R)) fail(ctx, VO_ERR_ARGUMENT, "radius must be finite and >= 0");
  if (R > 60000) fail(ctx, VO_ERR_ARGUMENT, "radius too large");
  const int rint = int(std::floor(R));
  const double R2 = R * R;
  int* status = static_cast<int*>(ws(ctx, kSTATUS, 64));
  int32_t* m_cnt = static_cast<int32_t*>(ws(ctx, kM_CNT, ncols * sizeof(int32_t)));
  int64_t* o_start = static_cast<int64_t*>(ws(ctx, kO_START, ncols * sizeof(int64_t)));
  int64_t* o_len = static_cast<int64_t*>(ws(ctx, kO_LEN, (ncols + 1) * sizeof(int64_t)));
  int64_t* out_off = static_cast<int64_t*>(ws(ctx, out_off_slot, (ncols + 1) * sizeof(int64_t)));
  // a pending input check of `in` (this is the operation's first dilation):
  // pass 1 does it when the z-sweep form runs, else its own kernel
  int* chk_word = ctx->fused.word;
  const int chk_range = ctx->fused.range;
  const double chk_lo = ctx->fused.lo, chk_hi = ctx->fused.hi;
  ctx->fused.word = nullptr;
  bool chk_init = chk_word != nullptr;  // the word is cleared by the first pass 1 that checks
  DevVol out;
  out.nx = in.nx;
  out.ny = in.ny;
  out.off = out_off;
  if (ncols == 0) {
    check(ctx, hipMemsetAsync(out_off, 0, sizeof(int64_t), ctx->stream), "memset");
    out.z = ws(ctx, out_z_slot, 16);
    return out;
  }
  std::vector<double> htab_host;  // outlives the uploads (the gather is followed by a synchronisation)
  int chunk1 = std::min(Shapes<T>::p1_cmax, 32);
  int chunk2 = Shapes<T>::p2_cmax;
  bool zsweep_failed = false;  // a window too dense for the z-sweep form: sweep along y instead
  bool paint_big = false;      // the painter's piece lists overflowed: its large-list form
  int64_t fast_total = -1;     // the output size, when the gather's one synchronisation settled it
  int64_t gather_end = 0;      // the gather form's part of the output arena (interleaved columns)
  bool copied = false;         // the output CSR was copied before the gather's synchronisation
  // the per-column capacities as this call found them: a call that gives up
  // leaves them so, so that failing calls do not compound their growth
  const int caps_at_entry = ctx->p1_caps, p2cap_at_entry = ctx->p2_cap_per_col;
  for (int attempt = 0;; ++attempt) {
    const int nch1 = (ny + chunk1 - 1) / chunk1;
    const int64_t grid1 = int64_t(nx) * nch1;
    const int caps = ctx->p1_caps;
    const int64_t slots = ncols * caps;
    T* m_lo = static_cast<T*>(ws(ctx, kM_LO, slots * sizeof(T)));
    T* m_hi = static_cast<T*>(ws(ctx, kM_HI, slots * sizeof(T)));
    uint16_t* m_dy = static_cast<uint16_t*>(ws(ctx, kM_DY, slots * sizeof(uint16_t)));
    T* s_lo = static_cast<T*>(ws(ctx, kM_SLO, size_t(slots / vo::kSlotBlock) * sizeof(T)));
    T* s_hi = static_cast<T*>(ws(ctx, kM_SHI, size_t(slots / vo::kSlotBlock) * sizeof(T)));
    T* c_hi = static_cast<T*>(ws(ctx, kM_CHI, size_t(ncols) * sizeof(T)));
    // pass 2: gather tiles of (x, 64 consecutive y), and the sweep form,
    // either over the whole volume in chunks of chunk2 columns of a row, or
    // (list mode) over the gather tiles that overflowed, one column a block
    const bool gather = rint <= vo::kGatherMaxR && !ctx->force_sweep;
    const int64_t ntiles = int64_t(nx) * ((ny + vo::kGatherTile - 1) / vo::kGatherTile);
    int gocap = ocap_for(ctx, rint);
    const int chunk2_eff = gather ? 1 : chunk2;
    const int64_t arena2 = int64_t(chunk2_eff) * ctx->p2_cap_per_col;  // endpoints per sweep block
    const int64_t sweep_blocks = gather ? ntiles * vo::kGatherTile : int64_t(ny) * ((nx + chunk2 - 1) / chunk2);
    int64_t garena = 0;  // endpoints per gather tile
    T* o_z = nullptr;
    auto size_arena = [&] {
      garena = gather ? int64_t(vo::kGatherTile) * 2 * gocap : 0;
      o_z = static_cast<T*>(ws(ctx, kO_Z, size_t(ntiles * garena + sweep_blocks * arena2) * sizeof(T)));
    };
    size_arena();
    int* tiles = static_cast<int*>(ws(ctx, kTILES, (ntiles + 1) * sizeof(int)));
    if (grid1 > INT32_MAX || ntiles > INT32_MAX || sweep_blocks > INT32_MAX)
      fail(ctx, VO_ERR_ARGUMENT, "grid too large");
    const bool zsweep = ctx->pass1_form == 0 && rint <= kZSweepMaxR && !zsweep_failed &&
                        vo::zsweep_offsets_fit(caps, ny, int(sizeof(T)));
    if (chk_init && !zsweep) {  // the check's own kernel, words and copy
      vo::InputCheck<T> c{in.off, static_cast<const T*>(in.z), chk_range, chk_lo, chk_hi};
      enqueue_input_check<T>(ctx, {c}, {ncols}, {0});
      chk_word = nullptr;
      chk_init = false;
    }
    const bool chk_here = chk_init && zsweep;
    hipLaunchKernelGGL(vo::reset_words_kernel, dim3(1), dim3(1), 0, ctx->stream, status, tiles, o_len + ncols,
                       chk_here ? chk_word : nullptr);
    check(ctx, hipGetLastError(), "reset words");

    vo::Pass1Args<T> a1{nx, ny, rint, chunk1, nch1, in.off, static_cast<const T*>(in.z), m_lo, m_hi, m_dy,
                        caps, m_cnt, status, ax.sx, ax.sy};
    check(ctx, hipEventRecord(ctx->ev[0], ctx->stream), "event");
    if (zsweep) {
      vo::Pass1ZArgs<T> az{nx, ny, rint, in.off, static_cast<const T*>(in.z), m_lo, m_hi, m_dy, caps, m_cnt,
                           status, s_lo, s_hi, c_hi, ax.sx, ax.sy};
      if (chk_here) {
        az.check_word = chk_word;
        az.check_range = chk_range;
        az.check_lo = chk_lo;
        az.check_hi = chk_hi;
        chk_init = false;
      }
      const int64_t gridz = int64_t(nx) * ((ny + vo::kZTile - 1) / vo::kZTile);
      launch_zsweep_events<T>(ctx, events_for(ctx, rint), rint, gridz, az);
    } else if (ctx->pass1_form == 2) {
      hipLaunchKernelGGL((vo::pass1_kernel<T, kP1Cap, kP1NCap, Shapes<T>::p1_arena, Shapes<T>::p1_cmax>),
                         dim3(unsigned(grid1)), dim3(vo::kWave), 0, ctx->stream, a1);
    } else if (paint_big) {
      hipLaunchKernelGGL((vo::pass1_paint_kernel<T, Shapes<T>::p1_cap_big, Shapes<T>::p1_arena, Shapes<T>::p1_cmax>),
                         dim3(unsigned(grid1)), dim3(vo::kWave), 0, ctx->stream, a1);
    } else {
      hipLaunchKernelGGL((vo::pass1_paint_kernel<T, kP1Cap, Shapes<T>::p1_arena, Shapes<T>::p1_cmax>),
                         dim3(unsigned(grid1)), dim3(vo::kWave), 0, ctx->stream, a1);
    }
    check(ctx, hipGetLastError(), "pass1 launch");
    check(ctx, hipEventRecord(ctx->ev[1], ctx->stream), "event");
    int nfall = 0;
    if (gather) {
      const double* ht = upload_htab(ctx, kG_HTAB, {R}, htab_host);
      for (;;) {
        double* lists =
            static_cast<double*>(ws(ctx, kG_LISTS, size_t(ntiles) * 2 * gocap * vo::kWave * sizeof(double)));
        // the summaries exist only when the z-sweep form wrote pass 1
        vo::Pass2GArgs<T> ag{nx, ny, rint, R2, m_lo, m_hi, m_dy, caps, m_cnt, o_z, garena, o_start, o_len,
                             tiles, status, gocap, ht, 0, lists, zsweep ? s_lo : nullptr, zsweep ? s_hi : nullptr, c_hi,
                             ax.sx, ax.sy};
        if (rint <= vo::kGatherLdsTableMax)
          hipLaunchKernelGGL((vo::pass2_gather_kernel<T, true>), dim3(unsigned(ntiles)), dim3(vo::kWave),
                             size_t(rint + 1) * vo::htab_width(rint) * sizeof(double), ctx->stream, ag);
        else
          hipLaunchKernelGGL((vo::pass2_gather_kernel<T, false>), dim3(unsigned(ntiles)), dim3(vo::kWave), 0,
                             ctx->stream, ag);
        check(ctx, hipGetLastError(), "pass2 gather launch");
        // The output scan is launched before the one synchronisation that
        // reads the overflow count and the status: when neither asks for
        // more work, its total is final and the dilation is done.
        check(ctx, hipEventRecord(ctx->ev[2], ctx->stream), "event");
        exclusive_scan(ctx, o_len, out_off, ncols + 1);
        hipLaunchKernelGGL(vo::report_words_kernel, dim3(1), dim3(1), 0, ctx->stream, out_off + ncols, status, tiles,
                           chk_word, ctx->host_words);
        check(ctx, hipGetLastError(), "report words");
        // The CSR copy goes before the synchronisation too, into an output
        // sized like the last result plus a quarter (columns that would not
        // fit are skipped): when the result fits, the dilation is complete at
        // the synchronisation.  (Not with the no-alloc scratch, whose output
        // buffer is the caller's.)
        int64_t spec = 0;
        if (!ctx->arena.on) {
          spec = std::max<int64_t>(ctx->spec_out, 16);
          T* oz = static_cast<T*>(ws(ctx, out_z_slot, size_t(spec) * sizeof(T)));
          hipLaunchKernelGGL(vo::gather_kernel<T>, dim3(blocks_for(ncols, 256)), dim3(256), 0, ctx->stream, ncols,
                             static_cast<const T*>(o_z), o_start, o_len, out_off, oz, ntiles * garena, spec);
          check(ctx, hipGetLastError(), "gather launch");
        }
        sync(ctx);
        std::memcpy(&nfall, ctx->host_words + 2, sizeof(int));
        if (nfall == 0 && int(ctx->host_words[1] & 0xffffffff) == 0) {
          fast_total = ctx->host_words[0];
          copied = fast_total <= spec;
        }
        // lists outgrew their capacity: the whole gather again with twice the
        // room is cheaper and more robust than the sweep form for many tiles,
        // as long as the lists and the output arena of all tiles stay within
        // kGatherRoomBytes (a column with hundreds of output intervals, on a
        // small grid, gets the room; a large grid keeps the sweep form for
        // its few dense tiles)
        const size_t room = size_t(ntiles) * vo::kWave * 2 * size_t(2 * gocap) * (sizeof(double) + sizeof(T));
        if (nfall == 0 || room > kGatherRoomBytes) break;
        gocap *= 2;
        size_arena();
        check(ctx, hipMemsetAsync(tiles, 0, sizeof(int), ctx->stream), "memset tiles");
      }
    }
    // the sweep form: the whole volume, or the tiles the gather form gave back
    const int64_t grid2 = gather ? int64_t(nfall) * vo::kGatherTile : sweep_blocks;
    if (grid2 > 0) {
      vo::Pass2Args<T> a2{nx, ny, rint, R2, chunk2_eff, 0, m_lo, m_hi, m_dy, caps, m_cnt, o_z, arena2, o_start,
                          o_len, status, gather ? tiles : nullptr, ntiles * garena, ax.sx, ax.sy};
      hipLaunchKernelGGL((vo::pass2_kernel<T, kP2SCap, kP2PCap, kP2NCap, kP2OCap, Shapes<T>::p2_farena,
                                           Shapes<T>::p2_cmax>),
                         dim3(unsigned(grid2)), dim3(vo::kWave), 0, ctx->stream, a2);
      check(ctx, hipGetLastError(), "pass2 launch");
    }
    ctx->last_fallback_tiles = nfall;
    gather_end = ntiles * garena;
    if (fast_total >= 0) {
      if (t1 || t2) {
        float ms1 = 0, ms2 = 0;
        check(ctx, hipEventElapsedTime(&ms1, ctx->ev[0], ctx->ev[1]), "elapsed");
        check(ctx, hipEventElapsedTime(&ms2, ctx->ev[1], ctx->ev[2]), "elapsed");
        if (t1) *t1 = ms1;
        if (t2) *t2 = ms2;
      }
      break;
    }
    check(ctx, hipEventRecord(ctx->ev[2], ctx->stream), "event");
    check(ctx, hipMemsetAsync(o_len + ncols, 0, sizeof(int64_t), ctx->stream), "memset");
    check(ctx, hipMemcpyAsync(ctx->host_words + 1, status, sizeof(int), hipMemcpyDeviceToHost, ctx->stream),
          "copy status");
    if (chk_word)
      check(ctx, hipMemcpyAsync(ctx->host_words + 3, chk_word, 2 * sizeof(int), hipMemcpyDeviceToHost, ctx->stream),
            "copy input check");
    sync(ctx);
    const int st = int(ctx->host_words[1] & 0xffffffff);
    if (st == 0) {
      if (t1 || t2) {
        float ms1 = 0, ms2 = 0;
        check(ctx, hipEventElapsedTime(&ms1, ctx->ev[0], ctx->ev[1]), "elapsed");
        check(ctx, hipEventElapsedTime(&ms2, ctx->ev[1], ctx->ev[2]), "elapsed");
        if (t1) *t1 = ms1;
        if (t2) *t2 = ms2;
      }
      break;
    }
    if ((st & vo::kStatusP1List) && ctx->pass1_form == 0 && !zsweep_failed && zsweep) {
      zsweep_failed = true;
      continue;
    }
    if ((st & vo::kStatusP1List) && !zsweep && ctx->pass1_form != 2 && !paint_big) {
      paint_big = true;  // one wavefront per CU, for the rare windows that need it
      continue;
    }
    if (st & vo::kStatusListOverflow) fail(ctx, VO_ERR_CAPACITY, status_text(st));
    // an arena overflowed: more room for what overflowed, and again
    if (attempt >= 6) {
      ctx->p1_caps = caps_at_entry;
      ctx->p2_cap_per_col = p2cap_at_entry;
      fail(ctx, VO_ERR_CAPACITY, "arena overflow after retries");
    }
    // (the sweep forms' LDS arenas want shorter chunks, the pass-1 slot
    // layout more slots per column, the sweep form of pass 2 a larger
    // output arena per column; the gather sizes its own)
    if (st & vo::kStatusArenaOverflow) {
      if (chunk1 > 8) chunk1 /= 2;
      if (chunk2 > 8) chunk2 /= 2;
    }
    if (st & vo::kStatusSlotOverflow) ctx->p1_caps *= 2;
    if (st & vo::kStatusP2Arena) ctx->p2_cap_per_col *= 2;
  }
  // output CSR (scanned already when the gather settled it)
  int64_t total = fast_total;
  if (total < 0) {
    exclusive_scan(ctx, o_len, out_off, ncols + 1);
    total = read_word(ctx, out_off + ncols);
  }
  T* oz = static_cast<T*>(ws(ctx, out_z_slot, size_t(std::max<int64_t>(total, 1)) * sizeof(T)));
  if (!copied) {
    hipLaunchKernelGGL(vo::gather_kernel<T>, dim3(blocks_for(ncols, 256)), dim3(256), 0, ctx->stream, ncols,
                       static_cast<const T*>(slot_ptr(ctx, kO_Z)), o_start, o_len, out_off, oz, gather_end,
                       std::max<int64_t>(total, 0));
    check(ctx, hipGetLastError(), "gather launch");
  }
  if (!ctx->arena.on) ctx->spec_out = total + total / 4;
  out.z = oz;
  out.count = total;
  ctx->last_outputs = total;
  return out;
}

// Device-side input check of the resident volumes of this call (device
// entry points; the host-buffer entry point checks on the host).  Nothing
// waits for it here: the result is copied to host_words[3] and looked at by
// every later synchronisation of the call (sync()), before anything else it
// reads, so a bad input fails with VO_ERR_ARGUMENT / VO_ERR_RANGE.  The
// kernels of the operation are memory-safe on any input whose offsets index
// the endpoint buffer, so they may run meanwhile.
template <class T>
vo::InputCheck<T> input_check_of(int op, double R, const vo_volume& v) {
  vo::InputCheck<T> c{v.offsets, static_cast<const T*>(v.z), 0, 0.0, 0.0};
  if (op == VO_OP_EROSION || op == VO_OP_OPENING) {  // Voronoi.cpp:27-42: the complement's ring range
    c.range = 1;
    c.lo = v.z_lo - 1;
    c.hi = v.z_hi + 1;
  } else if (op == VO_OP_CLOSING) {  // the dilation first, then the same range
    c.range = 1;
    c.lo = v.z_lo - 1 + R;
    c.hi = v.z_hi + 1 - R;
  }
  return c;
}

template <class T>
void enqueue_input_check(vo_context* ctx, const std::vector<vo::InputCheck<T>>& jobs,
                         const std::vector<int64_t>& ncols, const std::vector<int>& job_ids) {
  const int K = int(jobs.size());
  ctx->check_jobs = job_ids;
  int* word = static_cast<int*>(ws(ctx, kV_WORD, 2 * sizeof(int)));
  hipLaunchKernelGGL(vo::init_check_words_kernel, dim3(1), dim3(1), 0, ctx->stream, word);
  check(ctx, hipGetLastError(), "input check words");
  ctx->check_base.assign(size_t(K) + 1, 0);
  for (int k = 0; k < K; ++k) ctx->check_base[k + 1] = ctx->check_base[k] + ncols[k];
  const int64_t total = ctx->check_base[K];
  if (total > INT32_MAX) fail(ctx, VO_ERR_ARGUMENT, "batch too large");
  if (total > 0) {
    const dim3 grid(blocks_for(total, 256));
    if (K == 1) {
      // four wavefronts of 64 columns per block
      hipLaunchKernelGGL(vo::check_input_kernel<T>, dim3(blocks_for(total, 256)), dim3(256), 0, ctx->stream, jobs[0],
                         total, word);
    } else {
      const size_t tb = size_t(K) * sizeof(vo::InputCheck<T>);
      ctx->check_tab.resize(tb);
      std::memcpy(ctx->check_tab.data(), jobs.data(), tb);
      auto* dtab = static_cast<vo::InputCheck<T>*>(ws(ctx, kV_TAB, tb));
      auto* dbase = static_cast<int64_t*>(ws(ctx, kV_BASE, size_t(K + 1) * sizeof(int64_t)));
      check(ctx, hipMemcpyAsync(dtab, ctx->check_tab.data(), tb, hipMemcpyHostToDevice, ctx->stream), "upload");
      check(ctx, hipMemcpyAsync(dbase, ctx->check_base.data(), size_t(K + 1) * sizeof(int64_t),
                                hipMemcpyHostToDevice, ctx->stream),
            "upload");
      hipLaunchKernelGGL(vo::check_input_batch_kernel<T>, grid, dim3(256), 0, ctx->stream, dtab, dbase, K, word);
    }
    check(ctx, hipGetLastError(), "input check launch");
  }
  check(ctx, hipMemcpyAsync(ctx->host_words + 3, word, 2 * sizeof(int), hipMemcpyDeviceToHost, ctx->stream),
        "copy input check");
  ctx->check_input = true;
}

// The input check of an operation whose first step is a dilation of the
// input (dilation, closing): left to that dilation's pass 1, which reads
// every input column anyway (or to the check's own kernel when another pass-1
// form runs).  Until the dilation reports the words the host copy says "no
// failure", like the check kernel's words before its copy arrives.
template <class T>
void defer_input_check(vo_context* ctx, const vo::InputCheck<T>& c) {
  ctx->fused.word = static_cast<int*>(ws(ctx, kV_WORD, 2 * sizeof(int)));
  ctx->fused.range = c.range;
  ctx->fused.lo = c.lo;
  ctx->fused.hi = c.hi;
  ctx->check_jobs = {0};
  const int w[2] = {0, 0x7f7f7f7f};
  std::memcpy(ctx->host_words + 3, w, sizeof(w));
  ctx->check_input = true;
}

template <class T>
void check_input_of_op(vo_context* ctx, int op, double R, const vo_volume& v) {
  const vo::InputCheck<T> c = input_check_of<T>(op, R, v);
  if (op == VO_OP_DILATION || op == VO_OP_CLOSING)
    defer_input_check<T>(ctx, c);
  else
    enqueue_input_check<T>(ctx, {c}, {int64_t(v.nx) * v.ny}, {0});
}

// Per-column complement kernels (phase 0: lengths, phase 1: data): the ring
// complement grows the grid by one column on each side, the crop shrinks it.
template <class T>
DevVol column_transform(vo_context* ctx, bool ring, const DevVol& in, T lo, T hi, int off_slot, int z_slot) {
  const int onx = ring ? in.nx + 2 : in.nx - 2, ony = ring ? in.ny + 2 : in.ny - 2;
  const int64_t n = int64_t(onx) * ony;
  int64_t* len = static_cast<int64_t*>(ws(ctx, kO_LEN, (n + 1) * sizeof(int64_t)));
  int64_t* off = static_cast<int64_t*>(ws(ctx, off_slot, (n + 1) * sizeof(int64_t)));
  auto launch = [&](int phase, const int64_t* o, T* z) {
    const dim3 grid(blocks_for(std::max<int64_t>(n, 1), 256));
    vo::ColumnJob<T> job{ring ? in.nx : onx, ring ? in.ny : ony, in.off, static_cast<const T*>(in.z), lo, hi, len,
                         o, z};
    if (ring)
      hipLaunchKernelGGL(vo::ring_complement_kernel<T>, grid, dim3(256), 0, ctx->stream, job, phase);
    else
      hipLaunchKernelGGL(vo::crop_complement_kernel<T>, grid, dim3(256), 0, ctx->stream, job, phase);
    check(ctx, hipGetLastError(), "complement launch");
  };
  launch(0, nullptr, nullptr);
  check(ctx, hipMemsetAsync(len + n, 0, sizeof(int64_t), ctx->stream), "memset");
  exclusive_scan(ctx, len, off, n + 1);
  const int64_t total = read_word(ctx, off + n);
  T* z = static_cast<T*>(ws(ctx, z_slot, size_t(std::max<int64_t>(total, 1)) * sizeof(T)));
  launch(1, off, z);
  DevVol out;
  out.nx = onx;
  out.ny = ony;
  out.off = off;
  out.z = z;
  out.count = total;
  return out;
}

// erosion = crop-complement o dilation o ring-complement (Voronoi.cpp:8-17).
template <class T>
DevVol erode(vo_context* ctx, const DevVol& in, double R, double zlo, double zhi, int work_a, int work_b,
             int out_slot, double* t1, double* t2) {
  const double lo = zlo - 1, hi = zhi + 1;
  DevVol g = column_transform<T>(ctx, true, in, T(lo), T(hi), work_a, work_a + 1);
  DevVol d = dilate<T>(ctx, g, R, work_b, work_b + 1, t1, t2);
  return column_transform<T>(ctx, false, d, T(lo + 1), T(hi - 1), out_slot, out_slot + 1);
}

template <class T>
DevVol run_op(vo_context* ctx, int op, const DevVol& in, double R, double zlo, double zhi, double* t1,
              double* t2) {
  switch (op) {
    case VO_OP_DILATION:
      return dilate<T>(ctx, in, R, kV3_OFF, kV3_Z, t1, t2);
    case VO_OP_EROSION:
      return erode<T>(ctx, in, R, zlo, zhi, kV0_OFF, kV1_OFF, kV3_OFF, t1, t2);
    case VO_OP_CLOSING: {
      DevVol d = dilate<T>(ctx, in, R, kV2_OFF, kV2_Z, t1, t2);
      return erode<T>(ctx, d, R, zlo, zhi, kV0_OFF, kV1_OFF, kV3_OFF, t1, t2);
    }
    case VO_OP_OPENING: {
      DevVol e = erode<T>(ctx, in, R, zlo, zhi, kV0_OFF, kV1_OFF, kV2_OFF, t1, t2);
      return dilate<T>(ctx, e, R, kV3_OFF, kV3_Z, t1, t2);
    }
    default:
      fail(ctx, VO_ERR_ARGUMENT, "unknown operation");
  }
  return in;
}

void validate_shape(vo_context* ctx, const vo_volume* v);

// ---------------------------------------------------------------------------
// Batched form: many volumes per launch.  Every primitive (ring complement,
// dilation, crop complement) runs once per stage for all the volumes that
// need it; a volume whose lists overflow in the batched kernels is redone
// alone by the single-volume path, which has the sweep fallbacks.

template <class V>
V* upload_table(vo_context* ctx, int slot, const std::vector<V>& host) {
  V* d = static_cast<V*>(ws(ctx, slot, host.size() * sizeof(V)));
  check(ctx, hipMemcpyAsync(d, host.data(), host.size() * sizeof(V), hipMemcpyHostToDevice, ctx->stream),
        "upload table");
  return d;
}

// Keeps a copy of a single-volume result that must outlive later calls.
template <class T>
DevVol keep_alone(vo_context* ctx, const DevVol& r) {
  const int64_t ncols = int64_t(r.nx) * r.ny;
  const int64_t count = read_word(ctx, r.off + ncols);  // 0-based offsets
  void *off = nullptr, *z = nullptr;
  check(ctx, hipMalloc(&off, size_t(ncols + 1) * sizeof(int64_t)), "hipMalloc");
  ctx->batch_extra.push_back(off);
  check(ctx, hipMalloc(&z, size_t(std::max<int64_t>(count, 1)) * sizeof(T)), "hipMalloc");
  ctx->batch_extra.push_back(z);
  check(ctx, hipMemcpyAsync(off, r.off, size_t(ncols + 1) * sizeof(int64_t), hipMemcpyDeviceToDevice, ctx->stream),
        "copy");
  if (count)
    check(ctx, hipMemcpyAsync(z, r.z, size_t(count) * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream), "copy");
  DevVol k = r;
  k.off = static_cast<const int64_t*>(off);
  k.z = z;
  k.count = count;
  return k;
}

// The programs of the four operations (Voronoi.cpp:8-17, offset3d.cpp:124-133).
enum BatchStep { kStepNone = -1, kStepRing = 0, kStepDilate = 1, kStepCrop = 2 };

int batch_step(int op, int s) {
  static const int prog[5][4] = {
      {kStepNone, kStepNone, kStepNone, kStepNone},         // noop
      {kStepDilate, kStepNone, kStepNone, kStepNone},       // dilation
      {kStepRing, kStepDilate, kStepCrop, kStepNone},       // erosion
      {kStepDilate, kStepRing, kStepDilate, kStepCrop},     // closing
      {kStepRing, kStepDilate, kStepCrop, kStepDilate}};    // opening
  return prog[op][s];
}

void free_batch_extra(vo_context* ctx) {
  if (ctx->batch_extra.empty()) return;
  check(ctx, hipStreamSynchronize(ctx->stream), "sync");
  for (void* p : ctx->batch_extra) (void)hipFree(p);
  ctx->batch_extra.clear();
}

// ---------------------------------------------------------------------------
// Batched form as one program (no reference counterpart: it replaces a loop
// of VoronoiMorpho calls over many volumes, offset3d.cpp:116-136).  Each
// job's operation is a program over up to four stages of primitives -- ring
// complement, dilation, crop complement (Voronoi.cpp:8-17,
// offset3d.cpp:124-133) -- and every primitive runs once per stage for all
// the jobs that need it.  Every output size is bounded before anything runs
// (a complement adds at most two endpoints per column; a dilated column holds
// at most the gather lists' capacity of intervals), so all buffers and all
// kernel tables are set up front: the program reads the input sizes (one
// synchronisation), then runs to its end with nothing for the host to wait
// on.  A batch equal to the previous one (same jobs, pointers and sizes)
// runs from a HIP graph captured on its second call.  Jobs whose lists
// overflowed (a status bit) and jobs outside the batched kernels' range
// (R > 127, empty grids, forced kernel forms) go through the single-volume
// path afterwards.

// Kernel tables of a program, uploaded in one copy.
struct BlobBuilder {
  std::vector<char> host;
  template <class V>
  size_t put(const std::vector<V>& v) {
    const size_t at = (host.size() + 255) & ~size_t(255);
    host.resize(at + std::max<size_t>(v.size() * sizeof(V), 1));
    if (!v.empty()) std::memcpy(host.data() + at, v.data(), v.size() * sizeof(V));
    return at;
  }
};

void fill_words(vo_context* ctx, void* p, int64_t words, uint32_t v) {
  if (words <= 0) return;
  hipLaunchKernelGGL(vo::fill_kernel, dim3(blocks_for(words, 256)), dim3(256), 0, ctx->stream,
                     static_cast<uint32_t*>(p), words, v);
  check(ctx, hipGetLastError(), "fill launch");
}

// One primitive of one stage for several jobs.
struct ProgStep {
  int kind = kStepNone;
  std::vector<int> jobs;          // dilation: grouped by radius class
  std::vector<int64_t> col_base;  // output columns
  std::vector<int64_t> tile_base; // dilation: 64-position tiles
  int64_t C = 0, NT = 0, zcap = 0;
  int off_slot = 0, z_slot = 0;
  int64_t* off = nullptr;  // output CSR over the step's columns (absolute into z)
  void* z = nullptr;
  size_t t0 = 0, t1 = 0, t2 = 0, t3 = 0;  // table offsets in the blob
  int64_t cls_t0[3] = {0, 0, 0}, cls_t1[3] = {0, 0, 0};
  int cls_rmax[3] = {0, 0, 0};
};

// Whether job j's dilations fit the batched kernels (grid g at each one).
bool batchable(vo_context* ctx, int op, double R, int nx, int ny, int tsize) {
  if (op == VO_OP_NOOP) return true;  // only rebased
  const int rint = int(std::floor(R));
  if (rint > 127 || ctx->force_sweep || ctx->pass1_form != 0) return false;
  for (int s = 0; s < 4; ++s) {
    const int k = batch_step(op, s);
    if (k == kStepRing) {
      nx += 2;
      ny += 2;
    } else if (k == kStepCrop) {
      nx -= 2;
      ny -= 2;
    } else if (k == kStepDilate && !vo::zsweep_offsets_fit(ctx->p1_caps, axis_for(ctx, nx, ny).ny, tsize)) {
      return false;
    }
  }
  return true;
}

template <class T>
void morph_batch_device(vo_context* ctx, int n, const vo_job* jobs, vo_volume* out, int64_t* out_count, double* t1,
                        double* t2) {
  if (n < 0 || (n > 0 && (!jobs || !out))) fail(ctx, VO_ERR_ARGUMENT, "invalid batch");
  free_batch_extra(ctx);
  ctx->last_batch_alone = 0;
  std::vector<int> op(n);
  std::vector<double> R(n), zlo(n), zhi(n);
  std::vector<DevVol> vol(n);
  for (int j = 0; j < n; ++j) {
    const vo_job& jb = jobs[j];
    validate_shape(ctx, &jb.in);
    if (jb.op < VO_OP_NOOP || jb.op > VO_OP_OPENING) fail(ctx, VO_ERR_ARGUMENT, "unknown operation");
    if (!(jb.radius >= 0) || !std::isfinite(jb.radius) || jb.radius > 60000)
      fail(ctx, VO_ERR_ARGUMENT, "radius must be finite and >= 0");
    if ((jb.op == VO_OP_EROSION || jb.op == VO_OP_OPENING || jb.op == VO_OP_CLOSING) && !(jb.in.z_lo <= jb.in.z_hi))
      fail(ctx, VO_ERR_ARGUMENT, "z_lo must not exceed z_hi");
    op[j] = jb.op;
    R[j] = jb.radius;
    zlo[j] = jb.in.z_lo;
    zhi[j] = jb.in.z_hi;
    vol[j].nx = jb.in.nx;
    vol[j].ny = jb.in.ny;
    vol[j].off = jb.in.offsets;
    vol[j].z = jb.in.z;
  }
  if (t1) *t1 = 0;
  if (t2) *t2 = 0;
  {
    std::vector<vo::InputCheck<T>> chk;
    std::vector<int64_t> nc;
    std::vector<int> ids;
    for (int j = 0; j < n; ++j) {
      if (jobs[j].op == VO_OP_NOOP) continue;  // passed through untouched
      chk.push_back(input_check_of<T>(jobs[j].op, jobs[j].radius, jobs[j].in));
      nc.push_back(int64_t(jobs[j].in.nx) * jobs[j].in.ny);
      ids.push_back(j);
    }
    if (!chk.empty()) enqueue_input_check<T>(ctx, chk, nc, ids);
  }
  if (n == 0) return;
  const int tsize = int(sizeof(T));
  std::vector<char> inprog(n, 0);
  for (int j = 0; j < n; ++j) inprog[j] = batchable(ctx, op[j], R[j], vol[j].nx, vol[j].ny, tsize) ? 1 : 0;

  // 1. input endpoint counts: the program's one synchronisation before its end
  if (ctx->batch_host_n < size_t(3 * n + 8)) {
    if (ctx->batch_host) {
      check(ctx, hipStreamSynchronize(ctx->stream), "sync");
      check(ctx, hipHostFree(ctx->batch_host), "hipHostFree");
      ctx->batch_host = nullptr;
    }
    const size_t want = size_t(3 * n + 8) * 2;
    check(ctx, hipHostMalloc(reinterpret_cast<void**>(&ctx->batch_host), want * sizeof(int64_t), hipHostMallocDefault),
          "pinned");
    ctx->batch_host_n = want;
  }
  int64_t* hb = ctx->batch_host;
  std::vector<int64_t> E(n, 0);
  {
    std::vector<vo::FinalJob> cj(n);
    for (int j = 0; j < n; ++j) cj[j] = vo::FinalJob{vol[j].off, int64_t(vol[j].nx) * vol[j].ny, 0};
    const vo::FinalJob* dcj = upload_table(ctx, kB_TAB1, cj);
    int64_t* cnt = static_cast<int64_t*>(ws(ctx, kB_FIN_AUX, size_t(3 * n) * sizeof(int64_t)));
    hipLaunchKernelGGL(vo::job_counts_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, dcj, n, cnt);
    check(ctx, hipGetLastError(), "count launch");
    check(ctx, hipMemcpyAsync(hb, cnt, size_t(n) * sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream), "copy");
    sync(ctx);  // also where a bad input is reported
    for (int j = 0; j < n; ++j) E[j] = hb[j];
  }

  // 2. the program's steps, with every output bounded
  int ocap = 0;
  for (int j = 0; j < n; ++j)
    if (inprog[j] && op[j] != VO_OP_NOOP) ocap = std::max(ocap, ocap_for(ctx, int(std::floor(R[j]))));
  struct Cur {
    int nx, ny;
    int64_t E;
  };
  std::vector<Cur> cur(n);
  for (int j = 0; j < n; ++j) cur[j] = Cur{vol[j].nx, vol[j].ny, E[j]};
  std::vector<ProgStep> steps;
  int64_t maxC = 0, maxDilC = 0, maxNT = 0;
  for (int s = 0; s < 4; ++s) {
    for (int kind : {int(kStepRing), int(kStepDilate), int(kStepCrop)}) {
      ProgStep st;
      st.kind = kind;
      for (int j = 0; j < n; ++j)
        if (inprog[j] && batch_step(op[j], s) == kind) st.jobs.push_back(j);
      if (st.jobs.empty()) continue;
      if (kind == kStepDilate)  // grouped by radius class: one range of tiles per class
        std::stable_sort(st.jobs.begin(), st.jobs.end(), [&](int a, int b) {
          return radius_class(int(std::floor(R[a]))) < radius_class(int(std::floor(R[b])));
        });
      const int K = int(st.jobs.size());
      st.col_base.assign(K + 1, 0);
      st.tile_base.assign(K + 1, 0);
      for (int k = 0; k < K; ++k) {
        Cur& c = cur[st.jobs[k]];
        if (kind == kStepRing) {
          c.nx += 2;
          c.ny += 2;
        } else if (kind == kStepCrop) {
          c.nx -= 2;
          c.ny -= 2;
        }
        const int64_t cols = int64_t(c.nx) * c.ny;
        c.E = kind == kStepDilate ? 2 * int64_t(ocap) * cols : c.E + 2 * cols;
        st.col_base[k + 1] = st.col_base[k] + cols;
        st.zcap += c.E;
        if (kind == kStepDilate) {
          const Axis ax = axis_for(ctx, c.nx, c.ny);
          st.tile_base[k + 1] = st.tile_base[k] + int64_t(ax.nx) * ((ax.ny + vo::kGatherTile - 1) / vo::kGatherTile);
        }
      }
      st.C = st.col_base[K];
      st.NT = st.tile_base[K];
      if (st.C + 1 > INT32_MAX || st.NT > INT32_MAX) fail(ctx, VO_ERR_ARGUMENT, "batch too large");
      st.off_slot = kB_OUT + (s * 3 + kind) * 2;
      st.z_slot = st.off_slot + 1;
      maxC = std::max(maxC, st.C);
      if (kind == kStepDilate) {
        maxDilC = std::max(maxDilC, st.C);
        maxNT = std::max(maxNT, st.NT);
      }
      steps.push_back(std::move(st));
    }
  }

  // 3. workspace, sized for the largest step (steps run one after the other)
  const int caps = ctx->p1_caps;
  const int64_t garena = int64_t(vo::kGatherTile) * 2 * ocap;
  const int64_t list_tile = 2 * int64_t(ocap) * vo::kWave;
  for (auto& st : steps) {
    st.off = static_cast<int64_t*>(ws(ctx, st.off_slot, size_t(st.C + 1) * sizeof(int64_t)));
    st.z = ws(ctx, st.z_slot, size_t(std::max<int64_t>(st.zcap, 1)) * sizeof(T));
  }
  int64_t* len = static_cast<int64_t*>(ws(ctx, kO_LEN, size_t(maxC + 1) * sizeof(int64_t)));
  int* status = static_cast<int*>(ws(ctx, kB_STATUS, size_t(n) * sizeof(int)));
  // work-queue counters of the persistent form: eight per launch (two passes
  // and three radius classes per dilation step), cleared when the program starts
  int ndil = 0;
  for (auto& st : steps) ndil += st.kind == kStepDilate ? 1 : 0;
  const int64_t qwords = ctx->batch_queue ? int64_t(ndil) * 2 * 3 * vo::kXcds : 0;
  unsigned* queue = qwords ? static_cast<unsigned*>(ws(ctx, kB_QUEUE, size_t(qwords) * sizeof(unsigned))) : nullptr;
  auto queue_of = [&](int di, int pass, int c) { return queue ? queue + ((di * 2 + pass) * 3 + c) * vo::kXcds : nullptr; };
  T *m_lo = nullptr, *m_hi = nullptr, *s_lo = nullptr, *s_hi = nullptr, *c_hi = nullptr, *o_z = nullptr;
  uint16_t* m_dy = nullptr;
  int32_t* m_cnt = nullptr;
  int64_t* o_start = nullptr;
  double* lists = nullptr;
  if (maxDilC > 0) {
    m_lo = static_cast<T*>(ws(ctx, kM_LO, size_t(maxDilC) * caps * sizeof(T)));
    m_hi = static_cast<T*>(ws(ctx, kM_HI, size_t(maxDilC) * caps * sizeof(T)));
    m_dy = static_cast<uint16_t*>(ws(ctx, kM_DY, size_t(maxDilC) * caps * sizeof(uint16_t)));
    s_lo = static_cast<T*>(ws(ctx, kM_SLO, size_t(maxDilC) * (caps / vo::kSlotBlock) * sizeof(T)));
    s_hi = static_cast<T*>(ws(ctx, kM_SHI, size_t(maxDilC) * (caps / vo::kSlotBlock) * sizeof(T)));
    c_hi = static_cast<T*>(ws(ctx, kM_CHI, size_t(maxDilC) * sizeof(T)));
    m_cnt = static_cast<int32_t*>(ws(ctx, kM_CNT, size_t(maxDilC) * sizeof(int32_t)));
    o_z = static_cast<T*>(ws(ctx, kO_Z, size_t(maxNT) * garena * sizeof(T)));
    o_start = static_cast<int64_t*>(ws(ctx, kO_START, size_t(maxDilC) * sizeof(int64_t)));
    lists = static_cast<double*>(ws(ctx, kG_LISTS, size_t(maxNT) * list_tile * sizeof(double)));
  }
  {  // the scan's temporary storage, for the largest step
    size_t tmp = 0;
    check(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, len, len, int(maxC + 1), ctx->stream), "scan size");
    ws(ctx, kSCAN_TMP, tmp);
  }
  // final offsets of every program job, rebased to 0
  std::vector<int> pj;
  for (int j = 0; j < n; ++j)
    if (inprog[j]) pj.push_back(j);
  const int P = int(pj.size());
  std::vector<int64_t> entry_base(P + 1, 0);
  for (int k = 0; k < P; ++k)
    entry_base[k + 1] = entry_base[k] + int64_t(cur[pj[k]].nx) * cur[pj[k]].ny + 1;
  int64_t* rel = static_cast<int64_t*>(ws(ctx, kB_FIN_OFF, size_t(std::max<int64_t>(entry_base[P], 1)) * sizeof(int64_t)));
  int64_t* aux = static_cast<int64_t*>(ws(ctx, kB_FIN_AUX, size_t(3 * n) * sizeof(int64_t)));

  // 4. kernel tables (all pointers known now), in one blob
  BlobBuilder blob;
  std::vector<DevVol> v = vol;  // each job's current volume as the steps go
  for (auto& st : steps) {
    const int K = int(st.jobs.size());
    if (st.kind == kStepDilate) {
      std::vector<double> radii(K), htab;
      std::vector<int64_t> hbase(K, 0);
      for (int k = 0; k < K; ++k) {
        radii[k] = R[st.jobs[k]];
        const int rk = int(std::floor(radii[k]));
        if (k + 1 < K) hbase[k + 1] = hbase[k] + int64_t(rk + 1) * vo::htab_width(rk);
      }
      for (double Rk : radii) append_htab(Rk, htab);
      st.t3 = blob.put(htab);
      std::vector<vo::Pass1ZArgs<T>> a1(K);
      std::vector<vo::Pass2GArgs<T>> a2(K);
      for (int k = 0; k < K; ++k) {
        const int j = st.jobs[k];
        const DevVol& in = v[j];
        const int rint = int(std::floor(R[j]));
        const int64_t cb = st.col_base[k], sb = cb * caps, qb = cb * (caps / vo::kSlotBlock);
        const Axis ax = axis_for(ctx, in.nx, in.ny);
        // the H-table pointer is patched to the device blob at upload (index stays relative)
        a1[k] = vo::Pass1ZArgs<T>{ax.nx, ax.ny, rint, in.off, static_cast<const T*>(in.z), m_lo + sb, m_hi + sb, m_dy + sb,
                                  caps, m_cnt + cb, status + j, s_lo + qb, s_hi + qb, c_hi + cb, ax.sx, ax.sy};
        a2[k] = vo::Pass2GArgs<T>{ax.nx, ax.ny, rint, R[j] * R[j], m_lo + sb, m_hi + sb, m_dy + sb, caps,
                                  m_cnt + cb, o_z, garena, o_start + cb, len + cb, nullptr, status + j, ocap,
                                  reinterpret_cast<const double*>(st.t3 + size_t(hbase[k]) * sizeof(double)),
                                  st.tile_base[k] * garena, lists + st.tile_base[k] * list_tile, s_lo + qb, s_hi + qb,
                                  c_hi + cb, ax.sx, ax.sy};
        const int c = radius_class(rint);
        if (st.cls_t1[c] == 0) st.cls_t0[c] = st.tile_base[k];
        st.cls_t1[c] = st.tile_base[k + 1];
        st.cls_rmax[c] = std::max(st.cls_rmax[c], rint);
        v[j].off = st.off + cb;
        v[j].z = st.z;
      }
      st.t0 = blob.put(a1);
      st.t1 = blob.put(a2);
      st.t2 = blob.put(st.tile_base);
    } else {
      const bool ring = st.kind == kStepRing;
      std::vector<vo::ColumnJob<T>> t(K);
      for (int k = 0; k < K; ++k) {
        const int j = st.jobs[k];
        const DevVol& in = v[j];
        const double lo = ring ? zlo[j] - 1 : zlo[j], hi = ring ? zhi[j] + 1 : zhi[j];
        t[k] = vo::ColumnJob<T>{ring ? in.nx : in.nx - 2, ring ? in.ny : in.ny - 2, in.off, static_cast<const T*>(in.z),
                                T(lo), T(hi), len + st.col_base[k], st.off + st.col_base[k], static_cast<T*>(st.z)};
        v[j].nx = ring ? in.nx + 2 : in.nx - 2;
        v[j].ny = ring ? in.ny + 2 : in.ny - 2;
        v[j].off = st.off + st.col_base[k];
        v[j].z = st.z;
      }
      st.t0 = blob.put(t);
      st.t2 = blob.put(st.col_base);
    }
  }
  std::vector<vo::FinalJob> fj(P);
  for (int k = 0; k < P; ++k) {
    const int j = pj[k];
    fj[k] = vo::FinalJob{v[j].off, int64_t(v[j].nx) * v[j].ny, entry_base[k]};
  }
  const size_t tfj = blob.put(fj), teb = blob.put(entry_base);
  char* dblob = static_cast<char*>(ws(ctx, kB_TAB0, blob.host.size()));
  // the H-table pointers of the pass-2 tables are blob offsets: make them device addresses
  for (auto& st : steps) {
    if (st.kind != kStepDilate) continue;
    auto* a2 = reinterpret_cast<vo::Pass2GArgs<T>*>(blob.host.data() + st.t1);
    for (size_t k = 0; k < st.jobs.size(); ++k)
      a2[k].htab = reinterpret_cast<const double*>(dblob + reinterpret_cast<size_t>(a2[k].htab));
  }

  // 5. the program: enqueued, captured or replayed
  auto enqueue = [&] {
    fill_words(ctx, status, n, 0);
    if (queue) fill_words(ctx, queue, qwords, 0);
    int di = 0;  // dilation steps so far
    for (auto& st : steps) {
      const int K = int(st.jobs.size());
      const int64_t* dbase = reinterpret_cast<const int64_t*>(dblob + st.t2);
      if (st.kind == kStepDilate) {
        const auto* d1 = reinterpret_cast<const vo::Pass1ZArgs<T>*>(dblob + st.t0);
        const auto* d2 = reinterpret_cast<const vo::Pass2GArgs<T>*>(dblob + st.t1);
        // columns a failed job never writes keep length 0
        fill_words(ctx, len, 2 * (st.C + 1), 0);
        for (int c = 0; c < 3; ++c)
          if (st.cls_t1[c] > st.cls_t0[c])
            launch_zsweep_batch_events<T>(ctx, events_for(ctx, st.cls_rmax[c]), st.cls_rmax[c],
                                          st.cls_t1[c] - st.cls_t0[c], d1, dbase, K, st.cls_t0[c], queue_of(di, 0, c));
        check(ctx, hipGetLastError(), "batched pass1 launch");
        for (int c = 0; c < 3; ++c) {
          if (st.cls_t1[c] <= st.cls_t0[c]) continue;
          const int64_t nt = st.cls_t1[c] - st.cls_t0[c];
          unsigned* q = queue_of(di, 1, c);
          const dim3 block(vo::kWave);
          if (st.cls_rmax[c] <= vo::kGatherLdsTableMax) {
            auto* k = q ? vo::pass2_gather_batch_kernel<T, true, true> : vo::pass2_gather_batch_kernel<T, true, false>;
            const size_t lds = size_t(st.cls_rmax[c] + 1) * vo::htab_width(st.cls_rmax[c]) * sizeof(double);
            hipLaunchKernelGGL(k, dim3(unsigned(batch_grid(ctx, k, lds, nt, q))), block, lds, ctx->stream, d2, dbase, K,
                               st.cls_t0[c], nt, q);
          } else {
            auto* k = q ? vo::pass2_gather_batch_kernel<T, false, true> : vo::pass2_gather_batch_kernel<T, false, false>;
            hipLaunchKernelGGL(k, dim3(unsigned(batch_grid(ctx, k, 0, nt, q))), block, 0, ctx->stream, d2, dbase, K,
                               st.cls_t0[c], nt, q);
          }
        }
        check(ctx, hipGetLastError(), "batched pass2 launch");
        ++di;
        exclusive_scan(ctx, len, st.off, st.C + 1);
        if (st.C > 0) {  // (a step over empty grids only has its one offset)
          hipLaunchKernelGGL(vo::gather_kernel<T>, dim3(blocks_for(st.C, 256)), dim3(256), 0, ctx->stream, st.C,
                             static_cast<const T*>(o_z), o_start, len, st.off, static_cast<T*>(st.z), INT64_MAX,
                             INT64_MAX);
          check(ctx, hipGetLastError(), "gather launch");
        }
      } else {
        const auto* dt = reinterpret_cast<const vo::ColumnJob<T>*>(dblob + st.t0);
        const dim3 grid(blocks_for(std::max<int64_t>(st.C, 1), 256));
        auto launch = [&](int phase) {
          if (st.kind == kStepRing)
            hipLaunchKernelGGL(vo::ring_complement_batch_kernel<T>, grid, dim3(256), 0, ctx->stream, dt, dbase, K, phase);
          else
            hipLaunchKernelGGL(vo::crop_complement_batch_kernel<T>, grid, dim3(256), 0, ctx->stream, dt, dbase, K,
                               phase);
          check(ctx, hipGetLastError(), "batched complement launch");
        };
        launch(0);
        fill_words(ctx, len + st.C, 2, 0);
        exclusive_scan(ctx, len, st.off, st.C + 1);
        launch(1);
      }
    }
    if (P > 0) {
      hipLaunchKernelGGL(vo::finalize_offsets_kernel, dim3(blocks_for(entry_base[P], 256)), dim3(256), 0, ctx->stream,
                         reinterpret_cast<const vo::FinalJob*>(dblob + tfj),
                         reinterpret_cast<const int64_t*>(dblob + teb), P, rel, aux, aux + n);
      check(ctx, hipGetLastError(), "finalize launch");
    }
    check(ctx, hipMemcpyAsync(hb, aux, size_t(2 * n) * sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream), "copy");
    check(ctx, hipMemcpyAsync(hb + 2 * n, status, size_t(n) * sizeof(int), hipMemcpyDeviceToHost, ctx->stream), "copy");
  };
  // the graph key: the tables (every pointer and parameter the kernels see),
  // the launch shapes, and the buffers the program itself names
  std::vector<char> key = blob.host;
  auto add_key = [&](const void* p, size_t b) { key.insert(key.end(), static_cast<const char*>(p), static_cast<const char*>(p) + b); };
  for (auto& st : steps) {
    add_key(&st.C, sizeof(st.C));
    add_key(&st.NT, sizeof(st.NT));
    add_key(st.cls_t0, sizeof(st.cls_t0));
    add_key(st.cls_t1, sizeof(st.cls_t1));
    add_key(st.cls_rmax, sizeof(st.cls_rmax));
    add_key(&st.off, sizeof(st.off));
  }
  const void* bufs[] = {dblob, hb, len, status, rel, aux, ws(ctx, kSCAN_TMP, 16), queue};
  add_key(bufs, sizeof(bufs));
  for (int c = 0; c < kTuneBuckets; ++c) add_key(&ctx->tuning.events[c], sizeof(int));

  check(ctx, hipMemcpyAsync(dblob, blob.host.data(), blob.host.size(), hipMemcpyHostToDevice, ctx->stream), "upload");
  check(ctx, hipEventRecord(ctx->ev[0], ctx->stream), "event");
  BatchGraph& bg = ctx->bgraph;
  const bool same = bg.key == key;
  if (same && bg.exec) {
    check(ctx, hipGraphLaunch(bg.exec, ctx->stream), "graph launch");
  } else if (same && ctx->use_graphs) {  // second time: capture, then launch
    bool captured = false;
    if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) == hipSuccess) {
      bool ok = true;
      try {
        enqueue();
      } catch (const Failure&) {
        ok = false;
      }
      hipGraph_t g = nullptr;
      ok = hipStreamEndCapture(ctx->stream, &g) == hipSuccess && ok;
      hipGraphExec_t ex = nullptr;
      ok = ok && g && hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) == hipSuccess;
      if (g) (void)hipGraphDestroy(g);
      if (ok) {
        bg.exec = ex;
        check(ctx, hipGraphLaunch(bg.exec, ctx->stream), "graph launch");
        captured = true;
      } else if (ex) {
        (void)hipGraphExecDestroy(ex);
      }
    }
    if (!captured) {  // this device or runtime would not capture the program: run it as it is from now on
      (void)hipGetLastError();
      ctx->error.clear();
      ctx->use_graphs = false;
      enqueue();
    }
  } else {
    if (bg.exec) {
      (void)hipGraphExecDestroy(bg.exec);
      bg.exec = nullptr;
    }
    bg.key = std::move(key);
    enqueue();
  }
  check(ctx, hipEventRecord(ctx->ev[1], ctx->stream), "event");
  sync(ctx);
  float ms = 0;
  check(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]), "elapsed");
  if (t1) *t1 = ms;  // the whole program (passes and complements)

  // 6. results; failed and non-program jobs through the single-volume path
  const int* st_host = reinterpret_cast<const int*>(hb + 2 * n);
  std::vector<int64_t> first(n), count(n);
  for (int j = 0; j < n; ++j) {
    first[j] = hb[j];
    count[j] = hb[n + j];
  }
  for (int k = 0; k < P; ++k) {
    const int j = pj[k];
    out[j].nx = v[j].nx;
    out[j].ny = v[j].ny;
    out[j].offsets = rel + entry_base[k];
    out[j].z = static_cast<const T*>(v[j].z) + first[k];
    out[j].z_lo = zlo[j];
    out[j].z_hi = zhi[j];
    if (out_count) out_count[j] = count[k];
  }
  for (int j = 0; j < n; ++j) {
    if (inprog[j] && st_host[j] == 0) continue;
    if (op[j] == VO_OP_NOOP) {
      out[j] = jobs[j].in;
      if (out_count) out_count[j] = E[j];
      continue;
    }
    double a = 0, b = 0;
    DevVol r = run_op<T>(ctx, op[j], vol[j], R[j], zlo[j], zhi[j], &a, &b);
    if (t1) *t1 += a;
    if (t2) *t2 += b;
    DevVol kept = keep_alone<T>(ctx, r);
    out[j].nx = kept.nx;
    out[j].ny = kept.ny;
    out[j].offsets = kept.off;
    out[j].z = kept.z;
    out[j].z_lo = zlo[j];
    out[j].z_hi = zhi[j];
    if (out_count) out_count[j] = kept.count;
    ++ctx->last_batch_alone;
  }
  sync(ctx);
}

// ---------------------------------------------------------------------------
// Parity metrics (get_volume / numSegments, calculateXor) on device volumes.

template <class T>
void measure(vo_context* ctx, const vo_volume* v, double spacing, double* volume, int64_t* segments) {
  validate_shape(ctx, v);
  const int64_t n = int64_t(v->nx) * v->ny;
  double vol = 0;
  int64_t segs = 0;
  if (n > 0) {
    double* len = static_cast<double*>(ws(ctx, kX_LENGTH, size_t(n) * sizeof(double)));
    int64_t* cnt = static_cast<int64_t*>(ws(ctx, kX_SEGS, size_t(n) * sizeof(int64_t)));
    double* sum = static_cast<double*>(ws(ctx, kX_SUM, 2 * sizeof(double)));
    hipLaunchKernelGGL(vo::measure_kernel<T>, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, n, v->offsets,
                       static_cast<const T*>(v->z), len, cnt);
    check(ctx, hipGetLastError(), "measure launch");
    size_t tmp = 0;
    check(ctx, hipcub::DeviceReduce::Sum(nullptr, tmp, len, sum, int(n), ctx->stream), "reduce size");
    void* t = ws(ctx, kSCAN_TMP, tmp);
    check(ctx, hipcub::DeviceReduce::Sum(t, tmp, len, sum, int(n), ctx->stream), "reduce");
    check(ctx, hipMemcpyAsync(&vol, sum, sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "copy");
    check(ctx, hipStreamSynchronize(ctx->stream), "sync");
    size_t tmp2 = 0;
    int64_t* isum = reinterpret_cast<int64_t*>(sum + 1);
    check(ctx, hipcub::DeviceReduce::Sum(nullptr, tmp2, cnt, isum, int(n), ctx->stream), "reduce size");
    t = ws(ctx, kSCAN_TMP, tmp2);
    check(ctx, hipcub::DeviceReduce::Sum(t, tmp2, cnt, isum, int(n), ctx->stream), "reduce");
    segs = read_word(ctx, isum);
  }
  if (volume) *volume = vol * spacing * spacing * spacing;
  if (segments) *segments = segs;
}

template <class T>
void xor_device(vo_context* ctx, const vo_volume* a, const vo_volume* b, double spacing, vo_volume* out,
                int64_t* out_count, double* volume) {
  validate_shape(ctx, a);
  validate_shape(ctx, b);
  if (!out) fail(ctx, VO_ERR_ARGUMENT, "vo_xor_device: null output");
  if (a->nx != b->nx || a->ny != b->ny) fail(ctx, VO_ERR_ARGUMENT, "calculateXor needs equal grids");
  const int64_t n = int64_t(a->nx) * a->ny;
  int64_t* off = static_cast<int64_t*>(ws(ctx, kX_OFF, size_t(n + 1) * sizeof(int64_t)));
  T* z = nullptr;
  int64_t total = 0;
  if (n == 0) {
    check(ctx, hipMemsetAsync(off, 0, sizeof(int64_t), ctx->stream), "memset");
    z = static_cast<T*>(ws(ctx, kX_Z, 16));
  } else {
    // scratch: 2 * (endpoints of a + endpoints of b + 2 per column) doubles
    int64_t ea[2], eb[2];
    check(ctx, hipMemcpyAsync(&ea[0], a->offsets, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream), "copy");
    check(ctx, hipMemcpyAsync(&ea[1], a->offsets + n, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream), "copy");
    check(ctx, hipMemcpyAsync(&eb[0], b->offsets, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream), "copy");
    check(ctx, hipMemcpyAsync(&eb[1], b->offsets + n, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream), "copy");
    check(ctx, hipStreamSynchronize(ctx->stream), "sync");
    const int64_t scr = 2 * ((ea[1] - ea[0]) + (eb[1] - eb[0]) + 2 * n);
    double* scratch = static_cast<double*>(ws(ctx, kX_SCR, size_t(scr) * sizeof(double)));
    int64_t* n1 = static_cast<int64_t*>(ws(ctx, kX_N1, size_t(n) * sizeof(int64_t)));
    int64_t* n2 = static_cast<int64_t*>(ws(ctx, kX_N2, size_t(n) * sizeof(int64_t)));
    int64_t* len = static_cast<int64_t*>(ws(ctx, kX_LEN, size_t(n + 1) * sizeof(int64_t)));
    const double lo = a->z_lo, hi = a->z_hi;
    const dim3 grid(blocks_for(n, 256));
    hipLaunchKernelGGL(vo::xor_kernel<T>, grid, dim3(256), 0, ctx->stream, n, a->offsets, static_cast<const T*>(a->z),
                       b->offsets, static_cast<const T*>(b->z), lo, hi, 0, scratch, n1, n2, len, nullptr, nullptr);
    check(ctx, hipGetLastError(), "xor launch");
    check(ctx, hipMemsetAsync(len + n, 0, sizeof(int64_t), ctx->stream), "memset");
    exclusive_scan(ctx, len, off, n + 1);
    total = read_word(ctx, off + n);
    z = static_cast<T*>(ws(ctx, kX_Z, size_t(std::max<int64_t>(total, 1)) * sizeof(T)));
    hipLaunchKernelGGL(vo::xor_kernel<T>, grid, dim3(256), 0, ctx->stream, n, a->offsets, static_cast<const T*>(a->z),
                       b->offsets, static_cast<const T*>(b->z), lo, hi, 1, scratch, n1, n2, len, off, z);
    check(ctx, hipGetLastError(), "xor launch");
  }
  out->nx = a->nx;
  out->ny = a->ny;
  out->offsets = off;
  out->z = z;
  out->z_lo = a->z_lo;
  out->z_hi = a->z_hi;
  if (out_count) *out_count = total;
  if (volume) measure<T>(ctx, out, spacing, volume, nullptr);
  check(ctx, hipStreamSynchronize(ctx->stream), "sync");
}

// dexel_dump as a mesh, computed on the device from a host volume.
void dexel_mesh(vo_context* ctx, const vo_volume* v, double ox, double oy, double spacing, vo_mesh_result** out) {
  validate_shape(ctx, v);
  if (!out || !(spacing > 0)) fail(ctx, VO_ERR_ARGUMENT, "vo_dexel_mesh: invalid arguments");
  const int64_t n = int64_t(v->nx) * v->ny;
  std::unique_ptr<vo_mesh_result> res(new vo_mesh_result);
  const int64_t cnt = n ? v->offsets[n] - v->offsets[0] : 0;
  if (n == 0 || cnt == 0) {
    *out = res.release();
    return;
  }
  int64_t* doff = static_cast<int64_t*>(ws(ctx, kIN_OFF, size_t(n + 1) * sizeof(int64_t)));
  double* dz = static_cast<double*>(ws(ctx, kIN_Z, size_t(cnt) * sizeof(double)));
  std::vector<int64_t> off0(v->offsets, v->offsets + n + 1);
  for (auto& o : off0) o -= v->offsets[0];
  check(ctx, hipMemcpyAsync(doff, off0.data(), size_t(n + 1) * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream),
        "upload");
  check(ctx, hipMemcpyAsync(dz, static_cast<const double*>(v->z) + v->offsets[0], size_t(cnt) * sizeof(double),
                            hipMemcpyHostToDevice, ctx->stream),
        "upload");
  int64_t* pairs = static_cast<int64_t*>(ws(ctx, kM_PAIRS, size_t(n + 1) * sizeof(int64_t)));
  int64_t* seg_off = static_cast<int64_t*>(ws(ctx, kM_SEGOFF, size_t(n + 1) * sizeof(int64_t)));
  const dim3 grid(blocks_for(n, 256));
  hipLaunchKernelGGL(vo::pair_count_kernel, grid, dim3(256), 0, ctx->stream, n, doff, pairs);
  check(ctx, hipGetLastError(), "pair count launch");
  check(ctx, hipMemsetAsync(pairs + n, 0, sizeof(int64_t), ctx->stream), "memset");
  exclusive_scan(ctx, pairs, seg_off, n + 1);
  const int64_t segs = read_word(ctx, seg_off + n);
  double* verts = static_cast<double*>(ws(ctx, kM_VERT, size_t(std::max<int64_t>(segs, 1)) * 24 * sizeof(double)));
  int64_t* quads = static_cast<int64_t*>(ws(ctx, kM_QUAD, size_t(std::max<int64_t>(segs, 1)) * 24 * sizeof(int64_t)));
  hipLaunchKernelGGL(vo::dexel_mesh_kernel<double>, grid, dim3(256), 0, ctx->stream, v->nx, v->ny, doff,
                     static_cast<const double*>(dz), ox, oy, 0.0, spacing, seg_off, verts, quads);
  check(ctx, hipGetLastError(), "dexel mesh launch");
  res->vertices.resize(size_t(segs) * 24);
  res->quads.resize(size_t(segs) * 24);
  if (segs) {
    check(ctx, hipMemcpyAsync(res->vertices.data(), verts, res->vertices.size() * sizeof(double),
                              hipMemcpyDeviceToHost, ctx->stream),
          "download");
    check(ctx, hipMemcpyAsync(res->quads.data(), quads, res->quads.size() * sizeof(int64_t), hipMemcpyDeviceToHost,
                              ctx->stream),
          "download");
  }
  check(ctx, hipStreamSynchronize(ctx->stream), "sync");
  *out = res.release();
}

// ---------------------------------------------------------------------------
// Autotuner: per radius, the z-sweep event capacity (pass-1 time) and the
// gather list capacity (pass-2 time, fallbacks included) measured on a sample
// volume; the fastest of each is kept for the radius' bucket.

template <class T>
void autotune(vo_context* ctx, const vo_volume* sample, int n, const double* radii, int reps) {
  validate_shape(ctx, sample);
  if (n < 0 || (n > 0 && !radii) || reps < 1 || reps > 100) fail(ctx, VO_ERR_ARGUMENT, "vo_autotune: bad arguments");
  DevVol in;
  in.nx = sample->nx;
  in.ny = sample->ny;
  in.off = sample->offsets;
  in.z = sample->z;
  const int forced = ctx->zsweep_events;
  for (int i = 0; i < n; ++i) {
    const double R = radii[i];
    if (!(R >= 0) || !std::isfinite(R) || R > 127) fail(ctx, VO_ERR_ARGUMENT, "vo_autotune: radius out of [0, 127]");
    const int b = tune_bucket(int(std::floor(R)));
    auto time_of = [&](int which) {
      double best = INFINITY;
      for (int r = 0; r < reps; ++r) {
        double t1 = 0, t2 = 0;
        dilate<T>(ctx, in, R, kV3_OFF, kV3_Z, &t1, &t2);
        best = std::min(best, which == 1 ? t1 : t2);
      }
      return best;
    };
    // a candidate replaces the current choice only when it is more than 3 %
    // faster, so timing noise on a small sample does not move the defaults
    const int cur_e = ctx->tuning.events[b];
    ctx->zsweep_events = cur_e;
    int best_e = cur_e;
    double tb = time_of(1);
    for (int e : kEventChoices) {
      if (e == cur_e) continue;
      ctx->zsweep_events = e;
      const double t = time_of(1);
      if (t < 0.97 * tb) {
        tb = t;
        best_e = e;
      }
    }
    ctx->zsweep_events = forced;
    ctx->tuning.events[b] = best_e;
    const int cur_o = ctx->tuning.ocap[b];
    int best_o = cur_o;
    tb = time_of(2);
    for (int o : {8, 16, 24, 32}) {
      if (o == cur_o) continue;
      ctx->tuning.ocap[b] = o;
      const double t = time_of(2);
      if (t < 0.97 * tb) {
        tb = t;
        best_o = o;
      }
    }
    ctx->tuning.ocap[b] = best_o;
  }
}

int guarded(vo_context* ctx, const std::function<void()>& body) {
  if (!ctx) return VO_ERR_ARGUMENT;
  // a call's input check never outlives it
  ctx->check_input = false;
  ctx->fused.word = nullptr;
  (void)hipGetLastError();  // an error left by an earlier, already reported call is not this call's
  try {
    check(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    body();
    ctx->check_input = false;
    ctx->fused.word = nullptr;
    return VO_OK;
  } catch (const Failure& f) {
    ctx->check_input = false;
    ctx->fused.word = nullptr;
    return f.code;
  } catch (const std::exception& e) {
    ctx->check_input = false;
    ctx->fused.word = nullptr;
    ctx->error = e.what();
    return VO_ERR_ARGUMENT;
  }
}

void validate_shape(vo_context* ctx, const vo_volume* v) {
  if (!v || v->nx < 0 || v->ny < 0 || (!v->offsets && int64_t(v->nx) * v->ny > 0))
    fail(ctx, VO_ERR_ARGUMENT, "invalid volume");
}

// Host-side checks of the CompressedVolume invariants the sweeps rely on.
// f(begin, end) over [0, n) on up to 16 host threads (the calling thread
// takes the first chunk); small ranges run on the calling thread alone.
template <class F>
void host_parallel(int64_t n, F&& f) {
  const int64_t kMinChunk = int64_t(1) << 16;
  const int hw = int(std::max(1u, std::thread::hardware_concurrency()));
  const int nt = int(std::min<int64_t>(std::min(hw, 16), std::max<int64_t>(1, n / kMinChunk)));
  if (nt <= 1) {
    f(int64_t(0), n);
    return;
  }
  std::vector<std::thread> pool;
  const int64_t step = (n + nt - 1) / nt;
  for (int t = 1; t < nt; ++t) {
    const int64_t b = std::min(n, t * step), e = std::min(n, b + step);
    pool.emplace_back([&f, b, e] { f(b, e); });
  }
  f(int64_t(0), std::min(n, step));
  for (auto& th : pool) th.join();
}

// The host entry point's input checks.  Offsets first (column order, so the
// endpoints are read only within them), then the endpoints of every column
// in parallel; the lowest failing column is reported, with the first of its
// failures in endpoint order.
void validate_host(vo_context* ctx, const vo_volume* v, int op, double radius) {
  const int64_t n = int64_t(v->nx) * v->ny;
  const double* z = static_cast<const double*>(v->z);
  if (v->offsets[0] != 0) fail(ctx, VO_ERR_ARGUMENT, "offsets[0] must be 0");
  for (int64_t c = 0; c < n; ++c) {
    const int64_t s = v->offsets[c], e = v->offsets[c + 1];
    if (e < s || (e - s) % 2) fail(ctx, VO_ERR_ARGUMENT, "column " + std::to_string(c) + " has odd length");
  }
  std::mutex mu;
  int64_t bad_col = n;
  int bad_code = VO_OK;
  std::string bad_text;
  host_parallel(n, [&](int64_t c0, int64_t c1) {
    for (int64_t c = c0; c < c1; ++c) {
      const int64_t s = v->offsets[c], e = v->offsets[c + 1];
      int code = VO_OK;
      std::string text;
      for (int64_t i = s; i < e && code == VO_OK; ++i) {
        if (!std::isfinite(z[i])) {
          code = VO_ERR_ARGUMENT;
          text = "non-finite endpoint";
        } else if (i > s && z[i] < z[i - 1]) {
          code = VO_ERR_ARGUMENT;
          text = "column " + std::to_string(c) + " is unsorted";
        } else if ((op == VO_OP_EROSION || op == VO_OP_OPENING) && (z[i] < v->z_lo - 1 || z[i] > v->z_hi + 1)) {
          code = VO_ERR_RANGE;
          text = "erosion input endpoint outside [z_lo - 1, z_hi + 1]";
        } else if (op == VO_OP_CLOSING && (z[i] - radius < v->z_lo - 1 || z[i] + radius > v->z_hi + 1)) {
          code = VO_ERR_RANGE;
          text = "closing: the dilated volume would leave [z_lo - 1, z_hi + 1]";
        }
      }
      if (code != VO_OK) {
        std::lock_guard<std::mutex> lock(mu);
        if (c < bad_col) {
          bad_col = c;
          bad_code = code;
          bad_text = text;
        }
        return;  // the rest of this chunk is above c
      }
    }
  });
  if (bad_code != VO_OK) fail(ctx, bad_code, bad_text);
}

// Mesh -> dexels (create_dexels' ray casting, Dexelize.cpp:166-225); see
// vo_raster.hpp for the steps.
void rasterize(vo_context* ctx, const double* V, int64_t nv, const int32_t* F, int64_t nf, int nx, int ny,
               double ox, double oy, double spacing, vo_result** out) {
  if (!out || nv < 0 || nf < 0 || nx < 0 || ny < 0 || (nv > 0 && !V) || (nf > 0 && !F))
    fail(ctx, VO_ERR_ARGUMENT, "vo_rasterize: invalid arguments");
  if (!(spacing > 0) || !std::isfinite(spacing) || !std::isfinite(ox) || !std::isfinite(oy))
    fail(ctx, VO_ERR_ARGUMENT, "vo_rasterize: spacing must be finite and > 0, origin finite");
  for (int64_t i = 0; i < 3 * nf; ++i)
    if (F[i] < 0 || F[i] >= nv) fail(ctx, VO_ERR_ARGUMENT, "vo_rasterize: triangle index out of range");
  for (int64_t i = 0; i < 3 * nv; ++i)
    if (!std::isfinite(V[i])) fail(ctx, VO_ERR_ARGUMENT, "vo_rasterize: non-finite vertex");
  const int64_t ncols = int64_t(nx) * ny;
  std::unique_ptr<vo_result> res(new vo_result);
  res->nx = nx;
  res->ny = ny;
  res->off.assign(size_t(ncols) + 1, 0);
  if (ncols == 0 || nf == 0) {
    *out = res.release();
    return;
  }
  vo::RasterGrid g{nx, ny, ox, oy, spacing, (nx + vo::kRasterTile - 1) / vo::kRasterTile,
                   (ny + vo::kRasterTile - 1) / vo::kRasterTile};
  const int64_t ntiles = int64_t(g.tx) * g.ty;
  if (ntiles > INT32_MAX || nf > int64_t(INT32_MAX) * 256) fail(ctx, VO_ERR_ARGUMENT, "vo_rasterize: too large");
  double* dV = static_cast<double*>(ws(ctx, kR_V, size_t(nv) * 3 * sizeof(double)));
  int32_t* dF = static_cast<int32_t*>(ws(ctx, kR_F, size_t(nf) * 3 * sizeof(int32_t)));
  int64_t* tcnt = static_cast<int64_t*>(ws(ctx, kR_TCNT, size_t(ntiles + 1) * sizeof(int64_t)));
  int64_t* toff = static_cast<int64_t*>(ws(ctx, kR_TOFF, size_t(ntiles + 1) * sizeof(int64_t)));
  int64_t* tcur = static_cast<int64_t*>(ws(ctx, kR_TCUR, size_t(ntiles + 1) * sizeof(int64_t)));
  int64_t* len = static_cast<int64_t*>(ws(ctx, kR_LEN, size_t(ncols + 1) * sizeof(int64_t)));
  int64_t* off = static_cast<int64_t*>(ws(ctx, kR_OFF, size_t(ncols + 1) * sizeof(int64_t)));
  check(ctx, hipMemcpyAsync(dV, V, size_t(nv) * 3 * sizeof(double), hipMemcpyHostToDevice, ctx->stream), "upload");
  check(ctx, hipMemcpyAsync(dF, F, size_t(nf) * 3 * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream), "upload");
  check(ctx, hipMemsetAsync(tcnt, 0, size_t(ntiles + 1) * sizeof(int64_t), ctx->stream), "memset");
  const dim3 tgrid(blocks_for(nf, 256));
  hipLaunchKernelGGL(vo::raster_bin_kernel, tgrid, dim3(256), 0, ctx->stream, dV, dF, nf, g, 0, tcnt, nullptr,
                     nullptr);
  check(ctx, hipGetLastError(), "raster bin launch");
  exclusive_scan(ctx, tcnt, toff, ntiles + 1);
  const int64_t pairs = read_word(ctx, toff + ntiles);
  int64_t* ttris = static_cast<int64_t*>(ws(ctx, kR_TTRIS, size_t(std::max<int64_t>(pairs, 1)) * sizeof(int64_t)));
  check(ctx, hipMemcpyAsync(tcur, toff, size_t(ntiles + 1) * sizeof(int64_t), hipMemcpyDeviceToDevice, ctx->stream),
        "copy");
  hipLaunchKernelGGL(vo::raster_bin_kernel, tgrid, dim3(256), 0, ctx->stream, dV, dF, nf, g, 1, nullptr, tcur,
                     ttris);
  check(ctx, hipGetLastError(), "raster bin launch");
  hipLaunchKernelGGL(vo::raster_cast_kernel, dim3(unsigned(ntiles)), dim3(vo::kWave), 0, ctx->stream, dV, dF, g,
                     toff, ttris, 0, len, nullptr, nullptr);
  check(ctx, hipGetLastError(), "raster cast launch");
  check(ctx, hipMemsetAsync(len + ncols, 0, sizeof(int64_t), ctx->stream), "memset");
  exclusive_scan(ctx, len, off, ncols + 1);
  const int64_t hits = read_word(ctx, off + ncols);
  double* dz = static_cast<double*>(ws(ctx, kR_Z, size_t(std::max<int64_t>(hits, 1)) * sizeof(double)));
  hipLaunchKernelGGL(vo::raster_cast_kernel, dim3(unsigned(ntiles)), dim3(vo::kWave), 0, ctx->stream, dV, dF, g,
                     toff, ttris, 1, nullptr, off, dz);
  check(ctx, hipGetLastError(), "raster cast launch");
  res->z.resize(size_t(hits));
  check(ctx, hipMemcpyAsync(res->off.data(), off, size_t(ncols + 1) * sizeof(int64_t), hipMemcpyDeviceToHost,
                            ctx->stream),
        "download");
  if (hits)
    check(ctx, hipMemcpyAsync(res->z.data(), dz, size_t(hits) * sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
          "download");
  check(ctx, hipStreamSynchronize(ctx->stream), "sync");
  *out = res.release();
}

// FP32 travels relative to an integer shift of the z-range centre.
template <class T>
double storage_shift(const vo_volume* in) {
  return std::is_same<T, float>::value ? std::round(0.5 * (in->z_lo + in->z_hi)) : 0.0;
}

// The context's pinned staging buffer, at least `bytes` long (grown with
// the stream drained: no copy from the old buffer is in flight).
void* staging(vo_context* ctx, size_t bytes) {
  if (bytes > ctx->staging_bytes) {
    const size_t want = std::max(bytes, ctx->staging_bytes + ctx->staging_bytes / 2);
    check(ctx, hipStreamSynchronize(ctx->stream), "sync");
    if (ctx->staging) check(ctx, hipHostFree(ctx->staging), "free staging");
    ctx->staging = nullptr;
    ctx->staging_bytes = 0;
    check(ctx, hipHostMalloc(&ctx->staging, want, hipHostMallocDefault), "pinned staging");
    ctx->staging_bytes = want;
  }
  return ctx->staging;
}

// A checked host volume (double z) uploaded in storage precision T, minus
// `shift`, through the staging buffer (which the call's later
// synchronisation frees for reuse).
template <class T>
DevVol upload_host(vo_context* ctx, const vo_volume* in, double shift) {
  const int64_t n = int64_t(in->nx) * in->ny;
  const int64_t cnt = n ? in->offsets[n] : 0;
  const size_t zn = size_t(std::max<int64_t>(cnt, 1));
  // offsets and converted endpoints go through the pinned staging buffer
  char* st = static_cast<char*>(staging(ctx, (n + 1) * sizeof(int64_t) + zn * sizeof(T)));
  int64_t* so = reinterpret_cast<int64_t*>(st);
  T* sz = reinterpret_cast<T*>(st + (n + 1) * sizeof(int64_t));
  const double* z = static_cast<const double*>(in->z);
  host_parallel(n + 1, [&](int64_t b, int64_t e) {
    std::memcpy(so + b, in->offsets + b, size_t(e - b) * sizeof(int64_t));
  });
  host_parallel(cnt, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) sz[i] = T(z[i] - shift);
  });
  if (cnt == 0) sz[0] = T(0);
  int64_t* doff = static_cast<int64_t*>(ws(ctx, kIN_OFF, (n + 1) * sizeof(int64_t)));
  T* dz = static_cast<T*>(ws(ctx, kIN_Z, zn * sizeof(T)));
  check(ctx, hipMemcpyAsync(doff, so, (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream), "upload");
  check(ctx, hipMemcpyAsync(dz, sz, zn * sizeof(T), hipMemcpyHostToDevice, ctx->stream), "upload");
  DevVol v;
  v.nx = in->nx;
  v.ny = in->ny;
  v.off = doff;
  v.z = dz;
  v.count = cnt;
  return v;
}

template <class T>
void morph_host(vo_context* ctx, int op, double R, const vo_volume* in, vo_result** out, double* t1,
                double* t2) {
  validate_shape(ctx, in);
  validate_host(ctx, in, op, R);
  const double shift = storage_shift<T>(in);
  DevVol v = upload_host<T>(ctx, in, shift);
  DevVol r = v;
  if (op != VO_OP_NOOP) r = run_op<T>(ctx, op, v, R, in->z_lo - shift, in->z_hi - shift, t1, t2);
  auto* res = new vo_result;
  res->nx = r.nx;
  res->ny = r.ny;
  const int64_t rn = int64_t(r.nx) * r.ny;
  res->off.resize(rn + 1);
  // through the staging buffer: the upload from it ran before these copies
  const size_t zn = size_t(std::max<int64_t>(r.count, 1));
  char* st = static_cast<char*>(staging(ctx, (rn + 1) * sizeof(int64_t) + zn * sizeof(T)));
  int64_t* so = reinterpret_cast<int64_t*>(st);
  T* sz = reinterpret_cast<T*>(st + (rn + 1) * sizeof(int64_t));
  check(ctx, hipMemcpyAsync(so, r.off, (rn + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream), "download");
  check(ctx, hipMemcpyAsync(sz, r.z, zn * sizeof(T), hipMemcpyDeviceToHost, ctx->stream), "download");
  check(ctx, hipStreamSynchronize(ctx->stream), "sync");
  res->z.resize(r.count);
  host_parallel(rn + 1, [&](int64_t b, int64_t e) {
    std::memcpy(res->off.data() + b, so + b, size_t(e - b) * sizeof(int64_t));
  });
  host_parallel(r.count, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) res->z[i] = double(sz[i]) + shift;
  });
  *out = res;
}

// The pass-1 intermediate of a dilation of a host volume, as CSR triples.
// The dilation runs whole; its pass-1 buffers stay in the workspace and are
// exported from there.
template <class T>
void pass1_host(vo_context* ctx, double R, const vo_volume* in, vo_triples* res) {
  validate_shape(ctx, in);
  validate_host(ctx, in, VO_OP_DILATION, R);
  const double shift = storage_shift<T>(in);
  DevVol v = upload_host<T>(ctx, in, shift);
  // the reference's axis: lines of constant x, swept along y
  const int forced = ctx->axis_forced;
  ctx->axis_forced = 0;
  try {
    dilate<T>(ctx, v, R, kV3_OFF, kV3_Z, nullptr, nullptr);
  } catch (...) {
    ctx->axis_forced = forced;
    throw;
  }
  ctx->axis_forced = forced;
  const int64_t n = int64_t(in->nx) * in->ny;
  res->nx = in->nx;
  res->ny = in->ny;
  res->off.assign(size_t(n) + 1, 0);
  if (n == 0) return;
  int64_t* len = static_cast<int64_t*>(ws(ctx, kX_LEN, size_t(n + 1) * sizeof(int64_t)));
  int64_t* off = static_cast<int64_t*>(ws(ctx, kX_OFF, size_t(n + 1) * sizeof(int64_t)));
  vo::TripleExport<T> g{in->nx, in->ny, ctx->p1_caps, static_cast<const T*>(slot_ptr(ctx, kM_LO)),
                        static_cast<const T*>(slot_ptr(ctx, kM_HI)), static_cast<const uint16_t*>(slot_ptr(ctx, kM_DY)),
                        static_cast<const int32_t*>(slot_ptr(ctx, kM_CNT)), R * R, shift, len, off, nullptr, nullptr,
                        nullptr};
  const dim3 grid(blocks_for(n, 256));
  hipLaunchKernelGGL(vo::export_triples_kernel<T>, grid, dim3(256), 0, ctx->stream, g, 0);
  check(ctx, hipGetLastError(), "export launch");
  check(ctx, hipMemsetAsync(len + n, 0, sizeof(int64_t), ctx->stream), "memset");
  exclusive_scan(ctx, len, off, n + 1);
  const int64_t total = read_word(ctx, off + n);
  const int64_t stride = std::max<int64_t>(total, 1);
  double* y = static_cast<double*>(ws(ctx, kX_SCR, size_t(stride) * 3 * sizeof(double)));
  g.y1 = y;
  g.y2 = y + stride;
  g.r = y + 2 * stride;
  hipLaunchKernelGGL(vo::export_triples_kernel<T>, grid, dim3(256), 0, ctx->stream, g, 1);
  check(ctx, hipGetLastError(), "export launch");
  res->y1.resize(size_t(total));
  res->y2.resize(size_t(total));
  res->r.resize(size_t(total));
  check(ctx, hipMemcpyAsync(res->off.data(), off, size_t(n + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream),
        "download");
  if (total) {
    const size_t bytes = size_t(total) * sizeof(double);
    check(ctx, hipMemcpyAsync(res->y1.data(), g.y1, bytes, hipMemcpyDeviceToHost, ctx->stream), "download");
    check(ctx, hipMemcpyAsync(res->y2.data(), g.y2, bytes, hipMemcpyDeviceToHost, ctx->stream), "download");
    check(ctx, hipMemcpyAsync(res->r.data(), g.r, bytes, hipMemcpyDeviceToHost, ctx->stream), "download");
  }
  sync(ctx);
}

// vo_morph_device with every buffer in this device's scratch and the result
// in the caller's buffers: the operation runs as in vo_morph_device (same
// kernels, same results), its workspace handed out by ScratchArena.
template <class T>
void morph_noalloc(vo_context* ctx, int op, double R, const vo_volume* in, int64_t* out_offsets, void* out_z,
                   int64_t out_capacity, int64_t* out_count, double* t1, double* t2) {
  validate_shape(ctx, in);
  if (op < VO_OP_DILATION || op > VO_OP_OPENING) fail(ctx, VO_ERR_ARGUMENT, "unknown operation");
  if (!out_offsets || !out_z || !out_count || out_capacity < 1)
    fail(ctx, VO_ERR_ARGUMENT, "no-alloc: null output or capacity below 1");
  if (!ctx->scratch) fail(ctx, VO_ERR_DEVICE, "no-alloc: no scratch on this device");
  const int64_t n = int64_t(in->nx) * in->ny;
  ScratchArena& a = ctx->arena;
  a = ScratchArena();
  a.on = true;
  a.base = static_cast<char*>(ctx->scratch);
  a.cap = kScratchBytes;
  a.slot[kV3_OFF] = Buffer{out_offsets, size_t(n + 1) * sizeof(int64_t)};
  a.fixed[kV3_OFF] = true;
  a.slot[kV3_Z] = Buffer{out_z, size_t(out_capacity) * sizeof(T)};
  a.fixed[kV3_Z] = true;
  DevVol v;
  v.nx = in->nx;
  v.ny = in->ny;
  v.off = in->offsets;
  v.z = in->z;
  check_input_of_op<T>(ctx, op, R, *in);
  DevVol r = run_op<T>(ctx, op, v, R, in->z_lo, in->z_hi, t1, t2);
  if (r.off != out_offsets || (r.count > 0 && r.z != out_z))
    fail(ctx, VO_ERR_DEVICE, "no-alloc: the result is not in the caller's buffers");
  sync(ctx);
  *out_count = r.count;
}

// Why this thread's last vo_create failed (vo_last_error(NULL)).
thread_local std::string g_create_error;

}  // namespace

extern "C" {

int vo_create(vo_context** out, int device) {
  if (!out) return VO_ERR_ARGUMENT;
  *out = nullptr;
  auto* ctx = new vo_context;
  ctx->device = device;
  if (const char* e = std::getenv("VO_PASS2")) ctx->force_sweep = std::string(e) == "sweep";
  if (const char* e = std::getenv("VO_ZSWEEP_EVENTS")) ctx->zsweep_events = std::atoi(e);
  if (const char* e = std::getenv("VO_INJECT_HIP_ERROR")) ctx->inject = e;
  if (const char* e = std::getenv("VO_BATCH_GRAPH")) ctx->use_graphs = std::string(e) != "0";
  if (const char* e = std::getenv("VO_BATCH_QUEUE")) ctx->batch_queue = std::string(e) == "1";
  if (const char* e = std::getenv("VO_AXIS")) ctx->axis_forced = std::string(e) == "x" ? 1 : (std::string(e) == "y" ? 0 : -1);
  if (const char* e = std::getenv("VO_PASS1"))
    ctx->pass1_form = std::string(e) == "cull" ? 2 : (std::string(e) == "paint" ? 1 : 0);
  int rc = guarded(ctx, [&] {
    int n = 0;
    check(ctx, hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (device < 0 || device >= n) fail(ctx, VO_ERR_ARGUMENT, "no such HIP device");
    check(ctx, hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "stream");
    for (auto& e : ctx->ev) check(ctx, hipEventCreate(&e), "event");
    check(ctx, hipDeviceGetAttribute(&ctx->num_cus, hipDeviceAttributeMultiprocessorCount, device), "CU count");
    // coherent: kernels write the control words into it directly
    check(ctx, hipHostMalloc(reinterpret_cast<void**>(&ctx->host_words), 8 * sizeof(int64_t), hipHostMallocCoherent),
          "pinned");
    if (device < kMaxDevices) {  // this device's copy of the no-alloc scratch (device current: guarded)
      std::lock_guard<std::mutex> lock(g_scratch_mu);
      if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        check(ctx, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[device] = p;
      }
      ctx->scratch = g_scratch_ptr[device];
    }
  });
  if (rc != VO_OK) {
    g_create_error = ctx->error;
    vo_destroy(ctx);
    (void)hipGetLastError();  // reported here, not by the next launch check of another context
    return rc;
  }
  g_create_error.clear();
  *out = ctx;
  return VO_OK;
}

void vo_destroy(vo_context* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (void* p : ctx->batch_extra) (void)hipFree(p);
  for (auto& b : ctx->buf)
    if (b.ptr) (void)hipFree(b.ptr);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->host_words) (void)hipHostFree(ctx->host_words);
  if (ctx->batch_host) (void)hipHostFree(ctx->batch_host);
  if (ctx->staging) (void)hipHostFree(ctx->staging);
  if (ctx->bgraph.exec) (void)hipGraphExecDestroy(ctx->bgraph.exec);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* vo_last_error(const vo_context* ctx) {
  if (ctx) return ctx->error.c_str();
  return g_create_error.empty() ? "no context" : g_create_error.c_str();
}

int vo_morph(vo_context* ctx, int op, int precision, double radius, const vo_volume* in, vo_result** out,
             double* time_1_ms, double* time_2_ms) {
  return guarded(ctx, [&] {
    if (!out) fail(ctx, VO_ERR_ARGUMENT, "null result pointer");
    if (op < VO_OP_NOOP || op > VO_OP_OPENING) fail(ctx, VO_ERR_ARGUMENT, "unknown operation");
    if (precision == VO_FP32)
      morph_host<float>(ctx, op, radius, in, out, time_1_ms, time_2_ms);
    else if (precision == VO_FP64)
      morph_host<double>(ctx, op, radius, in, out, time_1_ms, time_2_ms);
    else
      fail(ctx, VO_ERR_ARGUMENT, "unknown precision");
  });
}

int vo_result_shape(const vo_result* r, int32_t* nx, int32_t* ny, int64_t* count) {
  if (!r) return VO_ERR_ARGUMENT;
  if (nx) *nx = r->nx;
  if (ny) *ny = r->ny;
  if (count) *count = int64_t(r->z.size());
  return VO_OK;
}

int vo_result_copy(const vo_result* r, int64_t* offsets, double* z) {
  if (!r) return VO_ERR_ARGUMENT;
  if (offsets) std::memcpy(offsets, r->off.data(), r->off.size() * sizeof(int64_t));
  if (z && !r->z.empty()) std::memcpy(z, r->z.data(), r->z.size() * sizeof(double));
  return VO_OK;
}

void vo_result_free(vo_result* r) { delete r; }

int vo_pass1(vo_context* ctx, int precision, double radius, const vo_volume* in, vo_triples** out) {
  if (!out) return VO_ERR_ARGUMENT;
  *out = nullptr;
  std::unique_ptr<vo_triples> t(new vo_triples);
  const int rc = guarded(ctx, [&] {
    if (precision == VO_FP32)
      pass1_host<float>(ctx, radius, in, t.get());
    else if (precision == VO_FP64)
      pass1_host<double>(ctx, radius, in, t.get());
    else
      fail(ctx, VO_ERR_ARGUMENT, "unknown precision");
  });
  if (rc == VO_OK) *out = t.release();
  return rc;
}

int vo_triples_shape(const vo_triples* t, int32_t* nx, int32_t* ny, int64_t* count) {
  if (!t) return VO_ERR_ARGUMENT;
  if (nx) *nx = t->nx;
  if (ny) *ny = t->ny;
  if (count) *count = int64_t(t->y1.size());
  return VO_OK;
}

int vo_triples_copy(const vo_triples* t, int64_t* offsets, double* y1, double* y2, double* r) {
  if (!t) return VO_ERR_ARGUMENT;
  if (offsets) std::memcpy(offsets, t->off.data(), t->off.size() * sizeof(int64_t));
  const size_t n = t->y1.size() * sizeof(double);
  if (n && y1) std::memcpy(y1, t->y1.data(), n);
  if (n && y2) std::memcpy(y2, t->y2.data(), n);
  if (n && r) std::memcpy(r, t->r.data(), n);
  return VO_OK;
}

void vo_triples_free(vo_triples* t) { delete t; }

int vo_morph_device_noalloc(vo_context* ctx, int op, int precision, double radius, const vo_volume* in,
                            int64_t* out_offsets, void* out_z, int64_t out_capacity, int64_t* out_count,
                            double* time_1_ms, double* time_2_ms) {
  const int rc = guarded(ctx, [&] {
    if (precision == VO_FP32)
      morph_noalloc<float>(ctx, op, radius, in, out_offsets, out_z, out_capacity, out_count, time_1_ms, time_2_ms);
    else if (precision == VO_FP64)
      morph_noalloc<double>(ctx, op, radius, in, out_offsets, out_z, out_capacity, out_count, time_1_ms, time_2_ms);
    else
      fail(ctx, VO_ERR_ARGUMENT, "unknown precision");
  });
  if (ctx) {
    // an output buffer too small: *out_count is the capacity needed
    if (rc == VO_ERR_CAPACITY && ctx->arena.short_by && out_count)
      *out_count = int64_t(ctx->arena.short_by / (precision == VO_FP32 ? sizeof(float) : sizeof(double)));
    ctx->arena.on = false;
  }
  return rc;
}

int vo_copy_device(vo_context* ctx, int n, void* const* dst, const void* const* src, const int64_t* bytes) {
  return guarded(ctx, [&] {
    if (n < 0 || (n > 0 && (!dst || !src || !bytes))) fail(ctx, VO_ERR_ARGUMENT, "vo_copy_device: bad arguments");
    for (int i = 0; i < n; ++i) {
      if (bytes[i] < 0) fail(ctx, VO_ERR_ARGUMENT, "vo_copy_device: negative size");
      if (bytes[i] > 0)
        check(ctx, hipMemcpyAsync(dst[i], src[i], size_t(bytes[i]), hipMemcpyDeviceToDevice, ctx->stream), "copy");
    }
    sync(ctx);
  });
}


int vo_rasterize(vo_context* ctx, const double* vertices, int64_t num_vertices, const int32_t* triangles,
                 int64_t num_triangles, int32_t nx, int32_t ny, double origin_x, double origin_y, double spacing,
                 vo_result** out) {
  vo_result* r = nullptr;
  const int rc = guarded(ctx, [&] {
    rasterize(ctx, vertices, num_vertices, triangles, num_triangles, nx, ny, origin_x, origin_y, spacing, &r);
  });
  if (rc != VO_OK) {
    delete r;
    return rc;
  }
  *out = r;
  return VO_OK;
}

int vo_morph_device(vo_context* ctx, int op, int precision, double radius, const vo_volume* in, vo_volume* out,
                    int64_t* out_count, double* time_1_ms, double* time_2_ms) {
  return guarded(ctx, [&] {
    validate_shape(ctx, in);
    if (!out || !out_count) fail(ctx, VO_ERR_ARGUMENT, "null output");
    DevVol v;
    v.nx = in->nx;
    v.ny = in->ny;
    v.off = in->offsets;
    v.z = in->z;
    DevVol r = v;
    if (op == VO_OP_NOOP) {
      const int64_t n = int64_t(in->nx) * in->ny;
      r.count = n ? read_word(ctx, in->offsets + n) : 0;
    } else if (precision == VO_FP32) {
      if (op < VO_OP_NOOP || op > VO_OP_OPENING) fail(ctx, VO_ERR_ARGUMENT, "unknown operation");
      check_input_of_op<float>(ctx, op, radius, *in);
      r = run_op<float>(ctx, op, v, radius, in->z_lo, in->z_hi, time_1_ms, time_2_ms);
    } else if (precision == VO_FP64) {
      if (op < VO_OP_NOOP || op > VO_OP_OPENING) fail(ctx, VO_ERR_ARGUMENT, "unknown operation");
      check_input_of_op<double>(ctx, op, radius, *in);
      r = run_op<double>(ctx, op, v, radius, in->z_lo, in->z_hi, time_1_ms, time_2_ms);
    } else {
      fail(ctx, VO_ERR_ARGUMENT, "unknown precision");
    }
    sync(ctx);
    out->nx = r.nx;
    out->ny = r.ny;
    out->offsets = r.off;
    out->z = r.z;
    out->z_lo = in->z_lo;
    out->z_hi = in->z_hi;
    *out_count = r.count;
  });
}

int vo_morph_batch_device(vo_context* ctx, int precision, int n, const vo_job* jobs, vo_volume* out,
                          int64_t* out_count, double* time_1_ms, double* time_2_ms) {
  return guarded(ctx, [&] {
    if (precision == VO_FP32)
      morph_batch_device<float>(ctx, n, jobs, out, out_count, time_1_ms, time_2_ms);
    else if (precision == VO_FP64)
      morph_batch_device<double>(ctx, n, jobs, out, out_count, time_1_ms, time_2_ms);
    else
      fail(ctx, VO_ERR_ARGUMENT, "unknown precision");
  });
}

int vo_measure_device(vo_context* ctx, int precision, const vo_volume* v, double spacing, double* volume,
                      int64_t* segments) {
  return guarded(ctx, [&] {
    if (precision == VO_FP32)
      measure<float>(ctx, v, spacing, volume, segments);
    else if (precision == VO_FP64)
      measure<double>(ctx, v, spacing, volume, segments);
    else
      fail(ctx, VO_ERR_ARGUMENT, "unknown precision");
  });
}

int vo_xor_device(vo_context* ctx, int precision, const vo_volume* a, const vo_volume* b, double spacing,
                  vo_volume* out, int64_t* out_count, double* volume) {
  return guarded(ctx, [&] {
    if (precision == VO_FP32)
      xor_device<float>(ctx, a, b, spacing, out, out_count, volume);
    else if (precision == VO_FP64)
      xor_device<double>(ctx, a, b, spacing, out, out_count, volume);
    else
      fail(ctx, VO_ERR_ARGUMENT, "unknown precision");
  });
}

int vo_dexel_mesh(vo_context* ctx, const vo_volume* v, double origin_x, double origin_y, double spacing,
                  vo_mesh_result** out) {
  vo_mesh_result* r = nullptr;
  const int rc = guarded(ctx, [&] { dexel_mesh(ctx, v, origin_x, origin_y, spacing, &r); });
  if (rc != VO_OK) {
    delete r;
    return rc;
  }
  *out = r;
  return VO_OK;
}

int vo_mesh_shape(const vo_mesh_result* r, int64_t* num_vertices, int64_t* num_quads) {
  if (!r) return VO_ERR_ARGUMENT;
  if (num_vertices) *num_vertices = int64_t(r->vertices.size() / 3);
  if (num_quads) *num_quads = int64_t(r->quads.size() / 4);
  return VO_OK;
}

int vo_mesh_copy(const vo_mesh_result* r, double* vertices, int64_t* quads) {
  if (!r) return VO_ERR_ARGUMENT;
  if (vertices && !r->vertices.empty()) std::memcpy(vertices, r->vertices.data(), r->vertices.size() * sizeof(double));
  if (quads && !r->quads.empty()) std::memcpy(quads, r->quads.data(), r->quads.size() * sizeof(int64_t));
  return VO_OK;
}

void vo_mesh_free(vo_mesh_result* r) { delete r; }

int vo_autotune(vo_context* ctx, int precision, const vo_volume* sample, int num_radii, const double* radii, int reps) {
  return guarded(ctx, [&] {
    if (precision == VO_FP32)
      autotune<float>(ctx, sample, num_radii, radii, reps);
    else if (precision == VO_FP64)
      autotune<double>(ctx, sample, num_radii, radii, reps);
    else
      fail(ctx, VO_ERR_ARGUMENT, "unknown precision");
  });
}

int vo_tuning_get(const vo_context* ctx, int bucket, int* events, int* ocap) {
  if (!ctx || bucket < 0 || bucket >= kTuneBuckets) return VO_ERR_ARGUMENT;
  if (events) *events = ctx->tuning.events[bucket];
  if (ocap) *ocap = ctx->tuning.ocap[bucket];
  return VO_OK;
}

int vo_tuning_save(const vo_context* ctx, const char* path) {
  if (!ctx || !path) return VO_ERR_ARGUMENT;
  FILE* f = std::fopen(path, "w");
  if (!f) return VO_ERR_ARGUMENT;
  std::fprintf(f, "voroffset_amd tuning 1\n");
  for (int b = 0; b < kTuneBuckets; ++b) std::fprintf(f, "%d %d %d\n", b, ctx->tuning.events[b], ctx->tuning.ocap[b]);
  return std::fclose(f) == 0 ? VO_OK : VO_ERR_ARGUMENT;
}

int vo_tuning_load(vo_context* ctx, const char* path) {
  if (!ctx || !path) return VO_ERR_ARGUMENT;
  FILE* f = std::fopen(path, "r");
  if (!f) {
    ctx->error = std::string("cannot open ") + path;
    return VO_ERR_ARGUMENT;
  }
  Tuning t = ctx->tuning;
  int version = 0, ok = std::fscanf(f, "voroffset_amd tuning %d", &version) == 1 && version == 1;
  for (int i = 0; ok && i < kTuneBuckets; ++i) {
    int b, e, o;
    ok = std::fscanf(f, "%d %d %d", &b, &e, &o) == 3 && b >= 0 && b < kTuneBuckets &&
         std::count(std::begin(kEventChoices), std::end(kEventChoices), e) == 1 && o >= 8 && o <= 64;
    if (ok) {
      t.events[b] = e;
      t.ocap[b] = o;
    }
  }
  std::fclose(f);
  if (!ok) {
    ctx->error = std::string("malformed tuning file ") + path;
    return VO_ERR_ARGUMENT;
  }
  ctx->tuning = t;
  return VO_OK;
}

int vo_last_batch_alone(const vo_context* ctx, int* jobs) {
  if (!ctx || !jobs) return VO_ERR_ARGUMENT;
  *jobs = ctx->last_batch_alone;
  return VO_OK;
}

int vo_last_stats(const vo_context* ctx, int64_t* output_endpoints, int64_t* fallback_tiles) {
  if (!ctx) return VO_ERR_ARGUMENT;
  if (output_endpoints) *output_endpoints = ctx->last_outputs;
  if (fallback_tiles) *fallback_tiles = ctx->last_fallback_tiles;
  return VO_OK;
}

}  // extern "C"
