/* voroffset_hip.h -- C ABI of the MI355X dexel morphology library.
 *
 * Replaces, for the `ours` method of geometryprocessing/voroffset, the plugin
 * interface `vor3d::VoronoiMorpho` (src/vor3d/Voronoi.h:13-44):
 *   - VoronoiMorpho::dilation(CompressedVolume input, CompressedVolume& result,
 *       double radius, double& time_1, double& time_2)   Voronoi.h:18
 *   - VoronoiMorpho::erosion(...)                          Voronoi.h:29, Voronoi.cpp:8-17
 *   - closing / opening as composed by the CLI            app/cli3d/offset3d.cpp:124-133
 * and its data layout `vor3d::CompressedVolume` (src/vor3d/CompressedVolume.h:10-41):
 * a grid nx x ny of z-columns indexed x + nx*y, each a sorted list of endpoints
 * [b0, e0, b1, e1, ...] in dexel units of absolute z.  Here a volume is a CSR
 * pair: `offsets` (nx*ny + 1 int64) and `z` (float or double endpoints).
 *
 * Conventions (mirroring the reference):
 *   - radius is in dexels; time_1 / time_2 report pass 1 / pass 2 in ms;
 *   - erosion complements on a grid grown by one ring of full columns over
 *     [z_lo - 1, z_hi + 1], dilates, crops and complements within [z_lo, z_hi],
 *     where z_lo = origin_z / spacing and z_hi = z_lo + 2*padding + extent_z / spacing;
 *   - output columns are sorted and disjoint, touching intervals fused.
 * Errors are reported as a nonzero status; vo_last_error() gives the text
 * (the reference throws std::runtime_error from vor_assert, Common.cpp:8-18).
 * One context per host thread; calls on a context are serialised on its stream.
 */
#ifndef VOROFFSET_HIP_H
#define VOROFFSET_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vo_context vo_context;
typedef struct vo_result vo_result;
typedef struct vo_mesh_result vo_mesh_result;
typedef struct vo_triples vo_triples;

enum vo_status {
  VO_OK = 0,
  VO_ERR_ARGUMENT = 1, /* bad shape, radius, unsorted or odd column, ... */
  VO_ERR_DEVICE = 2,   /* HIP runtime error */
  VO_ERR_CAPACITY = 3, /* a sweep exceeded its working-set capacity */
  VO_ERR_RANGE = 4     /* erosion input endpoint outside [z_lo - 1, z_hi + 1] */
};

enum vo_op { VO_OP_NOOP = 0, VO_OP_DILATION = 1, VO_OP_EROSION = 2, VO_OP_CLOSING = 3, VO_OP_OPENING = 4 };

/* Storage precision of the endpoints on the device.  Predicates and endpoint
 * arithmetic are FP64 in both modes; VO_FP32 halves the bytes moved. */
enum vo_precision { VO_FP32 = 0, VO_FP64 = 1 };

typedef struct {
  int32_t nx, ny;          /* grid size */
  const int64_t* offsets;  /* nx*ny + 1 column offsets into z */
  const void* z;           /* endpoints: float (VO_FP32) or double (VO_FP64) */
  double z_lo, z_hi;       /* z-range used by erosion (see above) */
} vo_volume;

/* Context bound to HIP device `device`: owns a stream and a workspace pool.
 * Environment read here: VO_INJECT_HIP_ERROR=<label> makes every HIP call of
 * the context whose internal label contains <label> ("pass1 launch", "sync",
 * "upload", ...) report VO_ERR_DEVICE after it ran -- a fault hook for the
 * error paths of callers (the reference's vor_assert failures). */
int vo_create(vo_context** ctx, int device);
void vo_destroy(vo_context* ctx);
/* Text of the last error on `ctx`; with ctx = NULL, why this thread's last
 * vo_create failed ("no context" when it did not). */
const char* vo_last_error(const vo_context* ctx);

/* Host volume in (z as double), host result out.  With VO_FP32 the endpoints
 * travel as float relative to an integer z-shift (the rounded centre of
 * [z_lo, z_hi]) and come back as double.  Replaces VoronoiMorpho::dilation /
 * erosion and the CLI's closing / opening composition. */
int vo_morph(vo_context* ctx, int op, int precision, double radius, const vo_volume* in,
             vo_result** out, double* time_1_ms, double* time_2_ms);
int vo_result_shape(const vo_result* r, int32_t* nx, int32_t* ny, int64_t* count);
int vo_result_copy(const vo_result* r, int64_t* offsets, double* z);
void vo_result_free(vo_result* r);

/* The pass-1 intermediate of a dilation of a host volume by `radius`: replaces
 * reading the vor3d::CompressedVolumeWithRadii that VoronoiMorphoVorPower
 * builds between its passes (src/vor3d/VoronoiVorPower.cpp:39-66,
 * CompressedVolumeWithRadii.h:10-38, Common.h:119-155).  Per column (index
 * x + nx*y) the z-sorted, z-disjoint triples (y1, y2, r): for every z in
 * [y1, y2] the nearest line y' = y +- dy (|dy| <= radius) of the same x whose
 * input covers z is at distance dy, and r = sqrt(radius^2 - dy^2).  The GPU
 * keeps every such piece; the reference also drops pieces its sweep finds
 * dominated, so counts differ while the envelope max_t (r^2 - dist_z(z, t)^2)
 * is the same.  A library-owned host result (shape / copy / free). */
int vo_pass1(vo_context* ctx, int precision, double radius, const vo_volume* in, vo_triples** out);
int vo_triples_shape(const vo_triples* t, int32_t* nx, int32_t* ny, int64_t* count);
int vo_triples_copy(const vo_triples* t, int64_t* offsets, double* y1, double* y2, double* r);
void vo_triples_free(vo_triples* t);

/* Device volume in (HBM pointers, z of the given precision), device result
 * out: `out` receives pointers into the context's workspace, valid until the
 * next call on `ctx`.  `*out_count` is the number of endpoints.  The input is
 * checked on the device as vo_morph checks it on the host: decreasing
 * offsets, an odd column, unsorted or non-finite endpoints -> VO_ERR_ARGUMENT;
 * for erosion / opening an endpoint outside [z_lo - 1, z_hi + 1], for a
 * closing one whose dilation by `radius` leaves it -> VO_ERR_RANGE (the
 * reference's preconditions, Voronoi.cpp:27-42, Common.h:228).  The offsets
 * must index the z buffer. */
int vo_morph_device(vo_context* ctx, int op, int precision, double radius, const vo_volume* in,
                    vo_volume* out, int64_t* out_count, double* time_1_ms, double* time_2_ms);

/* Batched form (no single reference counterpart: it replaces a loop of
 * VoronoiMorpho::dilation / erosion calls and the CLI's closing / opening,
 * Voronoi.h:18-29, offset3d.cpp:116-136, over many volumes).  Each job names
 * an op, a radius and a device volume whose z has the batch's precision; all
 * jobs run together, one launch per pass and stage for the whole batch.
 * Inputs are checked on the device as for vo_morph_device; the error text
 * names the first failing job.
 * out[i] receives device pointers into the context's workspace (offsets
 * starting at 0), valid until the next call on `ctx`; out_count[i] is its
 * number of endpoints.  The batch runs as one program: output sizes are
 * bounded before it starts, so the host waits once for the input sizes and
 * once at its end; a batch identical to the previous one (same jobs,
 * pointers, sizes) runs from a HIP graph captured on its second call
 * (VO_BATCH_GRAPH=0 at vo_create disables that).  time_1 is the program's
 * device time (all primitives), time_2 that of the jobs redone one by one.
 * Volumes whose working lists outgrow the batched kernels are redone one by
 * one with the single-volume fallbacks; results do not depend on it. */
typedef struct {
  int op;         /* vo_op */
  double radius;  /* dexels */
  vo_volume in;   /* device pointers */
} vo_job;

int vo_morph_batch_device(vo_context* ctx, int precision, int n, const vo_job* jobs, vo_volume* out,
                          int64_t* out_count, double* time_1_ms, double* time_2_ms);

/* vo_morph_device without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * operations, kernels and results as vo_morph_device; the workspace is one
 * 64 MiB array per GPU (a static device array, held until the process ends),
 * and a volume whose workspace would not fit in it is rejected with
 * VO_ERR_CAPACITY (use vo_morph_device for it).  The caller passes the output
 * buffers: out_offsets (nx*ny + 1 int64, the result's offsets from 0) and
 * out_z (out_capacity >= 1 endpoints of the storage type); when out_z is too
 * small the call returns VO_ERR_CAPACITY with *out_count the capacity
 * needed.  All calls on one GPU share its scratch, which keeps the last
 * call's contents: calls that overlap on one GPU (other contexts, other
 * threads) race and must be serialised by the caller. */
int vo_morph_device_noalloc(vo_context* ctx, int op, int precision, double radius, const vo_volume* in,
                            int64_t* out_offsets, void* out_z, int64_t out_capacity, int64_t* out_count,
                            double* time_1_ms, double* time_2_ms);

/* n device-to-device copies dst[i] <- src[i] (bytes[i] each) on the
 * context's stream, then one synchronisation: how a caller keeps results of
 * vo_morph_batch_device / vo_morph_device beyond the next call, into buffers
 * it owns.  No reference counterpart (the reference returns results by
 * value, Voronoi.h:18). */
int vo_copy_device(vo_context* ctx, int n, void* const* dst, const void* const* src, const int64_t* bytes);

/* Number of jobs of the last batch on `ctx` that were redone one by one
 * (a tuning signal).  No reference counterpart. */
int vo_last_batch_alone(const vo_context* ctx, int* jobs);

/* Parity metrics on a device volume (z of the given precision): the measure
 * spacing^3 * sum of segment lengths and the number of segments, as
 * CompressedVolume::get_volume / numSegments (src/vor3d/CompressedVolume.cpp:61-83;
 * an odd last endpoint of a column is ignored). */
int vo_measure_device(vo_context* ctx, int precision, const vo_volume* v, double spacing, double* volume,
                      int64_t* segments);

/* VoronoiMorpho::calculateXor (src/vor3d/Voronoi.cpp:91-111): per column the
 * symmetric difference of a and b (equal grids) within [a->z_lo, a->z_hi],
 * composed as calculate_ray_xor does (MorphologyOperators.cpp:319-374,
 * segments shorter than 1e-10 dropped).  `out` receives device pointers into
 * the workspace (valid until the next call), `*volume` its measure. */
int vo_xor_device(vo_context* ctx, int precision, const vo_volume* a, const vo_volume* b, double spacing,
                  vo_volume* out, int64_t* out_count, double* volume);

/* Volume -> hexahedral boundary mesh: replaces dexel_dump
 * (src/vor3d/Dexelize.cpp:312-351), computed on the device.  Host volume in
 * (double z); every segment [z_i, z_{i+1}] of column (x, y) becomes the box
 * [origin_x + x*s, origin_x + (x+1)*s] x [origin_y + y*s, ...] x [z_i*s, z_{i+1}*s]:
 * 8 vertices in the reference's corner order and 6 outward quads (vertex
 * indices), as a library-owned host result. */
int vo_dexel_mesh(vo_context* ctx, const vo_volume* v, double origin_x, double origin_y, double spacing,
                  vo_mesh_result** out);
int vo_mesh_shape(const vo_mesh_result* r, int64_t* num_vertices, int64_t* num_quads);
int vo_mesh_copy(const vo_mesh_result* r, double* vertices, int64_t* quads);
void vo_mesh_free(vo_mesh_result* r);

/* Autotuner (no reference counterpart): for each radius (at most 127), times
 * the z-sweep event capacity per slab (512, 768, 1024, 1536) and the gather
 * list capacity (8, 16, 24, 32) on `sample` (a device volume of the given
 * precision), `reps` runs each, and keeps the fastest per radius bucket
 * ([0,8), [8,16), [16,32), [32,64), [64,128)) in `ctx`; results do not
 * depend on the choice.  The table can be saved, loaded and read back. */
int vo_autotune(vo_context* ctx, int precision, const vo_volume* sample, int num_radii, const double* radii,
                int reps);
int vo_tuning_save(const vo_context* ctx, const char* path);
int vo_tuning_load(vo_context* ctx, const char* path);
int vo_tuning_get(const vo_context* ctx, int bucket, int* events, int* ocap);

/* Mesh -> dexels: replaces the ray casting of vor3d::create_dexels
 * (src/vor3d/Dexelize.cpp:166-225; predicates :56-162).  `vertices` is
 * num_vertices x 3 doubles (x, y, z), `triangles` num_triangles x 3 vertex
 * indices.  Column x + nx*y of the host result holds the sorted z / spacing of
 * every triangle whose xy-projection contains the dexel centre
 * (origin_x + (x + 0.5) * spacing, origin_y + (y + 0.5) * spacing), with the
 * reference's SOS tie-break on shared edges and vertices; consecutive pairs are
 * the column's segments (even-odd).  The grid (nx, ny, origin, spacing) is the
 * caller's, as create_dexels derives it from the mesh bounding box
 * (Dexelize.cpp:252-270, CompressedVolume.cpp:11-23). */
int vo_rasterize(vo_context* ctx, const double* vertices, int64_t num_vertices, const int32_t* triangles,
                 int64_t num_triangles, int32_t nx, int32_t ny, double origin_x, double origin_y, double spacing,
                 vo_result** out);

/* Statistics of the last dilation on `ctx`: the number of output endpoints
 * and the number of 64-column pass-2 tiles whose interval lists outgrew LDS
 * and were recomputed by the sweep form (a tuning signal; results do not
 * depend on it).  No reference counterpart. */
int vo_last_stats(const vo_context* ctx, int64_t* output_endpoints, int64_t* fallback_tiles);

#ifdef __cplusplus
}
#endif

#endif /* VOROFFSET_HIP_H */
