/*
 * agwn.h -- C ABI of the MI355X (gfx950) batch generalized-winding-number library, libagwn.so.
 *
 * This is the boundary between host code and the hand-written HIP kernels. Every entry point takes plain pointers
 * and sizes, returns an int status (AGWN_OK == 0) and never throws; on failure agwn_last_error() returns a message
 * for the calling thread. No C++ or torch types cross it. The C++20 plugin headers in include/antipodal_hip/ wrap it
 * behind the reference's own Dispatcher / Intersector / RobustMeshGwn shapes (references below are to the
 * MartensCedric/antipodal sources, relative to its repository root).
 *
 * Memory layouts are the reference's in-memory layouts, so std::span data passes zero-copy:
 *   positions  n   x {x,y,z}                      (vec3<T>, math/common.hh:32-45)
 *   seg7       nseg x {p0.xyz, p1.xyz, weight}    (weighted_segment3<T>, weight LAST, math/common.hh:50-65)
 *   verts      nv  x {x,y,z};  idx  3*ntri int32  (the spans NaiveIntersector holds, intersector/intersector.hh:58-59)
 *
 * Entry points without a suffix take HOST pointers (inputs are copied to the device, outputs back; the call is
 * synchronous). Entry points with the _dev suffix take DEVICE pointers and a stream (hipStream_t, NULL = default
 * stream) and are asynchronous with respect to the host. Counts are int64; the reference casts them to int
 * (gwn_mesh.hh:75), so any count it accepts is accepted here.
 *
 * Threading: handles are immutable after creation; concurrent calls on one handle from several host threads are
 * safe (the reference requires this of intersectors, intersector/intersector.hh:28-30, and of RobustMeshGwn,
 * gwn_mesh_robust.hh:161-162).
 */
#ifndef AGWN_H
#define AGWN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AGWN_OK 0
#define AGWN_ERR_INVALID 1     /* null pointer, negative count, size mismatch, unknown enum */
#define AGWN_ERR_HIP 2         /* a HIP runtime call or kernel launch failed */
#define AGWN_ERR_NODEVICE 3    /* no usable HIP device */
#define AGWN_ERR_ALLOC 4       /* device allocation failed */
#define AGWN_ERR_UNSUPPORTED 5 /* argument combination not supported by this build */

/* Precision policies of the robust evaluator: robust::int64_grid / int128_grid / exact_float
 * (math/robust_predicates.hh:186-215). */
#define AGWN_INT64_GRID 0
#define AGWN_INT128_GRID 1
#define AGWN_EXACT_FLOAT 2

typedef void* agwn_stream; /* hipStream_t */

/* Message of the last failing call on this thread ("" if none). */
const char* agwn_last_error(void);
/* Library version, major*10000 + minor*100 + patch. */
int agwn_version(void);
/* 1 if a HIP device is usable from this process, else 0 (sets the last error). No kernel is launched. */
int agwn_device_available(void);

/* ---------------------------------------------------------------------------------------------------------------
 * Fractional term. Replaces eval_gwnr_mesh_batch_fractional (gwn_mesh.hh:67-83) and the per-pair kernel
 * signed_spherical_tri_area_half_unorm (math/common.hh:181-189): out[i] = sum_s w_s * area(-x0, p0_s - p_i,
 * p1_s - p_i) / (2 pi). x0 must be unit length, as in the reference (gwn_mesh.hh:30-31).
 * --------------------------------------------------------------------------------------------------------------- */
int agwn_frac_batch_f32(const float* pos, int64_t n, const float* seg7, int64_t nseg, const float* x0, float* out);
int agwn_frac_batch_f64(const double* pos, int64_t n, const double* seg7, int64_t nseg, const double* x0, double* out);
int agwn_frac_batch_f32_dev(const float* pos, int64_t n, const float* seg7, int64_t nseg, const float* x0, float* out,
                            agwn_stream stream);
int agwn_frac_batch_f64_dev(const double* pos, int64_t n, const double* seg7, int64_t nseg, const double* x0,
                            double* out, agwn_stream stream);
/* Allocation-free variant of agwn_frac_batch_f64_dev (same kernels, same results, bit for bit). It has no counterpart
 * in the reference: it is for callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Its one workspace, the boundary rotated into the pole frame (64 B per
 * segment), lives in a 32 MiB device array that the library holds on every GPU it runs on until the process ends, so
 * nseg may not exceed 524,288 (larger boundaries are rejected with AGWN_ERR_INVALID). The caller still passes every
 * input and output. All calls on one GPU share that array: calls that overlap on one GPU (other host threads, or
 * streams without ordering between them) race and must be serialised by the caller. */
int agwn_frac_batch_f64_dev_noalloc(const double* pos, int64_t n, const double* seg7, int64_t nseg, const double* x0,
                                    double* out, agwn_stream stream);

/* ---------------------------------------------------------------------------------------------------------------
 * Prepared triangle mesh: a GPU-built LBVH (Morton codes, LDS radix sort, Karras hierarchy) over the triangles.
 * Replaces the Intersector concept's implementations NaiveIntersector<T> (intersector/intersector.hh:55-101) and
 * EmbreeIntersector (intersector/intersector_embree.hh:56-143). The handle owns device copies (the reference's
 * intersectors hold non-owning spans, intersector.hh:31-35; owning is strictly more permissive).
 * verts/idx are host pointers. scalar type is fixed at creation (f32 or f64).
 * --------------------------------------------------------------------------------------------------------------- */
typedef struct agwn_mesh agwn_mesh;
int agwn_mesh_create_f32(const float* verts, int64_t nv, const int32_t* idx, int64_t nidx, agwn_mesh** out);
int agwn_mesh_create_f64(const double* verts, int64_t nv, const int32_t* idx, int64_t nidx, agwn_mesh** out);
int agwn_mesh_destroy(agwn_mesh* mesh);
/* ntri, number of BVH nodes, scalar bytes (4 or 8). Any output may be NULL. */
int agwn_mesh_info(const agwn_mesh* mesh, int64_t* ntri, int64_t* nnodes, int* scalar_bytes);

/* Integer term: out[i] = signed ray-crossing count of the ray (pos[i], dir) -- NaiveIntersector::
 * signed_intersection_count (intersector.hh:61-100), bit-exact: same Moeller-Trumbore arithmetic per triangle, the
 * BVH only culls triangles that test cannot accept. Replaces eval_gwnr_mesh_batch_integer (gwn_mesh.hh:127-143). */
int agwn_int_batch_f32(const agwn_mesh* mesh, const float* pos, int64_t n, const float* dir, int32_t* out);
int agwn_int_batch_f64(const agwn_mesh* mesh, const double* pos, int64_t n, const double* dir, int32_t* out);
int agwn_int_batch_f32_dev(const agwn_mesh* mesh, const float* pos, int64_t n, const float* dir, int32_t* out,
                           agwn_stream stream);
int agwn_int_batch_f64_dev(const agwn_mesh* mesh, const double* pos, int64_t n, const double* dir, int32_t* out,
                           agwn_stream stream);

/* Full GWN: out[i] = fractional + integer, eval_gwnr_mesh_batch (gwn_mesh.hh:201-219). seg7 may be NULL iff nseg == 0
 * (a closed mesh, gwn_mesh.hh:26-27). */
int agwn_full_batch_f32(const agwn_mesh* mesh, const float* pos, int64_t n, const float* x0, const float* seg7,
                        int64_t nseg, float* out);
int agwn_full_batch_f64(const agwn_mesh* mesh, const double* pos, int64_t n, const double* x0, const double* seg7,
                        int64_t nseg, double* out);
int agwn_full_batch_f32_dev(const agwn_mesh* mesh, const float* pos, int64_t n, const float* x0, const float* seg7,
                            int64_t nseg, float* out, agwn_stream stream);
int agwn_full_batch_f64_dev(const agwn_mesh* mesh, const double* pos, int64_t n, const double* x0,
                            const double* seg7, int64_t nseg, double* out, agwn_stream stream);

/* ---------------------------------------------------------------------------------------------------------------
 * Robust evaluator. Replaces RobustMeshGwn<T, Precision> (gwn_mesh_robust.hh:170-424) and
 * eval_gwnr_mesh_batch_robust (gwn_mesh_robust.hh:441-456). scalar_bytes = sizeof(T) (4 or 8); precision one of
 * AGWN_INT64_GRID / AGWN_INT128_GRID / AGWN_EXACT_FLOAT (the reference's defaults: int64_grid for float, exact_float
 * for double, robust_predicates.hh:220-221). verts/seg7 are T arrays (host); the constructor's quantization frame,
 * swapped quantized boundary and exact predicates are the reference's (gwn_mesh_robust.hh:177-209, 334-378). The
 * ray is baked in (-x, gwn_mesh_robust.hh:224-226); the handle owns its device copies (gwn_mesh_robust.hh:152-156).
 * --------------------------------------------------------------------------------------------------------------- */
typedef struct agwn_robust agwn_robust;
int agwn_robust_create(int scalar_bytes, int precision, const void* verts, int64_t nv, const int32_t* idx,
                       int64_t nidx, const void* seg7, int64_t nseg, agwn_robust** out);
int agwn_robust_destroy(agwn_robust* robust);
/* Any output may be NULL. center3/scale: the quantization frame (center 0, scale 1 for exact_float). */
int agwn_robust_info(const agwn_robust* robust, int* scalar_bytes, int* precision, int64_t* ntri, int64_t* nseg,
                     double* center3, double* scale);
/* eval: RobustMeshGwn::eval (gwn_mesh_robust.hh:328) per position -> T out[n] */
int agwn_robust_eval_batch(const agwn_robust* robust, const void* pos, int64_t n, void* out);
/* fractional: RobustMeshGwn::fractional (gwn_mesh_robust.hh:264-325) -> T out[n] */
int agwn_robust_fractional_batch(const agwn_robust* robust, const void* pos, int64_t n, void* out);
/* count: RobustMeshGwn::signed_intersection_count (gwn_mesh_robust.hh:231-258) -> int32 out[n], bit-exact */
int agwn_robust_count_batch(const agwn_robust* robust, const void* pos, int64_t n, int32_t* out);
int agwn_robust_eval_batch_dev(const agwn_robust* robust, const void* pos, int64_t n, void* out, agwn_stream stream);
int agwn_robust_fractional_batch_dev(const agwn_robust* robust, const void* pos, int64_t n, void* out,
                                     agwn_stream stream);
int agwn_robust_count_batch_dev(const agwn_robust* robust, const void* pos, int64_t n, int32_t* out,
                                agwn_stream stream);

/* ---------------------------------------------------------------------------------------------------------------
 * Dense-grid evaluator. Replaces OpenGLMeshGwn (gwn_mesh_opengl.hh:426-809): set_scene (:471-510) uploads the mesh
 * (building its LBVH) and the boundary; eval_grid (:519-643) evaluates the full GWN at every lattice point
 * p = start + ix*dx + iy*dy + iz*dz, out[((iz*ny)+iy)*nx+ix] (host: agwn_grid_eval; device, written straight to HBM
 * on `stream`: agwn_grid_eval_dev). FP32 throughout, as the reference's evaluator.
 * --------------------------------------------------------------------------------------------------------------- */
typedef struct agwn_grid agwn_grid;
int agwn_grid_create(agwn_grid** out);
int agwn_grid_destroy(agwn_grid* grid);
int agwn_grid_set_scene(agwn_grid* grid, const float* verts, int64_t nv, const int32_t* idx, int64_t nidx,
                        const float* seg7, int64_t nseg);
int agwn_grid_eval(agwn_grid* grid, const float* start, const float* dx, const float* dy, const float* dz, int nx,
                   int ny, int nz, float* out);
int agwn_grid_eval_dev(agwn_grid* grid, const float* start, const float* dx, const float* dy, const float* dz,
                       int nx, int ny, int nz, float* out, agwn_stream stream);

/* ---------------------------------------------------------------------------------------------------------------
 * Open boundary. Replaces build_boundary_segments (math/common.hh:123-174): the same multiset of oriented weighted
 * segments; the order is sorted by (min index, max index) -- the reference's order is unordered_map order and so is
 * unspecified. Writes min(count, cap) segments to out7 (host) and the total count to *count.
 * --------------------------------------------------------------------------------------------------------------- */
int agwn_boundary_f32(const float* verts, int64_t nv, const int32_t* idx, int64_t nidx, float* out7, int64_t cap,
                      int64_t* count);
int agwn_boundary_f64(const double* verts, int64_t nv, const int32_t* idx, int64_t nidx, double* out7, int64_t cap,
                      int64_t* count);
/* The same on the GPU (device verts/idx/out7, host count; returns when done): half-edge keys radix-sorted, net
 * direction per undirected edge, nonzero edges compacted -- element for element the host functions' output. */
int agwn_boundary_f32_dev(const float* verts, int64_t nv, const int32_t* idx, int64_t nidx, float* out7, int64_t cap,
                          int64_t* count, agwn_stream stream);
int agwn_boundary_f64_dev(const double* verts, int64_t nv, const int32_t* idx, int64_t nidx, double* out7,
                          int64_t cap, int64_t* count, agwn_stream stream);

/* ---------------------------------------------------------------------------------------------------------------
 * Batched evaluation of many meshes with uneven query counts: eval_gwnr_mesh_batch (gwn_mesh.hh:201-219) of every
 * mesh's own queries against that mesh, for the whole batch in two persistent launches (integer term, then
 * fractional term + counts) captured in a HIP graph and replayed while the inputs stay the same. Per query the
 * arithmetic is the single-mesh path's (agwn_full_batch_f32 with the same x0), so results are identical.
 * Meshes are concatenated: verts (host float[3 * vert_off[k]]), idx (host int32, indices local to each mesh,
 * idx_off[k + 1], multiples of 3), boundaries seg7 (host float[7 * seg_off[k]], integer weights with |w| <= 16, e.g.
 * build_boundary_segments output). Queries (device float[3 * q_off[k]]) are concatenated per mesh in the same order,
 * q_off host int64[k + 1]; out device float[q_off[k]]; x0 host float[3] (unit), shared by the batch.
 * --------------------------------------------------------------------------------------------------------------- */
typedef struct agwn_batch agwn_batch;
int agwn_batch_create(int64_t k, const float* verts, const int64_t* vert_off, const int32_t* idx,
                      const int64_t* idx_off, const float* seg7, const int64_t* seg_off, agwn_batch** out);
int agwn_batch_destroy(agwn_batch* batch);
/* The autotuner's search space. Fractional kernel shapes v in [0, AGWN_BATCH_FRAC_SHAPES): queries per lane, lanes
 * per workgroup and stream points per LDS stage (in multiples of the workgroup size), see agwn_batch_frac_shape.
 * Leaf-bucket sizes l in [0, AGWN_BATCH_LEAF_SIZES) of the counting traversal (LBVH subtrees of at most that many
 * triangles tested as one leaf), see agwn_batch_leaf_size. Results do not depend on either choice. */
#define AGWN_BATCH_FRAC_SHAPES 15
#define AGWN_BATCH_LEAF_SIZES 5
int agwn_batch_frac_shape(int v, int* frac_q, int* frac_block, int* frac_stage);
int agwn_batch_leaf_size(int l, int* leaf);
/* Autotuner: times each fractional kernel shape on every 8th mesh, then each leaf-bucket size on the whole batch's
 * integer term, and keeps the fastest of each (run by the first agwn_batch_eval_dev unless tuned or set before).
 * `out` is used as scratch. Synchronous. */
int agwn_batch_tune(agwn_batch* batch, const float* pos, const int64_t* q_off, const float* x0, float* out,
                    agwn_stream stream);
/* Sets the shape and leaf-bucket size instead of tuning (e.g. a table saved from an earlier agwn_batch_info). */
int agwn_batch_set_tuning(agwn_batch* batch, int frac_shape, int leaf);
/* out[i] = full GWN of query i against its mesh. Ordered after prior work on `stream`; returns when done. */
int agwn_batch_eval_dev(agwn_batch* batch, const float* pos, const int64_t* q_off, const float* x0, float* out,
                        agwn_stream stream);
/* Mesh count, triangles, unit boundary segments, the chosen shape and leaf-bucket size indexes (-1 before tuning),
 * and the tuner's timings in ms (frac_ms[AGWN_BATCH_FRAC_SHAPES] per shape, on the sample; leaf_ms
 * [AGWN_BATCH_LEAF_SIZES] per size, the integer term of the whole batch; 0 where not timed); any output may be
 * NULL. */
int agwn_batch_info(const agwn_batch* batch, int64_t* k, int64_t* ntri, int64_t* nunit, int* frac_shape, int* leaf,
                    double* frac_ms, double* leaf_ms);

/* ---------------------------------------------------------------------------------------------------------------
 * Learned inside/outside surrogate (no counterpart in the reference: an accelerator in front of the exact path, whose
 * labels stay authoritative). A per-mesh MLP -- 192 Fourier features of the position normalised to a box (32 frequencies
 * max_freq^(k/31) per axis, sin and cos), 4 hidden layers x 256 hard-SiLU (z clamp(z/4 + 1/2, 0, 1)), linear output -- with BF16 weights/activations and
 * FP32 accumulation, trained on GPU-computed exact GWN values by the library's own MFMA forward/backward kernels and
 * Adam (FP32 master weights). Labels are GWN > 0.5 ("inside").
 * --------------------------------------------------------------------------------------------------------------- */
typedef struct agwn_mlp agwn_mlp;
/* Number of FP32 parameters (W0, b0, ..., W3, b3, wo, bo; W_l row-major [fan_in x 256]). */
int64_t agwn_mlp_num_params(void);
/* New network for positions in the box [lo, hi] (host float[3] each), top frequency max_freq in [1, 4096] (periods
 * over the box's full width / 2), deterministic init from seed. */
int agwn_mlp_create(const float* lo, const float* hi, float max_freq, uint64_t seed, agwn_mlp** out);
int agwn_mlp_destroy(agwn_mlp* mlp);
/* The box and top frequency (any output may be NULL). */
int agwn_mlp_info(const agwn_mlp* mlp, float* lo, float* hi, float* max_freq);
/* Host copies of the FP32 parameters (set resets the Adam state). */
int agwn_mlp_get_params(const agwn_mlp* mlp, float* out);
int agwn_mlp_set_params(agwn_mlp* mlp, const float* in);
/* Adam (b1 0.9, b2 0.999, eps 1e-8, cosine decay lr -> lr/20) on the mean squared error against target[i] over
 * pts[i] (device float[3n], float[n]); each epoch visits every sample once in a seeded order, batch a multiple of
 * 128. loss_per_epoch (host float[epochs], may be NULL) receives each epoch's mean loss. Synchronous. */
int agwn_mlp_train_dev(agwn_mlp* mlp, const float* pts, const float* target, int64_t n, int batch, int epochs,
                       float lr, uint64_t seed, float* loss_per_epoch, agwn_stream stream);
/* One forward + backward pass over host samples (n <= 2^20) without an update: grad (float[num_params]), pred
 * (float[n]), loss (mean squared error); outputs may be NULL. For gradient checks. */
int agwn_mlp_grad(agwn_mlp* mlp, const float* pts, const float* target, int64_t n, float* grad, float* pred,
                  float* loss);
/* Fused inference (device arrays): out[i] = network(pts[i]). */
int agwn_mlp_predict_dev(const agwn_mlp* mlp, const float* pts, int64_t n, float* out, agwn_stream stream);
/* Fused inference over the lattice of agwn_grid_eval (host start/dx/dy/dz float[3]; out: device float[nx ny nz] in its
 * layout). When each coordinate steps along one lattice axis at most (axis-aligned lattices, any axis order) and
 * nx + ny + nz <= 65536, the first layer is a sum of three per-axis tables of the same BF16 products (the points'
 * features depend on one index each), so outputs may differ from agwn_mlp_predict_dev's on the same points by the
 * order of FP32 sums in that layer; otherwise it is agwn_mlp_predict_dev's per-point kernel. */
int agwn_mlp_predict_grid_dev(const agwn_mlp* mlp, const float* start, const float* dx, const float* dy,
                              const float* dz, int nx, int ny, int nz, float* out, agwn_stream stream);
/* Near-surface mask of the network's box: voxels of edge `voxel` (<= 0: eps/2) within eps of a triangle of the mesh
 * (host verts/idx), marked conservatively on the GPU. Points in marked voxels, or outside the box, are never labelled
 * by the network. */
int agwn_mlp_set_surface(agwn_mlp* mlp, const float* verts, int64_t nv, const int32_t* idx, int64_t nidx, float eps,
                         float voxel);
/* n training positions (device float[3n]) of the surrogate: with probability near_frac a point at a signed offset
 * band s u^power (s = +-1, u uniform) along the unit normal of an area-uniform point of the mesh (device verts
 * float[3nv], idx int32[3 ntri]; cdf: device double[ntri], the inclusive prefix sums of |(b - a) x (c - a)| in triangle
 * order), otherwise uniform in the box [lo, hi] (host float[3] each). Counter-based SplitMix64 streams (seed, 0..5):
 * point i depends on (seed, i) only. */
int agwn_training_samples_dev(const float* verts, const int32_t* idx, const double* cdf, int64_t ntri, const float* lo,
                              const float* hi, int64_t n, double near_frac, double band, double power, uint64_t seed,
                              float* out, agwn_stream stream);
/* out[i] = 1 if pts[i] (device) is in the mask or outside the box, else 0 (device int8). */
int agwn_mlp_near_surface_dev(const agwn_mlp* mlp, const float* pts, int64_t n, int8_t* out, agwn_stream stream);
/* Surrogate + fallback labelling of points: labels[i] = (GWN(pts[i]) > 0.5). The network labels every point; points
 * in the near-surface mask or outside the box, points with |pred - 0.5| <= delta, plus an audit sample of about
 * 1/audit_every of the others (0: none, hashed with audit_seed), go through the exact FP32 path (agwn_full_batch_f32
 * with this mesh, x0 and boundary) and take its label. stats (host int64[5], may be NULL): points routed by the band,
 * by the mask, audited, routed in all, audited points whose network label disagreed with the exact one. Device
 * pos/seg7/labels; host x0. Synchronous. */
int agwn_label_surrogate_dev(const agwn_mesh* mesh, const agwn_mlp* mlp, const float* pos, int64_t n, const float* x0,
                             const float* seg7, int64_t nseg, float delta, int64_t audit_every, uint64_t audit_seed,
                             int8_t* labels, int64_t* stats, agwn_stream stream);
/* The same over the lattice of agwn_grid_eval (labels in its layout): routed points take the labels of
 * agwn_grid_eval_dev's values, bit for bit. */
int agwn_grid_label_surrogate_dev(agwn_grid* grid, const agwn_mlp* mlp, const float* start, const float* dx,
                                  const float* dy, const float* dz, int nx, int ny, int nz, float delta,
                                  int64_t audit_every, uint64_t audit_seed, int8_t* labels, int64_t* stats,
                                  agwn_stream stream);

#ifdef __cplusplus
}
#endif

#endif /* AGWN_H */
