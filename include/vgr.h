/*
 * vgr.h — C ABI of the MI355X (gfx950) batched Gaussian-splat camera renderer.
 *
 * Every entry point takes plain device pointers, sizes and a hipStream_t (passed
 * as void*), allocates nothing, and returns 0 on success or a nonzero VGR_E_*
 * code; vgr_last_error() then names the failing argument.  All memory is owned
 * by the caller.  Work is enqueued on the given stream; nothing synchronizes
 * the host unless the function says so.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   vgr_render                  <- render(...)          src/isaacsim_gaussian_renderer/native/renderer.cpp:214-310
 *                                  render_cuda(...)     src/isaacsim_gaussian_renderer/native/renderer_cuda.cu:2536-3456
 *   vgr_precompute_covariances  <- precompute_covariances(...) renderer.cpp:89-124, renderer_cuda.cu:373-394
 *   vgr_sort_temp_bytes         <- sort_temp_bytes / radix_sort_double_buffer_temp_bytes
 *                                  renderer.cpp:75-88, renderer_cuda.cu:2463-2534
 *   vgr_emitted_sort_key_end_bit<- emitted_sort_key_end_bit renderer_cuda.cu:28-46
 *   vgr_prepare_chunked_active_ids <- prepare_chunked_active_ids renderer.cpp:126-190, renderer_cuda.cu:2397-2444
 *   vgr_aggregate_chunk_counters   <- aggregate_chunk_counters  renderer.cpp:192-212, renderer_cuda.cu:2446-2461
 *   vgr_last_error              <- TORCH_CHECK messages surfaced as RuntimeError (renderer.cpp:9-25)
 */
#ifndef VGR_H
#define VGR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VGR_ABI_VERSION 7

enum vgr_status {
  VGR_OK = 0,
  VGR_E_INVALID_ARGUMENT = 1,
  VGR_E_LAUNCH = 2,
  VGR_E_CAPACITY = 3,
};

/* Counter block written by vgr_render (int64[6]).  Entries 0-4 have the order
 * and meaning of the reference's COUNTER_NAMES (cuda_backend.py:22-28; the
 * reference requires int64[>=5]).  Entry 5 counts radix-sort look-back waits
 * that gave up (bounded spin): nonzero means the intersection order of this
 * render is wrong and its images must not be used; callers treat it like an
 * overflow (re-render or fail closed).  No reference counterpart (CUB's sort
 * has no such bound). */
enum vgr_counter {
  VGR_COUNTER_VISIBLE = 0,
  VGR_COUNTER_INTERSECTIONS = 1,
  VGR_COUNTER_VISIBLE_OVERFLOW = 2,
  VGR_COUNTER_INTERSECTION_OVERFLOW = 3,
  VGR_COUNTER_ACTIVE_TILES = 4,
  VGR_COUNTER_SORT_FAILURES = 5,
  VGR_COUNTER_COUNT = 6,
};

/* Packed scene, one entry per Gaussian over all registered scenes
 * (reference scene tensor list, renderer_cuda.cu:48-56). */
typedef struct vgr_scene {
  const float* means;               /* [G,3] */
  const float* covariances;         /* [G,6] upper triangle 00,01,02,11,12,22 */
  const float* opacities;           /* [G] activated */
  const float* features;            /* [G,F]; F = 3 * (sh_degree+1)^2, k-major */
  const int64_t* semantic_ids;      /* [G] */
  const int64_t* registered_scene_ids; /* [S] */
  const int64_t* scene_offsets;     /* [S+1] */
  int64_t gaussian_count;           /* G */
  int64_t max_scene_gaussians;      /* largest registered scene */
  int32_t feature_width;            /* F >= 3 */
  int32_t scene_count;              /* S */
  int32_t sh_degree;                /* 0..3; bands beyond F/3 are absent */
  int32_t reserved;
  /* Optional spatial ordering.  When the packed arrays above are stored in a
   * permuted (e.g. Morton) order, gaussian_ids[p] is the registration-order id
   * of packed entry p; records, semantic lookups and tie order use those ids,
   * and semantic_ids stays in registration order.  NULL = identity. */
  const int32_t* gaussian_ids;      /* [G] or NULL */
  /* Optional per-chunk bounds from vgr_chunk_bounds (256 Gaussians per chunk):
   * enables frustum culling of (camera group x chunk) work items.  NULL = off. */
  const float* chunk_bounds;        /* [ceil(G/256), 8] */
} vgr_scene;

/* Camera batch (reference camera tensor list, renderer_cuda.cu:58-64). */
typedef struct vgr_cameras {
  const float* viewmats;            /* [B,4,4] row-major world->OpenCV camera */
  const float* intrinsics;          /* [B,3,3] */
  const float* env_xforms;          /* [B,4,4] scene->world */
  const int64_t* scene_ids;         /* [B] */
  const int64_t* active_ids;        /* [A] or NULL (all cameras active) */
  int32_t batch;                    /* B */
  int32_t active_count;             /* A (== B when active_ids is NULL) */
  /* Per-camera resolution.  NULL means every camera is width x height. */
  const int32_t* widths;            /* [B] or NULL */
  const int32_t* heights;           /* [B] or NULL */
  const int64_t* pixel_offsets;     /* [B] first output pixel of each camera, or NULL */
  /* Sum of tile counts over all B cameras at options.tile_size; required when
   * widths/heights are given (the host cannot read device resolutions). */
  int64_t max_global_tiles;
  /* Largest tile count of one camera (0: derived from width/height, or
   * max_global_tiles with per-camera resolutions); sizes the LDS tile tables
   * of the tile-front ordering. */
  int64_t max_camera_tiles;
} vgr_cameras;

/* Outputs (reference output tensor list, renderer_cuda.cu:66-71).  With
 * pixel_offsets NULL the layout is [B,H,W,C]; otherwise camera b owns pixels
 * [pixel_offsets[b], pixel_offsets[b] + H_b*W_b) of a flat [P,C] buffer.
 * depth/alpha/semantic may be NULL when full_sensor_output is 0. */
typedef struct vgr_outputs {
  float* rgb;                       /* [.,3] */
  float* depth;                     /* [.,1] */
  float* alpha;                     /* [.,1] */
  int64_t* semantic;                /* [.,1] */
  /* Optional training checkpoints (both or neither; need full_sensor_output
   * and output_srgb == 0).  Per pixel: the final transmittance T and the
   * position, within the pixel's tile list, one past the last blended splat
   * (0 when none).  No reference counterpart: the reference has no backward
   * pass (KERNEL_ARCHITECTURE.md:454-460). */
  float* final_transmittance;       /* [.] */
  int32_t* last_blended;            /* [.] */
} vgr_outputs;

/* Visible-record layout (48 bytes, 16-byte aligned).  One record per visible
 * (camera, Gaussian) pair, in camera-then-Gaussian order for each camera. */
typedef struct vgr_record {
  float mean_x, mean_y;             /* pixel-space centre */
  float conic_0, conic_1, conic_2;  /* inverse 2D covariance (xx, xy, yy) */
  float opacity;                    /* clamped to [0,1] */
  float depth;                      /* camera-space z */
  float red, green, blue;           /* view-dependent colour (SH evaluated) */
  int32_t gaussian_id;              /* packed global Gaussian index */
  int32_t camera_id;                /* camera index in the batch */
} vgr_record;

/* Workspace: caller-owned device buffers (sizes from vgr_workspace_sizes). */
typedef struct vgr_workspace {
  vgr_record* records;              /* [visible_capacity] */
  int32_t* radii;                   /* [visible_capacity,2] */
  uint64_t* keys_in;                /* [intersection_capacity] */
  uint64_t* keys_out;
  int32_t* values_in;
  int32_t* values_out;
  int32_t* tile_starts;             /* [tile_capacity] */
  int32_t* tile_ends;
  int64_t* counters;                /* [VGR_COUNTER_COUNT] */
  void* scratch;                    /* [scratch_bytes] */
  int64_t visible_capacity;
  int64_t intersection_capacity;
  int64_t tile_capacity;            /* global tiles the tile arrays can hold */
  int64_t scratch_bytes;
  /* Training forwards (outputs.final_transmittance set) only, else NULL:
   * [intersection_capacity] map from intersection slot to visible record.  The
   * sorted values are then the slots (the emission position of each
   * intersection) instead of record indices, so the backward can store each
   * (record, tile) pair's gradient terms in its own row without atomics.  No
   * reference counterpart (the reference is forward-only). */
  int32_t* slot_records;
  /* Exact-ray renders (options.ray_evaluation) only, else NULL: [visible_capacity, 12]
   * per visible record, the camera-space precision matrix (6 floats: 00, 01, 02,
   * 11, 12, 22), precision x mean (3), mean^T precision mean (1) and 2 pad floats
   * (renderer_cuda.cu:795-860, kVisibleRayPrecisions / kVisibleRayPrecisionMeans). */
  float* ray_records;
} vgr_workspace;

/* Render options (reference scalar arguments, renderer.cpp:214-238). */
typedef struct vgr_options {
  int32_t height;                   /* used when cameras.heights is NULL */
  int32_t width;
  float near_plane;
  float far_plane;
  float gaussian_support_sigma;
  float covariance_epsilon;
  float semantic_min_alpha;
  int32_t antialiased;
  int32_t tile_size;                /* 1, 2, 4, 8 or 16 (power of two, renderer_cuda.cu tile_size) */
  int32_t output_srgb;
  int32_t deterministic;            /* accepted; ordering is always deterministic */
  int32_t full_sensor_output;       /* 1: rgb+depth+alpha+semantic, 0: rgb only */
  int32_t fixed_capacity_sort;      /* accepted; the sort never reads the count on the host */
  int32_t sort_digit_bits;          /* 0 = default (8, or 9 where that saves a pass); 4..9 */
  int32_t raster_batch;             /* Gaussians staged per LDS batch: 0 = default (256); 64..512 */
  int32_t cameras_per_block;        /* projection camera group: 0 = default (32); 1..64 */
  int32_t reuse_projection;         /* 1: composite only, reusing the sorted intersections and
                                       tile ranges of the previous render into this workspace
                                       (reference projection cache, cuda_backend.py:607-619) */
  /* Optional instrumentation (NULL / 0 to disable). */
  void* const* stage_events;        /* 5 hipEvent_t recorded on the stream: start, after projection +
                                       emission, after sort, after tile ranges, after compositing */
  int32_t stage_event_count;        /* 0 or 5 */
  int32_t sort_tile_items;          /* keys per thread of a 512-thread sort tile: 0 = default (16); 8, 12 or 16 */
  int64_t* statistics;              /* int64[4], zeroed then filled: [0] compositor (pixel, Gaussian)
                                       evaluations; [1] sort look-back give-ups (as counters[5]);
                                       [2] projection work items kept by chunk culling; [3] tiles the
                                       tile-front ordering had to sort in full */
  int32_t persistent_composite;     /* 0: one workgroup per (camera, tile); N > 0: N workgroups pull
                                       (camera, tile) items from a device queue */
  int32_t chunk_culling;            /* 1: cull (camera group x chunk) items against camera frusta using
                                       scene.chunk_bounds (scratch sized for it) */
  int32_t lookback_spin_limit;      /* radix-sort look-back polls of an unpublished predecessor before the
                                       wait gives up (counted in counters[VGR_COUNTER_SORT_FAILURES]):
                                       0 = default (2^22); small values make give-ups likely and -1 makes
                                       every wait give up at once (both for tests of the fail-closed path) */
  int32_t ordering;                 /* VGR_ORDERING_GLOBAL_RADIX (default) or VGR_ORDERING_TILE_FRONT */
  /* Dense per-pixel path (tile_size 1, renderer_cuda.cu:1093-1690, 1764-1805).
   * Intersections are (pixel, record) pairs whose alpha at the pixel centre
   * reaches 1/255; records (and workspace.radii) are written first, then the
   * pixel intersections are emitted per record.  depth_bucket_count > 0 adds
   * the optical-thickness cutoff: records are bucketed by depth (buckets
   * uniform over [near, far]), each pixel sums a conservative optical
   * thickness (8-term -log(1 - alpha) series x 0.999, in 1/65536 fixed point
   * so the sums do not depend on the order of the adds) bucket by bucket,
   * depth_bucket_group_size buckets per pass, and its intersections behind the
   * first bucket where the sum reaches 9.2103 (T <= 1e-4) are not emitted.  The
   * compositor stops before such entries in any case, so images are identical
   * with and without the cutoff; only the intersection count shrinks. */
  int32_t depth_bucket_count;       /* 0 = no cutoff; else the bucket count (reference default 4096) */
  int32_t depth_bucket_group_size;  /* buckets per cutoff pass: 0 = default (64) */
  /* Exact-ray evaluation (reference ray_gaussian_evaluation, renderer_cuda.cu:203-245,
   * 795-948, 2180-2262): a splat's alpha at a pixel is opacity x exp(-m^2 / 2) with m
   * the Mahalanobis distance of the camera-space Gaussian's closest point on the
   * pixel's ray (0 beyond gaussian_support_sigma, capped at 1, no 1/255 floor), and
   * compositing stops after the entry that takes transmittance to 1e-4.  Tiles are
   * binned over the screen ellipse of the support quadric; with 16x16 raster tiles the
   * bins are 32x32 (four raster tiles share one sorted list; tile counts, ranges and
   * workspace tile capacities are in bins).  Needs 8x8 or 16x16 tiles, uniform
   * resolutions, workspace.ray_records, no antialiasing, the global radix ordering and
   * no training checkpoints. */
  int32_t ray_evaluation;
} vgr_options;

/* Intersection ordering of a render.
 * GLOBAL_RADIX: every intersection of the batch is sorted by (tile, depth bits)
 *   with the onesweep radix sort (the reference's DeviceRadixSort call,
 *   renderer_cuda.cu:3215-3269); intersection capacity up to 2^31 - 1.
 * TILE_FRONT: no reference counterpart.  Intersections are bucketed per tile
 *   (key = depth bits relative to near, value = packed Gaussian index), each
 *   tile sorts only the nearest entries of its bucket (at most 4096, chosen by
 *   a depth threshold) in LDS, compositing recomputes the records of the
 *   entries it reads, and tiles where a pixel is still open at the end of a
 *   truncated list are sorted in full and composited again.  Images are
 *   identical to GLOBAL_RADIX; it pays when pixels saturate early (closed
 *   scenes).  Needs all cameras active, 16x16 tiles, at most 2048 tiles per
 *   camera, no training checkpoints; capacity up to INT32_MAX; counters[4]
 *   counts non-empty tiles; records are not written (visible_capacity may be 1);
 *   sorted lists are in values_out (selector 1) and keys are not kept. */
enum vgr_ordering {
  VGR_ORDERING_GLOBAL_RADIX = 0,
  VGR_ORDERING_TILE_FRONT = 1,
};

typedef struct vgr_workspace_sizes {
  int64_t scratch_bytes;            /* bytes vgr_render needs in workspace.scratch */
  int64_t tile_capacity;            /* global tiles for the batch */
  int32_t sort_key_bits;            /* significant bits of the compressed sort key */
  int32_t sort_passes;              /* radix passes the sort will run */
} vgr_workspace_sizes;

/* Library ABI version (VGR_ABI_VERSION). */
int32_t vgr_abi_version(void);

/* Message for the most recent nonzero return on this host thread. */
const char* vgr_last_error(void);

/* 32 + bit length of (global_tile_capacity - 1); -1 when the capacity is not in
 * [1, INT32_MAX].  Same table as renderer_cuda.cu:28-46. */
int32_t vgr_emitted_sort_key_end_bit(int64_t global_tile_capacity);

/* Scratch requirement of vgr_render for the given capacities. */
int32_t vgr_workspace_sizes_query(const vgr_cameras* cameras,
                                  const vgr_options* options,
                                  int64_t scene_gaussians,
                                  int64_t intersection_capacity,
                                  vgr_workspace_sizes* out);

/* Bytes of sort scratch for n_items keys over n_segments tiles (the caller may
 * size workspace.scratch with vgr_workspace_sizes_query instead). */
int64_t vgr_sort_temp_bytes(int64_t n_items, int64_t n_segments);

/* Bounds of each chunk of 256 consecutive packed Gaussians: min xyz, max xyz,
 * an upper bound of the largest standard deviation (sqrt of the covariance
 * trace) and 0.  No reference counterpart (the reference projects every
 * camera x Gaussian pair, renderer_cuda.cu:581-591); used for frustum culling. */
int32_t vgr_chunk_bounds(const float* means, const float* covariances, int64_t count, float* bounds,
                         void* stream);

/* covariances[g] = R(q_g) diag(s_g^2) R(q_g)^T, upper triangle (renderer_cuda.cu:492-540). */
int32_t vgr_precompute_covariances(const float* scales, const float* rotations,
                                   int64_t count, float* covariances, void* stream);

/* Full forward pass: projection + culling, tile emission, radix ordering, tile
 * ranges, fused RGB/depth/alpha/semantic compositing.  *selector_out receives
 * 1 when keys_out/values_out hold the sorted intersections, 0 when keys_in /
 * values_in do (reference DoubleBuffer selector, cuda_backend.py:718-736).
 * The host is never synchronized. */
int32_t vgr_render(const vgr_scene* scene, const vgr_cameras* cameras,
                   const vgr_outputs* outputs, const vgr_workspace* workspace,
                   const vgr_options* options, void* stream, int64_t* selector_out);

/* Allocation-free variant of vgr_render.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Identical
 * to vgr_render except that workspace->scratch and workspace->scratch_bytes are
 * ignored: the scratch comes from a device array owned by the library,
 * VGR_STATIC_SCRATCH_BYTES per device, held from the first call on a device
 * until the process ends.  A render whose scratch requirement
 * (vgr_workspace_sizes_query) exceeds it fails with VGR_E_CAPACITY.  The
 * caller still passes the inputs, outputs and the other workspace buffers.
 * Calls on one device share the array: calls that overlap in time on one
 * device (two streams, two host threads) race and must be serialized by the
 * caller.  Calls nothing that allocates. */
#define VGR_STATIC_SCRATCH_BYTES (64LL << 20)
int64_t vgr_static_scratch_bytes(void);
int32_t vgr_render_static_scratch(const vgr_scene* scene, const vgr_cameras* cameras,
                                  const vgr_outputs* outputs, const vgr_workspace* workspace,
                                  const vgr_options* options, void* stream, int64_t* selector_out);

/* Gradients of one checkpointed forward pass.  Pixel-indexed inputs use the
 * forward's output layout; grad_depth / grad_alpha may be NULL (zero).  Scene
 * gradients are indexed like the packed scene arrays and are ADDED to (the
 * caller zero-fills them); grad_viewmats [B,4,4] may be NULL (no pose
 * gradient).  record_grads is workspace of 10 floats per intersection slot
 * (dL/d mean_x, mean_y, conic_0..2, opacity, depth, r, g, b).  camera_grads
 * holds VGR_POSE_REPLICAS copies of 24 pose partials per active camera: the
 * projection workgroups add into copy (workgroup % VGR_POSE_REPLICAS), so
 * that the adds of one camera spread over that many addresses. */
#define VGR_POSE_REPLICAS 32
typedef struct vgr_gradients {
  const float* grad_rgb;            /* [.,3] dL/d rgb */
  const float* grad_depth;          /* [.]   dL/d depth (ignored where depth is +inf) */
  const float* grad_alpha;          /* [.]   dL/d alpha */
  float* record_grads;              /* [intersection_capacity, 10] workspace (one row per slot, every row
                                       of a listed slot written by the pass: no zero fill needed) */
  float* grad_means;                /* [G,3] */
  float* grad_covariances;          /* [G,6] upper triangle, as scene.covariances */
  float* grad_opacities;            /* [G] */
  float* grad_features;             /* [G,feature_width]; the first 3*(d+1)^2 columns receive gradients */
  float* grad_viewmats;             /* [B,16] or NULL */
  float* camera_grads;              /* [VGR_POSE_REPLICAS * active_count, 24] workspace when grad_viewmats != NULL */
  int64_t sorted_selector;          /* the forward's *selector_out */
} vgr_gradients;

/* Backward pass of vgr_render.  Must follow the checkpointed vgr_render with
 * the same scene, cameras, options and workspace, untouched in between: it
 * re-uses the visible records, the sorted tile lists and the projection block
 * layout left there.  Compositing runs back to front from the checkpointed
 * final transmittance (no per-intersection state is stored) and stores each
 * (record, tile) pair's terms in the row of its intersection slot; projection,
 * EWA, SH and camera-pose gradients are recomputed per visible pair, which sums
 * its slots' rows (per-pair results are deterministic).  Scene gradients of a
 * Gaussian seen by several camera groups, and pose gradients, are merged with
 * float atomics (order-dependent to rounding).  Needs workspace.slot_records.
 * After a checkpointed VGR_ORDERING_TILE_FRONT forward (no slot_records) the
 * pass computes camera-pose gradients only: each tile list is walked with the
 * records recomputed from the scene, and every (Gaussian, tile) pair's terms go
 * straight through the pair's projection backward into camera_grads; scene
 * gradients and record_grads must be NULL there.
 * No reference counterpart (the reference is forward-only,
 * KERNEL_ARCHITECTURE.md:454-460). */
int32_t vgr_render_backward(const vgr_scene* scene, const vgr_cameras* cameras,
                            const vgr_outputs* outputs, const vgr_workspace* workspace,
                            const vgr_options* options, const vgr_gradients* gradients, void* stream);

/* Chain rule of vgr_precompute_covariances: from dL/d covariances to dL/d
 * scales [G,3] and dL/d rotations [G,4] (unnormalized WXYZ quaternions; the
 * normalization is differentiated).  Outputs are overwritten. */
int32_t vgr_covariance_backward(const float* scales, const float* rotations, const float* grad_covariances,
                                int64_t count, float* grad_scales, float* grad_rotations, void* stream);

/* Training-time camera poses: views[b] = exp(xi_b^) @ base_views[b] for se(3)
 * twists xi = (rho, phi) [B,6] and row-major 4x4 base views [B,16] (Rodrigues
 * rotation, left-Jacobian translation), and its backward: dL/d twists [B,6]
 * and, when grad_base_views is not NULL, dL/d base views [B,16] from
 * dL/d views.  Computed per camera in double.  No reference counterpart (the
 * reference is forward-only); it is the pose parameterisation of the
 * differentiable path. */
int32_t vgr_pose_compose(const float* twists, const float* base_views, int64_t count, float* views, void* stream);
int32_t vgr_pose_compose_backward(const float* twists, const float* base_views, const float* grad_views,
                                  int64_t count, float* grad_twists, float* grad_base_views, void* stream);

/* Logical-batch chunking helpers (renderer_cuda.cu:2307-2360). */
int32_t vgr_prepare_chunked_active_ids(const int64_t* active_ids, int64_t active_count,
                                       int32_t* active_mask, int64_t* chunk_active_ids,
                                       int64_t logical_batch, int64_t physical_batch,
                                       void* stream);
int32_t vgr_aggregate_chunk_counters(const int64_t* chunk_counters, int64_t counter_stride,
                                     int64_t chunk_count, int64_t* counters,
                                     int64_t* physical_capacity_counters, void* stream);

/* Fills n int64 values (graph-safe replacement for hipMemsetAsync). */
int32_t vgr_fill_i64(int64_t* data, int64_t n, int64_t value, void* stream);
int32_t vgr_fill_f32(float* data, int64_t n, float value, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* VGR_H */
