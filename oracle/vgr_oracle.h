/*
 * CPU oracle of the batched Gaussian-splat camera path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP renderer:
 * only the test suite, the smoke entry point and the CPU-baseline leg of the
 * benchmark may load it.  The product path never calls it.
 *
 * It restates, in plain C99, the reference renderer's forward algorithm:
 *   covariance precompute  renderer_cuda.cu:492-540 (quaternion 112-143)
 *   projection + culling   renderer_cuda.cu:542-1091 (screen-space branch)
 *   tile emission          renderer_cuda.cu:1691-1829 (tile_size > 1 branch)
 *   ordering               (tile, depth bits, Gaussian ID) as the deterministic
 *                          path renderer_cuda.cu:1831-1971, 3288-3381
 *   tile ranges            renderer_cuda.cu:1908-1942
 *   compositing            renderer_cuda.cu:1973-2305 (screen-space, FULL)
 *   sRGB                   renderer_cuda.cu:161-166
 * plus the SH colour rule this build adds (3DGS real-SH basis, degree <= 3).
 * Elementary functions come from vgr_math.h (exp/log built from IEEE basic
 * operations); everything else is written independently of the HIP sources.
 */
#ifndef VGR_ORACLE_H
#define VGR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_scene {
  const float* means;
  const float* covariances;
  const float* opacities;
  const float* features;
  const int64_t* semantic_ids;
  const int64_t* registered_scene_ids;
  const int64_t* scene_offsets;
  int64_t gaussian_count;
  int32_t feature_width;
  int32_t scene_count;
  int32_t sh_degree;
  int32_t reserved;
} oracle_scene;

typedef struct oracle_cameras {
  const float* viewmats;
  const float* intrinsics;
  const float* env_xforms;
  const int64_t* scene_ids;
  const int64_t* active_ids; /* NULL = all */
  int32_t batch;
  int32_t active_count;
  const int32_t* widths;     /* NULL = uniform */
  const int32_t* heights;
  const int64_t* pixel_offsets;
} oracle_cameras;

typedef struct oracle_options {
  int32_t height;
  int32_t width;
  float near_plane;
  float far_plane;
  float gaussian_support_sigma;
  float covariance_epsilon;
  float semantic_min_alpha;
  int32_t antialiased;
  int32_t tile_size;
  int32_t output_srgb;
  int32_t full_sensor_output;
  int32_t threads;           /* 0 = all online CPUs */
  /* tile_size 1 only: depth buckets of the optical-thickness cutoff (0 = none)
   * and buckets per pass (accepted; the cutoff does not depend on it). */
  int32_t depth_bucket_count;
  int32_t depth_bucket_group_size;
  /* Exact-ray evaluation (8x8 or 16x16 raster tiles; 16x16 bins in 32x32 tiles). */
  int32_t ray_evaluation;
} oracle_options;

typedef struct oracle_outputs {
  float* rgb;
  float* depth;
  float* alpha;
  int64_t* semantic;
  int64_t counters[5];       /* visible, intersections, 0, 0, active tiles */
  /* Optional diagnostics (NULL to skip). */
  int32_t* tile_counts;      /* [global tiles] intersections per tile */
  uint64_t* sorted_keys;     /* [sorted_capacity] (tile << 32 | depth bits) in sorted order */
  int32_t* sorted_gaussians; /* [sorted_capacity] Gaussian ID of each sorted intersection */
  int64_t sorted_capacity;
  int64_t global_tiles;      /* size of tile_counts */
} oracle_outputs;

int32_t oracle_precompute_covariances(const float* scales, const float* rotations, int64_t count,
                                      float* covariances);
int32_t oracle_render(const oracle_scene* scene, const oracle_cameras* cameras,
                      const oracle_options* options, oracle_outputs* outputs);
/* Elementary functions, exported for accuracy tests. */
float oracle_expf(float x);
float oracle_expf_nonpositive(float x);
float oracle_logf(float x);
float oracle_srgb(float x);
/* View-dependent colour of one Gaussian seen from origin (scene coordinates). */
void oracle_sh_colour(const float* features, int32_t degree, float dx, float dy, float dz, float* rgb);
const char* oracle_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
