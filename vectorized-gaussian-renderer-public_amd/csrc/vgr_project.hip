// Projection, culling, SH colour and tile-intersection emission for gfx950.
//
// Reference behaviour: project_visible_kernel (renderer_cuda.cu:542-1091) and
// bin_tiles_kernel<false> (renderer_cuda.cu:1691-1829).  The reference appends
// visible records and intersections with global atomics, so their order is
// scheduling-dependent.  Here the pass is split in two launches over the same
// (camera group x Gaussian chunk) grid:
//   count: every block counts its visible pairs and tile intersections,
//   scan : a device-wide exclusive scan turns the counts into block bases,
//   write: every block recomputes its pairs and writes records and keys at
//          deterministic positions (camera-then-Gaussian order inside a block).
// For any fixed camera the records therefore appear in increasing Gaussian
// order, and a stable sort on (tile, depth) yields the reference's
// deterministic (tile, depth, Gaussian ID) order without a 64-bit key pass.
//
// Each thread owns one Gaussian and loops over the cameras of its group, so the
// scene attributes are read from HBM once per group instead of once per
// (camera, Gaussian) pair; camera parameters are wave-uniform scalar loads.
#include "vgr_internal.h"

namespace vgr {
namespace {

constexpr float kShC1 = 0.4886025119029199f;
constexpr float kShC2_0 = 1.0925484305920792f;
constexpr float kShC2_1 = -1.0925484305920792f;
constexpr float kShC2_2 = 0.31539156525252005f;
constexpr float kShC2_3 = -1.0925484305920792f;
constexpr float kShC2_4 = 0.5462742152960396f;
constexpr float kShC3_0 = -0.5900435899266435f;
constexpr float kShC3_1 = 2.890611442640554f;
constexpr float kShC3_2 = -0.4570457994644658f;
constexpr float kShC3_3 = 0.3731763325901154f;
constexpr float kShC3_4 = -0.4570457994644658f;
constexpr float kShC3_5 = 1.445305721320277f;
constexpr float kShC3_6 = -0.5900435899266435f;

constexpr int kScanItems = 8;
constexpr int kScanTile = kBlockThreads * kScanItems;

__device__ __forceinline__ int float_to_int_clamped(float value) {
  // ceilf/floorf results of finite splat extents; clamp keeps the int
  // conversion defined for pathological covariances.
  if (!(value < 1073741824.0f)) return 1073741824;
  if (!(value > -1073741824.0f)) return -1073741824;
  return static_cast<int>(value);
}

struct Projected {
  float mean_x, mean_y;
  float conic_0, conic_1, conic_2;
  float depth, opacity;
  int radius_x, radius_y;
};

// Screen-space projection of one (camera, Gaussian) pair.  Operation order
// follows renderer_cuda.cu:614-1055 term by term (left-to-right sums, no
// contraction), with vgr_logf in place of __logf.
__device__ __forceinline__ bool project_pair(
    const CameraParams& cam, float m0, float m1, float m2, float opacity,
    const float* cov, float near_plane, float far_plane, float support_sigma,
    float covariance_epsilon, bool antialiased, Projected& out) {
  if (!(opacity >= VGR_ALPHA_THRESHOLD)) return false;
  const float* e = cam.env;
  const float env_x = e[0] * m0 + e[1] * m1 + e[2] * m2 + e[3];
  const float env_y = e[4] * m0 + e[5] * m1 + e[6] * m2 + e[7];
  const float env_z = e[8] * m0 + e[9] * m1 + e[10] * m2 + e[11];
  const float* v = cam.view;
  const float camera_x = v[0] * env_x + v[1] * env_y + v[2] * env_z + v[3];
  const float camera_y = v[4] * env_x + v[5] * env_y + v[6] * env_z + v[7];
  const float camera_z = v[8] * env_x + v[9] * env_y + v[10] * env_z + v[11];
  if (camera_z < near_plane || camera_z > far_plane) return false;

  const float inverse_z = 1.0f / camera_z;
  const float mean_x = cam.focal_x * camera_x * inverse_z + cam.center_x;
  const float mean_y = cam.focal_y * camera_y * inverse_z + cam.center_y;

  const float* r = cam.rotation;
  float rc[9];
#pragma unroll
  for (int row = 0; row < 3; ++row) {
    const float r0 = r[row * 3 + 0];
    const float r1 = r[row * 3 + 1];
    const float r2 = r[row * 3 + 2];
    rc[row * 3 + 0] = r0 * cov[0] + r1 * cov[1] + r2 * cov[2];
    rc[row * 3 + 1] = r0 * cov[1] + r1 * cov[3] + r2 * cov[4];
    rc[row * 3 + 2] = r0 * cov[2] + r1 * cov[4] + r2 * cov[5];
  }
  const float c00 = rc[0] * r[0] + rc[1] * r[1] + rc[2] * r[2];
  const float c01 = rc[0] * r[3] + rc[1] * r[4] + rc[2] * r[5];
  const float c02 = rc[0] * r[6] + rc[1] * r[7] + rc[2] * r[8];
  const float c11 = rc[3] * r[3] + rc[4] * r[4] + rc[5] * r[5];
  const float c12 = rc[3] * r[6] + rc[4] * r[7] + rc[5] * r[8];
  const float c22 = rc[6] * r[6] + rc[7] * r[7] + rc[8] * r[8];

  const float clamped_x = camera_z * fminf(cam.limit_x_pos, fmaxf(-cam.limit_x_neg, camera_x * inverse_z));
  const float clamped_y = camera_z * fminf(cam.limit_y_pos, fmaxf(-cam.limit_y_neg, camera_y * inverse_z));
  const float inverse_z_squared = inverse_z * inverse_z;
  const float j00 = cam.focal_x * inverse_z;
  const float j02 = -cam.focal_x * clamped_x * inverse_z_squared;
  const float j11 = cam.focal_y * inverse_z;
  const float j12 = -cam.focal_y * clamped_y * inverse_z_squared;
  const float raw00 = j00 * j00 * c00 + 2.0f * j00 * j02 * c02 + j02 * j02 * c22;
  const float cov01 = j00 * j11 * c01 + j00 * j12 * c02 + j02 * j11 * c12 + j02 * j12 * c22;
  const float raw11 = j11 * j11 * c11 + 2.0f * j11 * j12 * c12 + j12 * j12 * c22;
  const float cov00 = raw00 + covariance_epsilon;
  const float cov11 = raw11 + covariance_epsilon;
  const float determinant = cov00 * cov11 - cov01 * cov01;
  if (!(determinant > 0.0f)) return false;
  if (antialiased) {
    const float original = raw00 * raw11 - cov01 * cov01;
    const float compensation = sqrtf(fmaxf(
        VGR_MIN_ANTIALIAS_COMPENSATION * VGR_MIN_ANTIALIAS_COMPENSATION,
        original / determinant));
    opacity *= compensation;
    if (!(opacity >= VGR_ALPHA_THRESHOLD)) return false;
  }
  const float opacity_extend = sqrtf(2.0f * vgr_logf(opacity / VGR_ALPHA_THRESHOLD));
  const float extend = fminf(support_sigma, opacity_extend);
  const int radius_x = float_to_int_clamped(ceilf(extend * sqrtf(cov00)));
  const int radius_y = float_to_int_clamped(ceilf(extend * sqrtf(cov11)));
  if (radius_x <= 0 && radius_y <= 0) return false;
  if (mean_x + static_cast<float>(radius_x) <= 0.0f ||
      mean_y + static_cast<float>(radius_y) <= 0.0f ||
      mean_x - static_cast<float>(radius_x) >= static_cast<float>(cam.width) ||
      mean_y - static_cast<float>(radius_y) >= static_cast<float>(cam.height)) {
    return false;
  }
  const float inverse_determinant = 1.0f / determinant;
  out.mean_x = mean_x;
  out.mean_y = mean_y;
  out.conic_0 = cov11 * inverse_determinant;
  out.conic_1 = -cov01 * inverse_determinant;
  out.conic_2 = cov00 * inverse_determinant;
  out.depth = camera_z;
  out.opacity = fminf(fmaxf(opacity, 0.0f), 1.0f);
  out.radius_x = radius_x;
  out.radius_y = radius_y;
  return true;
}

struct TileRect {
  int min_x, min_y, max_x, max_y;
};

// Tile rectangle of a projected splat (renderer_cuda.cu:1729-1748).
__device__ __forceinline__ TileRect tile_rect(const Projected& p, int tile_size, int tiles_x, int tiles_y) {
  const float ts = static_cast<float>(tile_size);
  TileRect rect;
  rect.min_x = max(0, static_cast<int>(floorf(p.mean_x / ts - static_cast<float>(p.radius_x) / ts)));
  rect.min_y = max(0, static_cast<int>(floorf(p.mean_y / ts - static_cast<float>(p.radius_y) / ts)));
  rect.max_x = min(tiles_x, static_cast<int>(ceilf(p.mean_x / ts + static_cast<float>(p.radius_x) / ts)));
  rect.max_y = min(tiles_y, static_cast<int>(ceilf(p.mean_y / ts + static_cast<float>(p.radius_y) / ts)));
  return rect;
}

__device__ __forceinline__ int rect_tiles(const TileRect& rect) {
  if (rect.min_x >= rect.max_x || rect.min_y >= rect.max_y) return 0;
  return (rect.max_x - rect.min_x) * (rect.max_y - rect.min_y);
}

// View-dependent colour: base RGB plus real-SH bands 1..degree evaluated at the
// unit direction from the camera centre to the Gaussian (3DGS basis/constants),
// clamped at zero.  f[3k + c] holds band k, channel c (held in registers).
template <int kDegree>
__device__ __forceinline__ void sh_colour(const float (&f)[3 * (kDegree + 1) * (kDegree + 1)], float dx, float dy,
                                          float dz, float* rgb) {
  constexpr int kBands = (kDegree + 1) * (kDegree + 1);
  rgb[0] = f[0];
  rgb[1] = f[1];
  rgb[2] = f[2];
  if (kDegree <= 0) return;
  const float norm2 = dx * dx + dy * dy + dz * dz;
  float x = 0.0f, y = 0.0f, z = 0.0f;
  if (norm2 > 0.0f) {
    const float inverse_norm = 1.0f / sqrtf(norm2);
    x = dx * inverse_norm;
    y = dy * inverse_norm;
    z = dz * inverse_norm;
  }
  float basis[16];
  basis[1] = -kShC1 * y;
  basis[2] = kShC1 * z;
  basis[3] = -kShC1 * x;
  if (kDegree >= 2) {
    const float xx = x * x, yy = y * y, zz = z * z;
    const float xy = x * y, yz = y * z, xz = x * z;
    basis[4] = kShC2_0 * xy;
    basis[5] = kShC2_1 * yz;
    basis[6] = kShC2_2 * (2.0f * zz - xx - yy);
    basis[7] = kShC2_3 * xz;
    basis[8] = kShC2_4 * (xx - yy);
    if (kDegree >= 3) {
      basis[9] = kShC3_0 * y * (3.0f * xx - yy);
      basis[10] = kShC3_1 * xy * z;
      basis[11] = kShC3_2 * y * (4.0f * zz - xx - yy);
      basis[12] = kShC3_3 * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
      basis[13] = kShC3_4 * x * (4.0f * zz - xx - yy);
      basis[14] = kShC3_5 * z * (xx - yy);
      basis[15] = kShC3_6 * x * (xx - 3.0f * yy);
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float acc = 0.0f;
#pragma unroll
    for (int k = 1; k < kBands; ++k) acc = fmaf(basis[k], f[3 * k + c], acc);
    rgb[c] = fmaxf(f[c] + acc, 0.0f);
  }
}

// Loads the 3*(degree+1)^2 colour channels of Gaussian g into registers, with
// 16-byte loads when the row stride allows it.
template <int kDegree>
__device__ __forceinline__ void load_features(const float* __restrict__ features, int64_t g, int width,
                                              float (&f)[3 * (kDegree + 1) * (kDegree + 1)]) {
  constexpr int kChannels = 3 * (kDegree + 1) * (kDegree + 1);
  const float* row = features + g * width;
  if (kChannels % 4 == 0 && (width & 3) == 0 && (reinterpret_cast<uintptr_t>(features) & 15u) == 0) {
    const float4* row4 = reinterpret_cast<const float4*>(row);
#pragma unroll
    for (int i = 0; i < kChannels / 4; ++i) {
      const float4 v = row4[i];
      f[4 * i + 0] = v.x;
      f[4 * i + 1] = v.y;
      f[4 * i + 2] = v.z;
      f[4 * i + 3] = v.w;
    }
  } else {
#pragma unroll
    for (int i = 0; i < kChannels; ++i) f[i] = row[i];
  }
}

__global__ void build_camera_table_kernel(
    const float* __restrict__ viewmats, const float* __restrict__ intrinsics,
    const float* __restrict__ env_xforms, const int64_t* __restrict__ camera_scene_ids,
    const int64_t* __restrict__ active_ids, int active_count, int batch,
    const int32_t* __restrict__ widths, const int32_t* __restrict__ heights,
    const int64_t* __restrict__ pixel_offsets, int default_width, int default_height,
    int tile_size, const int64_t* __restrict__ registered_scene_ids,
    const int64_t* __restrict__ scene_offsets, int scene_count,
    CameraParams* __restrict__ table) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= active_count) return;
  CameraParams p;
  const int64_t camera64 = active_ids == nullptr ? a : active_ids[a];
  const bool in_batch = camera64 >= 0 && camera64 < batch;
  const int camera = in_batch ? static_cast<int>(camera64) : 0;
  p.camera = in_batch ? camera : -1;
  const float* env = env_xforms + static_cast<int64_t>(camera) * 16;
  const float* view = viewmats + static_cast<int64_t>(camera) * 16;
  for (int i = 0; i < 12; ++i) {
    p.env[i] = env[i];
    p.view[i] = view[i];
  }
  // R_view * R_env (renderer_cuda.cu:147-159 order).
  for (int row = 0; row < 3; ++row) {
    for (int column = 0; column < 3; ++column) {
      p.rotation[row * 3 + column] = view[row * 4 + 0] * env[0 * 4 + column] +
                                     view[row * 4 + 1] * env[1 * 4 + column] +
                                     view[row * 4 + 2] * env[2 * 4 + column];
    }
  }
  const int width = widths == nullptr ? default_width : widths[camera];
  const int height = heights == nullptr ? default_height : heights[camera];
  const float* k = intrinsics + static_cast<int64_t>(camera) * 9;
  p.focal_x = k[0];
  p.focal_y = k[4];
  p.center_x = k[2];
  p.center_y = k[5];
  const float tan_fov_x = 0.5f * static_cast<float>(width) / p.focal_x;
  const float tan_fov_y = 0.5f * static_cast<float>(height) / p.focal_y;
  p.limit_x_pos = (static_cast<float>(width) - p.center_x) / p.focal_x + 0.3f * tan_fov_x;
  p.limit_x_neg = p.center_x / p.focal_x + 0.3f * tan_fov_x;
  p.limit_y_pos = (static_cast<float>(height) - p.center_y) / p.focal_y + 0.3f * tan_fov_y;
  p.limit_y_neg = p.center_y / p.focal_y + 0.3f * tan_fov_y;

  // Camera centre in scene coordinates: solve (R_view R_env) o = -(R_view t_env + t_view)
  // in double precision with the adjugate inverse.
  double m[9];
  for (int row = 0; row < 3; ++row) {
    for (int column = 0; column < 3; ++column) {
      m[row * 3 + column] = static_cast<double>(view[row * 4 + 0]) * static_cast<double>(env[0 * 4 + column]) +
                            static_cast<double>(view[row * 4 + 1]) * static_cast<double>(env[1 * 4 + column]) +
                            static_cast<double>(view[row * 4 + 2]) * static_cast<double>(env[2 * 4 + column]);
    }
  }
  double t[3];
  for (int row = 0; row < 3; ++row) {
    t[row] = static_cast<double>(view[row * 4 + 0]) * static_cast<double>(env[3]) +
             static_cast<double>(view[row * 4 + 1]) * static_cast<double>(env[7]) +
             static_cast<double>(view[row * 4 + 2]) * static_cast<double>(env[11]) +
             static_cast<double>(view[row * 4 + 3]);
  }
  const double a00 = m[4] * m[8] - m[5] * m[7];
  const double a01 = m[2] * m[7] - m[1] * m[8];
  const double a02 = m[1] * m[5] - m[2] * m[4];
  const double a10 = m[5] * m[6] - m[3] * m[8];
  const double a11 = m[0] * m[8] - m[2] * m[6];
  const double a12 = m[2] * m[3] - m[0] * m[5];
  const double a20 = m[3] * m[7] - m[4] * m[6];
  const double a21 = m[1] * m[6] - m[0] * m[7];
  const double a22 = m[0] * m[4] - m[1] * m[3];
  const double det = m[0] * a00 + m[1] * a10 + m[2] * a20;
  if (det != 0.0) {
    p.origin_x = static_cast<float>(-(a00 * t[0] + a01 * t[1] + a02 * t[2]) / det);
    p.origin_y = static_cast<float>(-(a10 * t[0] + a11 * t[1] + a12 * t[2]) / det);
    p.origin_z = static_cast<float>(-(a20 * t[0] + a21 * t[1] + a22 * t[2]) / det);
  } else {
    p.origin_x = p.origin_y = p.origin_z = 0.0f;
  }

  p.width = width;
  p.height = height;
  p.tiles_x = (width + tile_size - 1) / tile_size;
  p.tiles_y = (height + tile_size - 1) / tile_size;
  if (widths == nullptr) {
    p.tile_base = static_cast<uint32_t>(camera) * static_cast<uint32_t>(p.tiles_x * p.tiles_y);
  } else {
    uint32_t base = 0;
    for (int b = 0; b < camera; ++b) {
      base += static_cast<uint32_t>(((widths[b] + tile_size - 1) / tile_size) *
                                    ((heights[b] + tile_size - 1) / tile_size));
    }
    p.tile_base = base;
  }
  p.pixel_offset = pixel_offsets != nullptr
                       ? pixel_offsets[camera]
                       : static_cast<int64_t>(camera) * width * height;
  int slot = -1;
  if (in_batch) {
    const int64_t requested = camera_scene_ids[camera];
    for (int s = 0; s < scene_count; ++s) {
      if (registered_scene_ids[s] == requested) {
        slot = s;
        break;
      }
    }
  }
  p.scene_slot = slot;
  p.gaussian_begin = slot >= 0 ? scene_offsets[slot] : 0;
  p.gaussian_end = slot >= 0 ? scene_offsets[slot + 1] : 0;
  int64_t active_tile_begin = 0;
  for (int previous = 0; previous < a; ++previous) {
    const int64_t other64 = active_ids == nullptr ? previous : active_ids[previous];
    const int other = (other64 >= 0 && other64 < batch) ? static_cast<int>(other64) : 0;
    const int other_width = widths == nullptr ? default_width : widths[other];
    const int other_height = heights == nullptr ? default_height : heights[other];
    active_tile_begin += static_cast<int64_t>((other_width + tile_size - 1) / tile_size) *
                         ((other_height + tile_size - 1) / tile_size);
  }
  p.active_tile_begin = active_tile_begin;
  p.pad0 = 0;
  p.pad1[0] = p.pad1[1] = 0;
  table[a] = p;
}

struct ProjectionArgs {
  const float* means;
  const float* covariances;
  const float* opacities;
  const float* features;
  const CameraParams* table;
  int64_t gaussian_count;
  int64_t chunks;
  int active_count;
  int cameras_per_block;
  int feature_width;
  int sh_degree;
  int tile_size;
  int antialiased;
  float near_plane, far_plane, support_sigma, covariance_epsilon;
  const int32_t* gaussian_ids;   // registration-order id per packed entry, or null (identity)
  const int32_t* item_blocks;    // kept (group, chunk) blocks in order, or null (every block)
  const uint64_t* item_masks;    // cameras of the group that may see the chunk
  const int64_t* item_count;     // device count of kept blocks
};

// Resolves the (group, chunk) block and camera mask processed by this workgroup.
// Returns false for workgroups beyond the kept item count.
__device__ __forceinline__ bool resolve_block(const ProjectionArgs& args, int64_t& block, uint64_t& mask) {
  block = blockIdx.x;
  mask = ~0ull;
  if (args.item_blocks == nullptr) return true;
  if (block >= *args.item_count) return false;
  mask = args.item_masks[block];
  block = args.item_blocks[block];
  return true;
}

// Pair results for one camera of the group; identical in both passes.
__device__ __forceinline__ bool evaluate_pair(const ProjectionArgs& args, const CameraParams& cam, int64_t g,
                                              bool loaded, float m0, float m1, float m2, float opacity,
                                              const float* cov, Projected& p, TileRect& rect, int& tiles) {
  tiles = 0;
  if (!loaded || cam.scene_slot < 0 || g < cam.gaussian_begin || g >= cam.gaussian_end) return false;
  if (!project_pair(cam, m0, m1, m2, opacity, cov, args.near_plane, args.far_plane, args.support_sigma,
                    args.covariance_epsilon, args.antialiased != 0, p)) {
    return false;
  }
  rect = tile_rect(p, args.tile_size, cam.tiles_x, cam.tiles_y);
  tiles = rect_tiles(rect);
  return true;
}

__device__ __forceinline__ void load_gaussian(const ProjectionArgs& args, int64_t g, bool& loaded, float& m0,
                                              float& m1, float& m2, float& opacity, float* cov) {
  loaded = g < args.gaussian_count;
  m0 = m1 = m2 = opacity = 0.0f;
  for (int i = 0; i < 6; ++i) cov[i] = 0.0f;
  if (!loaded) return;
  m0 = args.means[g * 3 + 0];
  m1 = args.means[g * 3 + 1];
  m2 = args.means[g * 3 + 2];
  opacity = args.opacities[g];
  for (int i = 0; i < 6; ++i) cov[i] = args.covariances[g * 6 + i];
}

__device__ __forceinline__ int64_t wave_sum_i64(int64_t value) {
  for (int offset = 32; offset > 0; offset >>= 1) value += __shfl_xor(value, offset, kWave);
  return value;
}

__global__ __launch_bounds__(kBlockThreads) void projection_count_kernel(ProjectionArgs args,
                                                                         int64_t* __restrict__ block_counts) {
  int64_t block;
  uint64_t mask;
  if (!resolve_block(args, block, mask)) {
    if (threadIdx.x == 0) {
      block_counts[static_cast<int64_t>(blockIdx.x) * 2 + 0] = 0;
      block_counts[static_cast<int64_t>(blockIdx.x) * 2 + 1] = 0;
    }
    return;
  }
  const int group = static_cast<int>(block / args.chunks);
  const int64_t chunk = block - static_cast<int64_t>(group) * args.chunks;
  const int64_t g = chunk * kBlockThreads + threadIdx.x;
  bool loaded;
  float m0, m1, m2, opacity, cov[6];
  load_gaussian(args, g, loaded, m0, m1, m2, opacity, cov);
  int64_t visible = 0, intersections = 0;
  const int first = group * args.cameras_per_block;
  const int last = min(args.active_count, first + args.cameras_per_block);
  for (int a = first; a < last; ++a) {
    if (((mask >> (a - first)) & 1ull) == 0) continue;
    const CameraParams& cam = args.table[a];
    Projected p;
    TileRect rect;
    int tiles;
    if (evaluate_pair(args, cam, g, loaded, m0, m1, m2, opacity, cov, p, rect, tiles)) {
      visible += 1;
      intersections += tiles;
    }
  }
  __shared__ int64_t partial[2][kBlockThreads / kWave];
  visible = wave_sum_i64(visible);
  intersections = wave_sum_i64(intersections);
  const int wave = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    partial[0][wave] = visible;
    partial[1][wave] = intersections;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t v = 0, t = 0;
    for (int w = 0; w < kBlockThreads / kWave; ++w) {
      v += partial[0][w];
      t += partial[1][w];
    }
    block_counts[static_cast<int64_t>(blockIdx.x) * 2 + 0] = v;
    block_counts[static_cast<int64_t>(blockIdx.x) * 2 + 1] = t;
  }
}

// Inclusive wave scan of two int32 values.
__device__ __forceinline__ void wave_inclusive_scan2(int& a, int& b) {
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int offset = 1; offset < kWave; offset <<= 1) {
    const int ta = __shfl_up(a, offset, kWave);
    const int tb = __shfl_up(b, offset, kWave);
    if (lane >= offset) {
      a += ta;
      b += tb;
    }
  }
}

template <int kDegree>
__global__ __launch_bounds__(kBlockThreads) void projection_write_kernel(
    ProjectionArgs args, const int64_t* __restrict__ block_bases, vgr_record* __restrict__ records,
    int32_t* __restrict__ radii, uint64_t* __restrict__ keys, int32_t* __restrict__ values,
    int64_t visible_capacity, int64_t intersection_capacity) {
  int64_t block;
  uint64_t mask;
  if (!resolve_block(args, block, mask)) return;
  const int group = static_cast<int>(block / args.chunks);
  const int64_t chunk = block - static_cast<int64_t>(group) * args.chunks;
  const int64_t g = chunk * kBlockThreads + threadIdx.x;
  bool loaded;
  float m0, m1, m2, opacity, cov[6];
  load_gaussian(args, g, loaded, m0, m1, m2, opacity, cov);
  constexpr int kChannels = 3 * (kDegree + 1) * (kDegree + 1);
  float f[kChannels];
#pragma unroll
  for (int i = 0; i < kChannels; ++i) f[i] = 0.0f;
  if (loaded && opacity >= VGR_ALPHA_THRESHOLD) load_features<kDegree>(args.features, g, args.feature_width, f);
  int64_t visible_base = block_bases[static_cast<int64_t>(blockIdx.x) * 2 + 0];
  int64_t intersection_base = block_bases[static_cast<int64_t>(blockIdx.x) * 2 + 1];
  const int32_t record_id = args.gaussian_ids != nullptr && loaded ? args.gaussian_ids[g] : static_cast<int32_t>(g);
  const int first = group * args.cameras_per_block;
  const int last = min(args.active_count, first + args.cameras_per_block);
  const int wave = threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  __shared__ int wave_totals[2][kBlockThreads / kWave];
  for (int a = first; a < last; ++a) {
    if (((mask >> (a - first)) & 1ull) == 0) continue;  // uniform: whole block skips the camera
    const CameraParams& cam = args.table[a];
    Projected p;
    TileRect rect;
    int tiles;
    const bool visible = evaluate_pair(args, cam, g, loaded, m0, m1, m2, opacity, cov, p, rect, tiles);
    int inclusive_visible = visible ? 1 : 0;
    int inclusive_tiles = tiles;
    wave_inclusive_scan2(inclusive_visible, inclusive_tiles);
    if (lane == kWave - 1) {
      wave_totals[0][wave] = inclusive_visible;
      wave_totals[1][wave] = inclusive_tiles;
    }
    __syncthreads();
    int64_t prefix_visible = 0, prefix_tiles = 0, total_visible = 0, total_tiles = 0;
#pragma unroll
    for (int w = 0; w < kBlockThreads / kWave; ++w) {
      if (w < wave) {
        prefix_visible += wave_totals[0][w];
        prefix_tiles += wave_totals[1][w];
      }
      total_visible += wave_totals[0][w];
      total_tiles += wave_totals[1][w];
    }
    __syncthreads();
    if (visible) {
      const int64_t record = visible_base + prefix_visible + inclusive_visible - 1;
      const bool stored = record < visible_capacity;
      if (stored) {
        float rgb[3];
        sh_colour<kDegree>(f, m0 - cam.origin_x, m1 - cam.origin_y, m2 - cam.origin_z, rgb);
        float4* out = reinterpret_cast<float4*>(records + record);
        out[0] = make_float4(p.mean_x, p.mean_y, p.conic_0, p.conic_1);
        out[1] = make_float4(p.conic_2, p.opacity, p.depth, rgb[0]);
        out[2] = make_float4(rgb[1], rgb[2], __int_as_float(record_id), __int_as_float(cam.camera));
        if (radii != nullptr) {
          radii[record * 2 + 0] = p.radius_x;
          radii[record * 2 + 1] = p.radius_y;
        }
      }
      int64_t slot = intersection_base + prefix_tiles + inclusive_tiles - tiles;
      if (tiles > 0) {
        const uint64_t depth_bits = vgr_float_bits(p.depth);
        for (int ty = rect.min_y; ty < rect.max_y; ++ty) {
          for (int tx = rect.min_x; tx < rect.max_x; ++tx, ++slot) {
            if (slot >= intersection_capacity) continue;
            if (stored) {
              const uint32_t tile = cam.tile_base + static_cast<uint32_t>(ty * cam.tiles_x + tx);
              keys[slot] = (static_cast<uint64_t>(tile) << 32) | depth_bits;
              values[slot] = static_cast<int32_t>(record);
            } else {
              keys[slot] = kSentinelKey;
              values[slot] = -1;
            }
          }
        }
      }
    }
    visible_base += total_visible;
    intersection_base += total_tiles;
  }
}

// ---- device-wide exclusive scan of (visible, intersection) block counts ----

__global__ __launch_bounds__(kBlockThreads) void scan_reduce_kernel(const int64_t* __restrict__ counts, int64_t n,
                                                                    int64_t* __restrict__ partials) {
  const int64_t start = static_cast<int64_t>(blockIdx.x) * kScanTile;
  int64_t a = 0, b = 0;
  for (int i = 0; i < kScanItems; ++i) {
    const int64_t index = start + i * kBlockThreads + threadIdx.x;
    if (index < n) {
      a += counts[index * 2 + 0];
      b += counts[index * 2 + 1];
    }
  }
  a = wave_sum_i64(a);
  b = wave_sum_i64(b);
  __shared__ int64_t partial[2][kBlockThreads / kWave];
  const int wave = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    partial[0][wave] = a;
    partial[1][wave] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t sa = 0, sb = 0;
    for (int w = 0; w < kBlockThreads / kWave; ++w) {
      sa += partial[0][w];
      sb += partial[1][w];
    }
    partials[blockIdx.x * 2 + 0] = sa;
    partials[blockIdx.x * 2 + 1] = sb;
  }
}

__device__ __forceinline__ void wave_inclusive_scan2_i64(int64_t& a, int64_t& b) {
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int offset = 1; offset < kWave; offset <<= 1) {
    const int64_t ta = __shfl_up(a, offset, kWave);
    const int64_t tb = __shfl_up(b, offset, kWave);
    if (lane >= offset) {
      a += ta;
      b += tb;
    }
  }
}

// Block-wide exclusive scan of int64 pairs; returns the block totals.
__device__ __forceinline__ void block_exclusive_scan2_i64(int64_t a, int64_t b, int64_t& ea, int64_t& eb,
                                                          int64_t& ta, int64_t& tb) {
  __shared__ int64_t totals[2][kBlockThreads / kWave];
  int64_t ia = a, ib = b;
  wave_inclusive_scan2_i64(ia, ib);
  const int wave = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == kWave - 1) {
    totals[0][wave] = ia;
    totals[1][wave] = ib;
  }
  __syncthreads();
  int64_t pa = 0, pb = 0;
  ta = 0;
  tb = 0;
  for (int w = 0; w < kBlockThreads / kWave; ++w) {
    if (w < wave) {
      pa += totals[0][w];
      pb += totals[1][w];
    }
    ta += totals[0][w];
    tb += totals[1][w];
  }
  __syncthreads();
  ea = pa + ia - a;
  eb = pb + ib - b;
}

// Single block: exclusive scan of the partials.  Writes the totals to
// `totals` when given, and the counter block when `counters` is given.
__global__ __launch_bounds__(kBlockThreads) void scan_partials_kernel(int64_t* __restrict__ partials, int64_t n,
                                                                      int64_t visible_capacity,
                                                                      int64_t intersection_capacity,
                                                                      int64_t* __restrict__ counters,
                                                                      int64_t* __restrict__ totals) {
  int64_t carry_a = 0, carry_b = 0;
  for (int64_t start = 0; start < n; start += kBlockThreads) {
    const int64_t index = start + threadIdx.x;
    const int64_t a = index < n ? partials[index * 2 + 0] : 0;
    const int64_t b = index < n ? partials[index * 2 + 1] : 0;
    int64_t ea, eb, ta, tb;
    block_exclusive_scan2_i64(a, b, ea, eb, ta, tb);
    if (index < n) {
      partials[index * 2 + 0] = carry_a + ea;
      partials[index * 2 + 1] = carry_b + eb;
    }
    carry_a += ta;
    carry_b += tb;
  }
  if (threadIdx.x == 0) {
    if (totals != nullptr) {
      totals[0] = carry_a;
      totals[1] = carry_b;
    }
    if (counters != nullptr) {
      counters[VGR_COUNTER_VISIBLE] = carry_a;
      counters[VGR_COUNTER_INTERSECTIONS] = carry_b;
      counters[VGR_COUNTER_VISIBLE_OVERFLOW] = carry_a > visible_capacity ? carry_a - visible_capacity : 0;
      counters[VGR_COUNTER_INTERSECTION_OVERFLOW] =
          carry_b > intersection_capacity ? carry_b - intersection_capacity : 0;
      counters[VGR_COUNTER_ACTIVE_TILES] = 0;
    }
  }
}

__global__ __launch_bounds__(kBlockThreads) void scan_downsweep_kernel(const int64_t* __restrict__ counts, int64_t n,
                                                                       const int64_t* __restrict__ partials,
                                                                       int64_t* __restrict__ bases) {
  const int64_t start = static_cast<int64_t>(blockIdx.x) * kScanTile;
  int64_t carry_a = partials[blockIdx.x * 2 + 0];
  int64_t carry_b = partials[blockIdx.x * 2 + 1];
  for (int i = 0; i < kScanItems; ++i) {
    const int64_t index = start + i * kBlockThreads + threadIdx.x;
    const int64_t a = index < n ? counts[index * 2 + 0] : 0;
    const int64_t b = index < n ? counts[index * 2 + 1] : 0;
    int64_t ea, eb, ta, tb;
    block_exclusive_scan2_i64(a, b, ea, eb, ta, tb);
    if (index < n) {
      bases[index * 2 + 0] = carry_a + ea;
      bases[index * 2 + 1] = carry_b + eb;
    }
    carry_a += ta;
    carry_b += tb;
  }
}

// Per-chunk bounds: min/max of the means and the largest sqrt(trace(cov)),
// which bounds every standard deviation of every Gaussian in the chunk.
__global__ __launch_bounds__(kBlockThreads) void chunk_bounds_kernel(const float* __restrict__ means,
                                                                     const float* __restrict__ covariances,
                                                                     int64_t count, float* __restrict__ bounds) {
  const int64_t g = static_cast<int64_t>(blockIdx.x) * kBlockThreads + threadIdx.x;
  float lo[3], hi[3], extent = 0.0f;
  if (g < count) {
    for (int i = 0; i < 3; ++i) lo[i] = hi[i] = means[g * 3 + i];
    const float* c = covariances + g * 6;
    extent = sqrtf(fmaxf(c[0] + c[3] + c[5], 0.0f));
  } else {
    for (int i = 0; i < 3; ++i) {
      lo[i] = __int_as_float(0x7f800000);
      hi[i] = -__int_as_float(0x7f800000);
    }
  }
  for (int offset = 32; offset > 0; offset >>= 1) {
    for (int i = 0; i < 3; ++i) {
      lo[i] = fminf(lo[i], __shfl_xor(lo[i], offset, kWave));
      hi[i] = fmaxf(hi[i], __shfl_xor(hi[i], offset, kWave));
    }
    extent = fmaxf(extent, __shfl_xor(extent, offset, kWave));
  }
  __shared__ float partial[kBlockThreads / kWave][7];
  const int wave = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    for (int i = 0; i < 3; ++i) {
      partial[wave][i] = lo[i];
      partial[wave][3 + i] = hi[i];
    }
    partial[wave][6] = extent;
  }
  __syncthreads();
  if (threadIdx.x < 8) {
    float value = 0.0f;
    if (threadIdx.x < 7) {
      value = partial[0][threadIdx.x];
      for (int w = 1; w < kBlockThreads / kWave; ++w) {
        const float other = partial[w][threadIdx.x];
        value = threadIdx.x < 3 ? fminf(value, other) : fmaxf(value, other);
      }
    }
    bounds[static_cast<int64_t>(blockIdx.x) * 8 + threadIdx.x] = value;
  }
}

// Conservative visibility of a chunk for one camera.  A Gaussian can only be
// emitted if its centre c (camera space) satisfies z in [near, far] and, for
// each image side, f*x/z + cx + r_x > 0 etc. with the splat radius bounded by
// r_x <= sigma*(s*(f/z)*sqrt(1 + lim^2)*|W| + sqrt(eps)) + 1, where s bounds the
// standard deviation, lim the Jacobian clamp and |W| the norm of the camera
// rotation.  Multiplying by z/f > 0 turns every side condition into a half
// space x + a z + b > 0, tested against a ball around the chunk.
__device__ __forceinline__ bool chunk_may_be_visible(const CameraParams& cam, const float* bounds,
                                                     float near_plane, float far_plane, float support_sigma,
                                                     float covariance_epsilon) {
  if (!(bounds[0] <= bounds[3])) return false;  // empty chunk
  const double cx_s = 0.5 * (static_cast<double>(bounds[0]) + bounds[3]);
  const double cy_s = 0.5 * (static_cast<double>(bounds[1]) + bounds[4]);
  const double cz_s = 0.5 * (static_cast<double>(bounds[2]) + bounds[5]);
  const double hx = 0.5 * (static_cast<double>(bounds[3]) - bounds[0]);
  const double hy = 0.5 * (static_cast<double>(bounds[4]) - bounds[1]);
  const double hz = 0.5 * (static_cast<double>(bounds[5]) - bounds[2]);
  const double half_diagonal = sqrt(hx * hx + hy * hy + hz * hz);
  // Scene -> camera: env then view affine; Frobenius norm bounds the scaling.
  const float* e = cam.env;
  const float* v = cam.view;
  const double wx = e[0] * cx_s + e[1] * cy_s + e[2] * cz_s + e[3];
  const double wy = e[4] * cx_s + e[5] * cy_s + e[6] * cz_s + e[7];
  const double wz = e[8] * cx_s + e[9] * cy_s + e[10] * cz_s + e[11];
  const double x = v[0] * wx + v[1] * wy + v[2] * wz + v[3];
  const double y = v[4] * wx + v[5] * wy + v[6] * wz + v[7];
  const double z = v[8] * wx + v[9] * wy + v[10] * wz + v[11];
  double norm2 = 0.0;
  for (int i = 0; i < 9; ++i) norm2 += static_cast<double>(cam.rotation[i]) * cam.rotation[i];
  const double scale = sqrt(norm2);
  const double radius = half_diagonal * scale * 1.001 + 1.0e-5;
  if (z + radius < near_plane || z - radius > far_plane) return false;
  const double sigma = support_sigma;
  const double margin = sigma * sqrt(static_cast<double>(covariance_epsilon)) + 2.0;
  const double extent = bounds[6];
  const double lim_x = fmax(cam.limit_x_pos, cam.limit_x_neg);
  const double lim_y = fmax(cam.limit_y_pos, cam.limit_y_neg);
  const double bx = sigma * extent * sqrt(1.0 + lim_x * lim_x) * scale * 1.001 + 1.0e-5;
  const double by = sigma * extent * sqrt(1.0 + lim_y * lim_y) * scale * 1.001 + 1.0e-5;
  const double fx = cam.focal_x, fy = cam.focal_y;
  // Half spaces: n.(x,y,z) + b > 0 with n = (+-1, 0, a) or (0, +-1, a).
  const double a_left = (cam.center_x + margin) / fx;
  const double a_right = (cam.width - cam.center_x + margin) / fx;
  const double a_top = (cam.center_y + margin) / fy;
  const double a_bottom = (cam.height - cam.center_y + margin) / fy;
  if (x + a_left * z + radius * sqrt(1.0 + a_left * a_left) + bx <= 0.0) return false;
  if (-x + a_right * z + radius * sqrt(1.0 + a_right * a_right) + bx <= 0.0) return false;
  if (y + a_top * z + radius * sqrt(1.0 + a_top * a_top) + by <= 0.0) return false;
  if (-y + a_bottom * z + radius * sqrt(1.0 + a_bottom * a_bottom) + by <= 0.0) return false;
  return true;
}

// One thread per (camera group, chunk) candidate: camera mask of the group's
// cameras that may see the chunk, and a 0/1 keep flag for the ordered compaction.
__global__ __launch_bounds__(kBlockThreads) void cull_chunks_kernel(ProjectionArgs args,
                                                                    const float* __restrict__ chunk_bounds,
                                                                    int64_t blocks, int64_t* __restrict__ flags,
                                                                    uint64_t* __restrict__ masks) {
  const int64_t block = static_cast<int64_t>(blockIdx.x) * kBlockThreads + threadIdx.x;
  if (block >= blocks) return;
  const int group = static_cast<int>(block / args.chunks);
  const int64_t chunk = block - static_cast<int64_t>(group) * args.chunks;
  const float* bounds = chunk_bounds + chunk * 8;
  const int64_t first_gaussian = chunk * kBlockThreads;
  const int64_t last_gaussian = first_gaussian + kBlockThreads;
  const int first = group * args.cameras_per_block;
  const int last = min(args.active_count, first + args.cameras_per_block);
  uint64_t mask = 0;
  for (int a = first; a < last; ++a) {
    const CameraParams& cam = args.table[a];
    if (cam.scene_slot < 0 || cam.gaussian_end <= first_gaussian || cam.gaussian_begin >= last_gaussian) continue;
    if (chunk_may_be_visible(cam, bounds, args.near_plane, args.far_plane, args.support_sigma,
                             args.covariance_epsilon)) {
      mask |= 1ull << (a - first);
    }
  }
  masks[block] = mask;
  flags[block * 2 + 0] = mask != 0 ? 1 : 0;
  flags[block * 2 + 1] = 0;
}

__global__ __launch_bounds__(kBlockThreads) void compact_items_kernel(const int64_t* __restrict__ flags,
                                                                      const int64_t* __restrict__ positions,
                                                                      const uint64_t* __restrict__ masks,
                                                                      int64_t blocks, int32_t* __restrict__ items,
                                                                      uint64_t* __restrict__ item_masks) {
  const int64_t block = static_cast<int64_t>(blockIdx.x) * kBlockThreads + threadIdx.x;
  if (block >= blocks || flags[block * 2] == 0) return;
  const int64_t position = positions[block * 2];
  items[position] = static_cast<int32_t>(block);
  item_masks[position] = masks[block];
}

__global__ void precompute_covariances_kernel(const float* __restrict__ scales, const float* __restrict__ rotations,
                                              int64_t count, float* __restrict__ covariances) {
  const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= count) return;
  float w = rotations[g * 4 + 0];
  float x = rotations[g * 4 + 1];
  float y = rotations[g * 4 + 2];
  float z = rotations[g * 4 + 3];
  // renderer_cuda.cu:112-143 with a correctly rounded 1/sqrt in place of rsqrtf.
  const float norm = 1.0f / sqrtf(fmaxf(w * w + x * x + y * y + z * z, 1.0e-20f));
  w *= norm;
  x *= norm;
  y *= norm;
  z *= norm;
  const float xx = x * x, yy = y * y, zz = z * z;
  const float xy = x * y, xz = x * z, yz = y * z;
  const float wx = w * x, wy = w * y, wz = w * z;
  float r[9];
  r[0] = 1.0f - 2.0f * (yy + zz);
  r[1] = 2.0f * (xy - wz);
  r[2] = 2.0f * (xz + wy);
  r[3] = 2.0f * (xy + wz);
  r[4] = 1.0f - 2.0f * (xx + zz);
  r[5] = 2.0f * (yz - wx);
  r[6] = 2.0f * (xz - wy);
  r[7] = 2.0f * (yz + wx);
  r[8] = 1.0f - 2.0f * (xx + yy);
  const float sx = scales[g * 3 + 0] * scales[g * 3 + 0];
  const float sy = scales[g * 3 + 1] * scales[g * 3 + 1];
  const float sz = scales[g * 3 + 2] * scales[g * 3 + 2];
  float* c = covariances + g * 6;
  c[0] = r[0] * r[0] * sx + r[1] * r[1] * sy + r[2] * r[2] * sz;
  c[1] = r[0] * r[3] * sx + r[1] * r[4] * sy + r[2] * r[5] * sz;
  c[2] = r[0] * r[6] * sx + r[1] * r[7] * sy + r[2] * r[8] * sz;
  c[3] = r[3] * r[3] * sx + r[4] * r[4] * sy + r[5] * r[5] * sz;
  c[4] = r[3] * r[6] * sx + r[4] * r[7] * sy + r[5] * r[8] * sz;
  c[5] = r[6] * r[6] * sx + r[7] * r[7] * sy + r[8] * r[8] * sz;
}

}  // namespace

int launch_camera_table(const vgr_scene& scene, const vgr_cameras& cameras, const vgr_options& options,
                        CameraParams* table, hipStream_t stream) {
  if (cameras.active_count <= 0) return 0;
  const int blocks = (cameras.active_count + 63) / 64;
  hipLaunchKernelGGL(build_camera_table_kernel, dim3(blocks), dim3(64), 0, stream, cameras.viewmats,
                     cameras.intrinsics, cameras.env_xforms, cameras.scene_ids, cameras.active_ids,
                     cameras.active_count, cameras.batch, cameras.widths, cameras.heights, cameras.pixel_offsets,
                     options.width, options.height, options.tile_size, scene.registered_scene_ids,
                     scene.scene_offsets, scene.scene_count, table);
  const hipError_t status = hipGetLastError();
  if (status != hipSuccess) {
    set_error("camera table launch failed: %s", hipGetErrorString(status));
    return VGR_E_LAUNCH;
  }
  return 0;
}

// Device-wide exclusive scan of n int64 pairs (3 launches); partials hold
// 2 * ceil(n / 2048) entries.  Totals go to `totals` and/or the counter block.
void scan_pairs(const int64_t* counts, int64_t n, int64_t* bases, int64_t* partials, int64_t visible_capacity,
                int64_t intersection_capacity, int64_t* counters, int64_t* totals, hipStream_t stream) {
  const int64_t scan_blocks = (n + kScanTile - 1) / kScanTile;
  if (n > 0) {
    hipLaunchKernelGGL(scan_reduce_kernel, dim3(static_cast<unsigned>(scan_blocks)), dim3(kBlockThreads), 0, stream,
                       counts, n, partials);
  }
  hipLaunchKernelGGL(scan_partials_kernel, dim3(1), dim3(kBlockThreads), 0, stream, partials, scan_blocks,
                     visible_capacity, intersection_capacity, counters, totals);
  if (n > 0) {
    hipLaunchKernelGGL(scan_downsweep_kernel, dim3(static_cast<unsigned>(scan_blocks)), dim3(kBlockThreads), 0,
                       stream, counts, n, partials, bases);
  }
}

int launch_projection(const vgr_scene& scene, const vgr_cameras& cameras, const vgr_options& options,
                      const vgr_workspace& ws, const ScratchLayout& layout, char* scratch, hipStream_t stream) {
  ProjectionArgs args;
  args.means = scene.means;
  args.covariances = scene.covariances;
  args.opacities = scene.opacities;
  args.features = scene.features;
  args.table = reinterpret_cast<const CameraParams*>(scratch + layout.camera_table);
  args.gaussian_count = scene.gaussian_count;
  args.chunks = layout.projection_chunks;
  args.active_count = cameras.active_count;
  args.cameras_per_block = layout.cameras_per_block;
  args.feature_width = scene.feature_width;
  args.sh_degree = scene.sh_degree;
  args.tile_size = options.tile_size;
  args.antialiased = options.antialiased;
  args.near_plane = options.near_plane;
  args.far_plane = options.far_plane;
  args.support_sigma = options.gaussian_support_sigma;
  args.covariance_epsilon = options.covariance_epsilon;
  args.gaussian_ids = scene.gaussian_ids;
  args.item_blocks = nullptr;
  args.item_masks = nullptr;
  args.item_count = nullptr;

  int64_t* block_counts = reinterpret_cast<int64_t*>(scratch + layout.block_counts);
  int64_t* block_bases = reinterpret_cast<int64_t*>(scratch + layout.block_bases);
  const int64_t blocks = layout.projection_blocks;
  // Partials live in the tail of block_bases' region, which holds 2*(blocks + scan_blocks) entries.
  int64_t* partials = block_bases + 2 * blocks;

  if (blocks > 0 && layout.culling && scene.chunk_bounds != nullptr) {
    int64_t* cull_flags = reinterpret_cast<int64_t*>(scratch + layout.cull_counts);
    int64_t* cull_positions = reinterpret_cast<int64_t*>(scratch + layout.cull_bases);
    uint64_t* cull_masks = reinterpret_cast<uint64_t*>(scratch + layout.cull_masks);
    int32_t* items = reinterpret_cast<int32_t*>(scratch + layout.item_blocks);
    uint64_t* item_masks = reinterpret_cast<uint64_t*>(scratch + layout.item_masks);
    int64_t* item_totals = reinterpret_cast<int64_t*>(scratch + layout.item_totals);
    const unsigned candidate_grid = static_cast<unsigned>((blocks + kBlockThreads - 1) / kBlockThreads);
    hipLaunchKernelGGL(cull_chunks_kernel, dim3(candidate_grid), dim3(kBlockThreads), 0, stream, args,
                       scene.chunk_bounds, blocks, cull_flags, cull_masks);
    scan_pairs(cull_flags, blocks, cull_positions, cull_positions + 2 * blocks, 0, 0, nullptr, item_totals, stream);
    hipLaunchKernelGGL(compact_items_kernel, dim3(candidate_grid), dim3(kBlockThreads), 0, stream, cull_flags,
                       cull_positions, cull_masks, blocks, items, item_masks);
    args.item_blocks = items;
    args.item_masks = item_masks;
    args.item_count = item_totals;
    if (options.statistics != nullptr) launch_copy_i64(item_totals, options.statistics + 2, stream);
  }

  if (blocks > 0) {
    hipLaunchKernelGGL(projection_count_kernel, dim3(static_cast<unsigned>(blocks)), dim3(kBlockThreads), 0,
                       stream, args, block_counts);
  }
  scan_pairs(block_counts, blocks, block_bases, partials, ws.visible_capacity, ws.intersection_capacity,
             ws.counters, nullptr, stream);
  if (blocks > 0) {
    const dim3 grid(static_cast<unsigned>(blocks)), threads(kBlockThreads);
    switch (args.sh_degree) {
      case 1:
        hipLaunchKernelGGL(projection_write_kernel<1>, grid, threads, 0, stream, args, block_bases, ws.records,
                           ws.radii, ws.keys_in, ws.values_in, ws.visible_capacity, ws.intersection_capacity);
        break;
      case 2:
        hipLaunchKernelGGL(projection_write_kernel<2>, grid, threads, 0, stream, args, block_bases, ws.records,
                           ws.radii, ws.keys_in, ws.values_in, ws.visible_capacity, ws.intersection_capacity);
        break;
      case 3:
        hipLaunchKernelGGL(projection_write_kernel<3>, grid, threads, 0, stream, args, block_bases, ws.records,
                           ws.radii, ws.keys_in, ws.values_in, ws.visible_capacity, ws.intersection_capacity);
        break;
      default:
        hipLaunchKernelGGL(projection_write_kernel<0>, grid, threads, 0, stream, args, block_bases, ws.records,
                           ws.radii, ws.keys_in, ws.values_in, ws.visible_capacity, ws.intersection_capacity);
        break;
    }
  }
  const hipError_t status = hipGetLastError();
  if (status != hipSuccess) {
    set_error("projection launch failed: %s", hipGetErrorString(status));
    return VGR_E_LAUNCH;
  }
  return 0;
}

}  // namespace vgr

extern "C" int32_t vgr_precompute_covariances(const float* scales, const float* rotations, int64_t count,
                                              float* covariances, void* stream) {
  if (count < 0 || (count > 0 && (scales == nullptr || rotations == nullptr || covariances == nullptr))) {
    vgr::set_error("vgr_precompute_covariances: invalid arguments (count=%lld)", static_cast<long long>(count));
    return VGR_E_INVALID_ARGUMENT;
  }
  if (count == 0) return 0;
  const int64_t blocks = (count + vgr::kBlockThreads - 1) / vgr::kBlockThreads;
  if (blocks > 0x7fffffff) {
    vgr::set_error("vgr_precompute_covariances: grid exceeds the launch limit");
    return VGR_E_INVALID_ARGUMENT;
  }
  hipLaunchKernelGGL(vgr::precompute_covariances_kernel, dim3(static_cast<unsigned>(blocks)),
                     dim3(vgr::kBlockThreads), 0, static_cast<hipStream_t>(stream), scales, rotations, count,
                     covariances);
  const hipError_t status = hipGetLastError();
  if (status != hipSuccess) {
    vgr::set_error("covariance launch failed: %s", hipGetErrorString(status));
    return VGR_E_LAUNCH;
  }
  return 0;
}

extern "C" int32_t vgr_chunk_bounds(const float* means, const float* covariances, int64_t count, float* bounds,
                                    void* stream) {
  if (count < 0 || (count > 0 && (means == nullptr || covariances == nullptr || bounds == nullptr))) {
    vgr::set_error("vgr_chunk_bounds: invalid arguments (count=%lld)", static_cast<long long>(count));
    return VGR_E_INVALID_ARGUMENT;
  }
  if (count == 0) return 0;
  const int64_t chunks = (count + vgr::kBlockThreads - 1) / vgr::kBlockThreads;
  hipLaunchKernelGGL(vgr::chunk_bounds_kernel, dim3(static_cast<unsigned>(chunks)), dim3(vgr::kBlockThreads), 0,
                     static_cast<hipStream_t>(stream), means, covariances, count, bounds);
  const hipError_t status = hipGetLastError();
  if (status != hipSuccess) {
    vgr::set_error("chunk bounds launch failed: %s", hipGetErrorString(status));
    return VGR_E_LAUNCH;
  }
  return 0;
}
