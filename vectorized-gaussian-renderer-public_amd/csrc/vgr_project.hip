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
#include "vgr_project_device.h"

namespace vgr {
namespace {

using namespace dev;

constexpr int kScanItems = 8;
constexpr int kScanTile = kBlockThreads * kScanItems;

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


// kMode: 0 tile rectangles, 1 dense per-pixel passes, 2 exact-ray extents.
template <int kMode>
__global__ __launch_bounds__(kBlockThreads) void projection_count_kernel(ProjectionArgs args,
                                                                         int64_t* __restrict__ block_counts,
                                                                         uint64_t* __restrict__ wave_masks) {
  int64_t block;
  uint64_t mask;
  const int64_t wave_slot = static_cast<int64_t>(blockIdx.x) * kProjectionWaves + threadIdx.x / kWave;
  if (!resolve_block(args, block, mask)) {
    if ((threadIdx.x & (kWave - 1)) == 0) {
      block_counts[wave_slot * 2 + 0] = 0;
      block_counts[wave_slot * 2 + 1] = 0;
      wave_masks[wave_slot] = 0;
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
  uint64_t wave_mask = 0;  // cameras with at least one visible pair among the wave's Gaussians
  const int first = group * args.cameras_per_block;
  const int last = min(args.active_count, first + args.cameras_per_block);
  for (int a = first; a < last; ++a) {
    if (((mask >> (a - first)) & 1ull) == 0) continue;
    const CameraParams& cam = args.table[a];
    Projected p;
    TileRect rect;
    int tiles;
    bool seen;
    if constexpr (kMode == 2) {
      RayData ray;
      seen = evaluate_ray_pair(args, cam, g, loaded, m0, m1, m2, opacity, cov, p, rect, tiles, ray);
    } else {
      seen = evaluate_pair(args, cam, g, loaded, m0, m1, m2, opacity, cov, p, rect, tiles);
    }
    if constexpr (kMode == 1) {
      // Pixel intersections, counted by the wave together (added once, on lane 0).
      tiles = 0;
      if (args.dense_mode != 1) {
        const int bucket = seen && args.depth_cutoff != nullptr
                               ? depth_bucket(p.depth, args.near_plane, args.far_plane, args.depth_buckets)
                               : 0;
        // Splats behind the cutoff of every pixel they reach are not walked (the
        // write pass applies the same test, so the counts and slots agree).
        const bool walk = seen && dense_rect_reaches(rect, cam.tile_base, cam.tiles_x, bucket, args.dense_block_cutoff);
        const int wave_pixels = dense_wave_pixels(walk, p, rect, cam.tile_base, cam.tiles_x, bucket,
                                                  args.depth_cutoff, [](bool, int, int, uint32_t, float) {});
        if ((threadIdx.x & (kWave - 1)) == 0) intersections += wave_pixels;
      }
    }
    if (seen) {
      visible += 1;
      intersections += tiles;
    }
    if (__ballot(seen) != 0ull) wave_mask |= 1ull << (a - first);
  }
  // Counts per wave: each wave of the write kernel then owns its output range and
  // runs without block barriers.
  visible = wave_sum_i64(visible);
  intersections = wave_sum_i64(intersections);
  if ((threadIdx.x & (kWave - 1)) == 0) {
    block_counts[wave_slot * 2 + 0] = visible;
    block_counts[wave_slot * 2 + 1] = intersections;
    wave_masks[wave_slot] = wave_mask;
  }
}

template <int kDegree, int kMode>
__global__ __launch_bounds__(kBlockThreads, 4) void projection_write_kernel(
    ProjectionArgs args, const int64_t* __restrict__ block_bases, const uint64_t* __restrict__ wave_masks,
    vgr_record* __restrict__ records, int32_t* __restrict__ radii, uint64_t* __restrict__ keys,
    int32_t* __restrict__ values, int32_t* __restrict__ slot_records, int64_t visible_capacity,
    int64_t intersection_capacity, float* __restrict__ ray_records) {
  constexpr bool kDense = kMode == 1;
  int64_t block;
  uint64_t mask;
  if (!resolve_block(args, block, mask)) return;
  const int group = static_cast<int>(block / args.chunks);
  const int64_t chunk = block - static_cast<int64_t>(group) * args.chunks;
  const int64_t g = chunk * kBlockThreads + threadIdx.x;
  const int lane = threadIdx.x & (kWave - 1);
  bool loaded;
  float m0, m1, m2, opacity, cov[6];
  load_gaussian(args, g, loaded, m0, m1, m2, opacity, cov);
  constexpr int kChannels = 3 * (kDegree + 1) * (kDegree + 1);
  float f[kChannels];
#pragma unroll
  for (int i = 0; i < kChannels; ++i) f[i] = 0.0f;
  if (loaded && opacity >= VGR_ALPHA_THRESHOLD) load_features<kDegree>(args.features, g, args.feature_width, f);
  const int64_t wave_slot = static_cast<int64_t>(blockIdx.x) * kProjectionWaves + threadIdx.x / kWave;
  int64_t visible_base = block_bases[wave_slot * 2 + 0];
  int64_t intersection_base = block_bases[wave_slot * 2 + 1];
  // Cameras that see none of the wave's Gaussians (per the count pass) are skipped
  // whole: they add no records and no intersections.
  const uint64_t wave_mask = wave_masks[wave_slot];
  const int32_t record_id = args.gaussian_ids != nullptr && loaded ? args.gaussian_ids[g] : static_cast<int32_t>(g);
  const int first = group * args.cameras_per_block;
  const int last = min(args.active_count, first + args.cameras_per_block);
  // Each wave owns its scanned output range, so the waves never synchronize.
  for (int a = first; a < last; ++a) {
    if (((mask >> (a - first)) & 1ull) == 0) continue;       // uniform: whole block skips the camera
    if (((wave_mask >> (a - first)) & 1ull) == 0) continue;  // uniform: the wave skips the camera
    const CameraParams& cam = args.table[a];
    Projected p;
    TileRect rect;
    int tiles;
    RayData ray;
    bool visible;
    if constexpr (kMode == 1) {
      visible = evaluate_pair(args, cam, g, loaded, m0, m1, m2, opacity, cov, p, rect, tiles);
      tiles = 0;  // pixel intersections are placed by the wave together below
    } else if constexpr (kMode == 2) {
      visible = evaluate_ray_pair(args, cam, g, loaded, m0, m1, m2, opacity, cov, p, rect, tiles, ray);
    } else {
      visible = evaluate_pair(args, cam, g, loaded, m0, m1, m2, opacity, cov, p, rect, tiles);
    }
    int inclusive_visible = visible ? 1 : 0;
    int inclusive_tiles = tiles;
    wave_inclusive_scan2(inclusive_visible, inclusive_tiles);
    const int total_visible = __shfl(inclusive_visible, kWave - 1, kWave);
    const int total_tiles = __shfl(inclusive_tiles, kWave - 1, kWave);
    if (visible && !(kDense && args.dense_mode == 2)) {  // the emission-only pass keeps the records
      const int64_t record = visible_base + inclusive_visible - 1;
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
        if constexpr (kMode == 2) {
          float4* row = reinterpret_cast<float4*>(ray_records + record * 12);
          row[0] = make_float4(ray.precision[0], ray.precision[1], ray.precision[2], ray.precision[3]);
          row[1] = make_float4(ray.precision[4], ray.precision[5], ray.precision_mean[0], ray.precision_mean[1]);
          row[2] = make_float4(ray.precision_mean[2], ray.mean_precision_mean, ray.cull_half[0], ray.cull_half[1]);
        }
      }
    }
    if constexpr (kDense) {
      // Pixel intersections, placed by the wave together in (lane, row-major)
      // order from the wave's scanned offset: the slots of a per-lane loop.
      if (args.dense_mode != 1) {
        const int64_t record = visible_base + inclusive_visible - 1;
        const int32_t value = visible && record < visible_capacity ? static_cast<int32_t>(record) : -1;
        const uint32_t depth_bits = visible ? vgr_float_bits(p.depth) : 0u;
        const int bucket = visible && args.depth_cutoff != nullptr
                               ? depth_bucket(p.depth, args.near_plane, args.far_plane, args.depth_buckets)
                               : 0;
        const bool walk =
            visible && dense_rect_reaches(rect, cam.tile_base, cam.tiles_x, bucket, args.dense_block_cutoff);
        const int wave_pixels = dense_wave_pixels(
            walk, p, rect, cam.tile_base, cam.tiles_x, bucket, args.depth_cutoff,
            [&](bool keep, int index, int owner, uint32_t pixel, float) {
              const int32_t owner_value = __shfl(value, owner, kWave);
              const uint32_t owner_depth = static_cast<uint32_t>(__shfl(static_cast<int>(depth_bits), owner, kWave));
              const int64_t slot = intersection_base + index;
              if (keep && slot < intersection_capacity) {
                keys[slot] =
                    owner_value >= 0 ? ((static_cast<uint64_t>(pixel) << 32) | owner_depth) : kSentinelKey;
                values[slot] = owner_value;
              }
            });
        intersection_base += wave_pixels;
      }
      visible_base += total_visible;
      continue;
    }
    // Tile intersections of the wave: the wave writes its total_tiles slots as
    // contiguous 64-slot runs (coalesced stores).  Output slot j belongs to the
    // lane whose [exclusive, inclusive) tile range holds j, found by binary
    // search over the lanes' inclusive counts; that lane's rectangle gives the
    // tile in row-major order, as a per-lane loop over the rectangle would.
    const int exclusive_tiles = inclusive_tiles - tiles;
    const int owner_min_x = visible ? rect.min_x : 0;
    const int owner_min_y = visible ? rect.min_y : 0;
    const int owner_width = visible ? rect.max_x - rect.min_x : 1;
    const int32_t owner_record =
        visible && visible_base + inclusive_visible - 1 < visible_capacity
            ? static_cast<int32_t>(visible_base + inclusive_visible - 1)
            : -1;
    const uint32_t owner_depth = visible ? vgr_float_bits(p.depth) : 0u;
    for (int run = 0; run < total_tiles; run += kWave) {
      const int j = run + lane;
      int owner = 0;
#pragma unroll
      for (int step = kWave / 2; step > 0; step >>= 1) {
        if (__shfl(inclusive_tiles, owner + step - 1, kWave) <= j) owner += step;
      }
      const int local = j - __shfl(exclusive_tiles, owner, kWave);
      const int min_x = __shfl(owner_min_x, owner, kWave);
      const int min_y = __shfl(owner_min_y, owner, kWave);
      const int width = __shfl(owner_width, owner, kWave);
      const int32_t value = __shfl(owner_record, owner, kWave);
      const uint32_t depth_bits = __shfl(owner_depth, owner, kWave);
      const int64_t slot = intersection_base + j;
      if (j < total_tiles && slot < intersection_capacity) {
        // Row within the rectangle: float quotient, corrected to the exact integer one.
        int dy = static_cast<int>(static_cast<float>(local) / static_cast<float>(width));
        if (dy * width > local) --dy;
        if ((dy + 1) * width <= local) ++dy;
        const int dx = local - dy * width;
        if (value >= 0) {
          const uint32_t tile = cam.tile_base + static_cast<uint32_t>((min_y + dy) * cam.tiles_x + min_x + dx);
          keys[slot] = (static_cast<uint64_t>(tile) << 32) | depth_bits;
          // Training forwards sort the slot itself (slot_records maps it back).
          values[slot] = slot_records != nullptr ? static_cast<int32_t>(slot) : value;
        } else {
          keys[slot] = kSentinelKey;
          values[slot] = -1;
        }
        if (slot_records != nullptr) slot_records[slot] = value;
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
      counters[VGR_COUNTER_SORT_FAILURES] = 0;
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

void make_projection_args(const vgr_scene& scene, const vgr_cameras& cameras, const vgr_options& options,
                          const ScratchLayout& layout, const char* scratch, ProjectionArgs& args) {
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
  args.dense_mode = 0;
  args.depth_buckets = 0;
  args.depth_cutoff = nullptr;
  args.dense_block_cutoff = nullptr;
  args.antialiased = options.antialiased;
  args.near_plane = options.near_plane;
  args.far_plane = options.far_plane;
  args.support_sigma = options.gaussian_support_sigma;
  args.covariance_epsilon = options.covariance_epsilon;
  args.gaussian_ids = scene.gaussian_ids;
  args.item_blocks = nullptr;
  args.item_masks = nullptr;
  args.item_count = nullptr;
  if (layout.projection_blocks > 0 && layout.culling && scene.chunk_bounds != nullptr) {
    args.item_blocks = reinterpret_cast<const int32_t*>(scratch + layout.item_blocks);
    args.item_masks = reinterpret_cast<const uint64_t*>(scratch + layout.item_masks);
    args.item_count = reinterpret_cast<const int64_t*>(scratch + layout.item_totals);
  }
}

int launch_cull(const ProjectionArgs& args, const vgr_scene& scene, const ScratchLayout& layout, char* scratch,
                int64_t* statistics, hipStream_t stream) {
  const int64_t blocks = layout.projection_blocks;
  if (blocks <= 0) return 0;
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
  if (statistics != nullptr) launch_copy_i64(item_totals, statistics + 2, stream);
  const hipError_t status = hipGetLastError();
  if (status != hipSuccess) {
    set_error("culling launch failed: %s", hipGetErrorString(status));
    return VGR_E_LAUNCH;
  }
  return 0;
}

namespace {

// One count -> scan -> write pass over the projection grid (records and/or intersections).
template <int kMode>
void projection_pass(const ProjectionArgs& args, const vgr_workspace& ws, int64_t* block_counts, int64_t* block_bases,
                     uint64_t* wave_masks, int64_t* partials, int64_t blocks, hipStream_t stream) {
  const int64_t waves = blocks * kProjectionWaves;
  if (blocks > 0) {
    hipLaunchKernelGGL(projection_count_kernel<kMode>, dim3(static_cast<unsigned>(blocks)), dim3(kBlockThreads), 0,
                       stream, args, block_counts, wave_masks);
  }
  scan_pairs(block_counts, waves, block_bases, partials, ws.visible_capacity, ws.intersection_capacity,
             ws.counters, nullptr, stream);
  if (blocks <= 0) return;
  const dim3 grid(static_cast<unsigned>(blocks)), threads(kBlockThreads);
  switch (args.sh_degree) {
    case 1:
      hipLaunchKernelGGL((projection_write_kernel<1, kMode>), grid, threads, 0, stream, args, block_bases, wave_masks,
                         ws.records, ws.radii, ws.keys_in, ws.values_in, ws.slot_records, ws.visible_capacity,
                         ws.intersection_capacity, ws.ray_records);
      break;
    case 2:
      hipLaunchKernelGGL((projection_write_kernel<2, kMode>), grid, threads, 0, stream, args, block_bases, wave_masks,
                         ws.records, ws.radii, ws.keys_in, ws.values_in, ws.slot_records, ws.visible_capacity,
                         ws.intersection_capacity, ws.ray_records);
      break;
    case 3:
      hipLaunchKernelGGL((projection_write_kernel<3, kMode>), grid, threads, 0, stream, args, block_bases, wave_masks,
                         ws.records, ws.radii, ws.keys_in, ws.values_in, ws.slot_records, ws.visible_capacity,
                         ws.intersection_capacity, ws.ray_records);
      break;
    default:
      hipLaunchKernelGGL((projection_write_kernel<0, kMode>), grid, threads, 0, stream, args, block_bases, wave_masks,
                         ws.records, ws.radii, ws.keys_in, ws.values_in, ws.slot_records, ws.visible_capacity,
                         ws.intersection_capacity, ws.ray_records);
      break;
  }
}

}  // namespace

int launch_projection(const vgr_scene& scene, const vgr_cameras& cameras, const vgr_options& options,
                      const vgr_workspace& ws, const ScratchLayout& layout, char* scratch, hipStream_t stream) {
  ProjectionArgs args;
  make_projection_args(scene, cameras, options, layout, scratch, args);
  const bool culled = args.item_blocks != nullptr;
  args.item_blocks = nullptr;  // the culling kernels below take the full candidate grid
  args.item_masks = nullptr;
  args.item_count = nullptr;

  int64_t* block_counts = reinterpret_cast<int64_t*>(scratch + layout.block_counts);
  int64_t* block_bases = reinterpret_cast<int64_t*>(scratch + layout.block_bases);
  uint64_t* wave_masks = reinterpret_cast<uint64_t*>(scratch + layout.wave_masks);
  const int64_t blocks = layout.projection_blocks;
  const int64_t waves = blocks * kProjectionWaves;
  // Partials live in the tail of block_bases' region, which holds 2*(waves + scan_blocks) entries.
  int64_t* partials = block_bases + 2 * waves;

  if (culled) {
    const int status = launch_cull(args, scene, layout, scratch, options.statistics, stream);
    if (status != 0) return status;
    args.item_blocks = reinterpret_cast<const int32_t*>(scratch + layout.item_blocks);
    args.item_masks = reinterpret_cast<const uint64_t*>(scratch + layout.item_masks);
    args.item_count = reinterpret_cast<const int64_t*>(scratch + layout.item_totals);
  }

  if (options.ray_evaluation != 0) {
    projection_pass<2>(args, ws, block_counts, block_bases, wave_masks, partials, blocks, stream);
  } else if (options.tile_size != 1) {
    projection_pass<0>(args, ws, block_counts, block_bases, wave_masks, partials, blocks, stream);
  } else if (options.depth_bucket_count <= 0) {
    args.dense_mode = 0;  // records and pixel intersections in one pass
    projection_pass<1>(args, ws, block_counts, block_bases, wave_masks, partials, blocks, stream);
  } else {
    // Records first, then the per-pixel depth cutoff from them, then the pixel
    // intersections that survive it (the second pass re-evaluates the same pairs,
    // so its scanned record indices are those of the first).
    args.dense_mode = 1;
    projection_pass<1>(args, ws, block_counts, block_bases, wave_masks, partials, blocks, stream);
    const int status = launch_dense_cutoff(cameras, options, ws, layout, scratch, stream);
    if (status != 0) return status;
    args.dense_mode = 2;
    args.depth_buckets = options.depth_bucket_count;
    args.depth_cutoff = reinterpret_cast<const int32_t*>(scratch + layout.dense_cutoff);
    args.dense_block_cutoff = reinterpret_cast<const int32_t*>(scratch + layout.dense_block_cutoff);
    projection_pass<1>(args, ws, block_counts, block_bases, wave_masks, partials, blocks, stream);
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
