// Fused RGB / expected-depth / alpha / semantic compositing for gfx950.
//
// Reference behaviour: composite_tiles_kernel<false,false,FULL>
// (renderer_cuda.cu:1973-2305).  One workgroup renders one (active camera,
// tile): tile_size^2 lanes, one pixel each (16x16 = four wave64s).  The
// depth-sorted intersection list of the tile is consumed in batches: every
// lane gathers whole 48-byte records (three 16-byte loads) for the next batch
// into registers while the current batch is composited out of LDS, so the
// gather latency overlaps the blend loop.  Inside the loop every lane reads
// the same LDS address, which the LDS serves as a broadcast.
//
// The per-pixel arithmetic is the reference's, term by term, with vgr_expf in
// place of __expf.  The semantic winner is tracked as a Gaussian index and the
// int64 semantic label is gathered once per pixel at the end.
#include <type_traits>

#include "vgr_composite_device.h"
#include "vgr_project_device.h"

namespace vgr {
namespace {

using namespace dev;

// ---- tile launch order -------------------------------------------------------
// Blocks sorted by descending list length (counting sort on length / 8, capped):
// the longest tiles start first and the short ones fill in behind them, instead of
// a long tile starting late and running alone at the end of the grid.  One
// workgroup; the order inside a length bucket is unspecified (it changes only the
// schedule: every tile's outputs are its own).  On 16 cameras of 512x512 (lists
// of ~950 entries on average, up to 5k) the forward goes 2.85 -> 2.24 ms and
// forward+backward 9.6 -> 8.0 ms.
constexpr int kOrderBuckets = 1024;
constexpr int kOrderThreads = 1024;

__device__ __forceinline__ int order_bucket(const CameraParams* __restrict__ table,
                                            const int32_t* __restrict__ tile_starts,
                                            const int32_t* __restrict__ tile_ends, int active_count,
                                            int uniform_tiles, int64_t block) {
  int active;
  int64_t image_tile;
  if (uniform_tiles > 0) {
    active = static_cast<int>(block / uniform_tiles);
    if (active >= active_count) return 0;
    image_tile = block - static_cast<int64_t>(active) * uniform_tiles;
  } else {
    active = find_active_camera(table, active_count, block);
    image_tile = block - table[active].active_tile_begin;
    if (image_tile >= static_cast<int64_t>(table[active].tiles_x) * table[active].tiles_y) return 0;
  }
  if (table[active].camera < 0) return 0;
  const uint32_t global_tile = table[active].tile_base + static_cast<uint32_t>(image_tile);
  const int64_t start = tile_starts[global_tile];
  const int64_t end = tile_ends[global_tile];
  const int64_t bucket = (start >= 0 && end > start ? end - start : 0) >> 3;
  return static_cast<int>(bucket < kOrderBuckets - 1 ? bucket : kOrderBuckets - 1);
}

__global__ __launch_bounds__(kOrderThreads) void tile_order_kernel(const CameraParams* __restrict__ table,
                                                                   const int32_t* __restrict__ tile_starts,
                                                                   const int32_t* __restrict__ tile_ends,
                                                                   int active_count, int uniform_tiles, int64_t blocks,
                                                                   int32_t* __restrict__ order) {
  __shared__ uint32_t s_next[kOrderBuckets];
  __shared__ uint32_t s_wave_totals[kOrderThreads / kWave];
  for (int i = threadIdx.x; i < kOrderBuckets; i += kOrderThreads) s_next[i] = 0;
  __syncthreads();
  for (int64_t b = threadIdx.x; b < blocks; b += kOrderThreads) {
    atomicAdd(&s_next[order_bucket(table, tile_starts, tile_ends, active_count, uniform_tiles, b)], 1u);
  }
  __syncthreads();
  // Exclusive scan in descending bucket order: thread t owns bucket kOrderBuckets-1-t.
  static_assert(kOrderBuckets == kOrderThreads, "one bucket per thread");
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int bucket = kOrderBuckets - 1 - static_cast<int>(threadIdx.x);
  const uint32_t count = s_next[bucket];
  uint32_t inclusive = count;
#pragma unroll
  for (int offset = 1; offset < kWave; offset <<= 1) {
    const uint32_t t = __shfl_up(inclusive, offset, kWave);
    if (lane >= offset) inclusive += t;
  }
  if (lane == kWave - 1) s_wave_totals[wave] = inclusive;
  __syncthreads();
  uint32_t prefix = 0;
  for (int w = 0; w < wave; ++w) prefix += s_wave_totals[w];
  s_next[bucket] = prefix + inclusive - count;
  __syncthreads();
  for (int64_t b = threadIdx.x; b < blocks; b += kOrderThreads) {
    const int k = order_bucket(table, tile_starts, tile_ends, active_count, uniform_tiles, b);
    order[atomicAdd(&s_next[k], 1u)] = static_cast<int32_t>(b);
  }
}

struct CompositeArgs {
  const CameraParams* table;
  const int32_t* tile_starts;
  const int32_t* tile_ends;
  const int32_t* sorted_values;
  const vgr_record* records;
  const int64_t* semantic_ids;
  float* rgb;
  float* depth;
  float* alpha;
  int64_t* semantic;
  int64_t* statistics;   // optional: [0] += evaluated (pixel, Gaussian) pairs
  int64_t visible_capacity;
  int64_t gaussian_count;
  uint32_t* work_counter;  // persistent queue counter (zeroed before the launch)
  const int32_t* tile_order;  // launch order of the (camera, tile) items
  int64_t total_items;     // (camera, tile) items in the launch
  int active_count;
  int uniform_tiles;     // tiles per camera when every camera has the same resolution, else 0
  int output_srgb;
  float semantic_min_alpha;
  float* final_transmittance;  // optional per-pixel checkpoint (training forward)
  int32_t* last_blended;       // optional: tile-list position one past the last blended splat
  const int32_t* slot_records;  // training forward: the sorted values are slots, mapped to records here
};

// Tile-front ordering: the lists hold packed Gaussian indices (the records are
// recomputed from the scene), a list may be a truncated front of the tile's
// bucket (front_lengths = full bucket sizes), and a tile with a pixel still
// open at the end of a truncated list is appended to open_tiles.
struct FrontCompositeArgs : CompositeArgs {
  ProjectionArgs projection;
  const int32_t* front_lengths;  // null: no open-tile detection (the recomposite pass)
  int32_t* open_tiles;
  uint32_t* open_count;
  const uint32_t* item_count;  // persistent launches: device item count (overrides total_items)
};

template <int kDegree>
using CompositeArgsFor = typename std::conditional<(kDegree < 0), CompositeArgs, FrontCompositeArgs>::type;

template <typename Args>
__device__ __forceinline__ int64_t item_total(const Args& args) {
  if constexpr (std::is_same<Args, FrontCompositeArgs>::value) {
    if (args.item_count != nullptr) return static_cast<int64_t>(*args.item_count);
  }
  return args.total_items;
}

// Gathers one batch: stored records (kDegree < 0) or records recomputed from
// the listed Gaussians (kDegree = SH degree, tile-front lists).
template <int kThreads, int kBatch, int kPerThread, int kDegree, bool kSlots = false>
__device__ __forceinline__ void gather_any(const CompositeArgsFor<kDegree>& args, const CameraParams& cam,
                                           int64_t batch_start,
                                           int64_t range_end, int lane, f32x4 (&a)[kPerThread],
                                           f32x4 (&b)[kPerThread], f32x4 (&c)[kPerThread]) {
  if constexpr (kDegree < 0) {
    gather_batch<kThreads, kBatch, kPerThread, kSlots>(args.sorted_values, args.records, args.visible_capacity,
                                                       batch_start, range_end, lane, a, b, c, args.slot_records);
  } else {
#pragma unroll
    for (int k = 0; k < kPerThread; ++k) {
      const int64_t index = batch_start + k * kThreads + lane;
      if (k * kThreads + lane < kBatch && index < range_end) {
        const int32_t g = args.sorted_values[index];
        if (g >= 0 && g < args.projection.gaussian_count) {
          record_from_gaussian<kDegree>(args.projection, cam, g, a[k], b[k], c[k]);
        } else {
          a[k] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
          b[k] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
          c[k] = f32x4{0.0f, 0.0f, __int_as_float(-1), 0.0f};
        }
      }
    }
  }
}

// Renders one (active camera, tile) work item with the whole workgroup.  With
// kCheckpoint the final transmittance and the list position one past the last
// blended splat are stored per pixel for the backward pass.
template <int kTileSize, int kBatch, bool kFull, bool kCount, bool kCheckpoint, int kDegree = -1>
__device__ __forceinline__ void composite_tile(const CompositeArgsFor<kDegree>& args, int64_t block) {
  using Lists = WaveLists<kTileSize, kBatch>;
  constexpr int kThreads = Lists::kThreads;
  constexpr int kPerThread = Lists::kPerThread;
  static_assert(kThreads % kWave == 0, "tile must be a whole number of waves");
  __shared__ Lists s;

  int active;
  int64_t image_tile;
  if (args.uniform_tiles > 0) {
    active = static_cast<int>(block / args.uniform_tiles);
    if (active >= args.active_count) return;
    image_tile = block - static_cast<int64_t>(active) * args.uniform_tiles;
  } else {
    active = find_active_camera(args.table, args.active_count, block);
    image_tile = block - args.table[active].active_tile_begin;
    if (image_tile >= static_cast<int64_t>(args.table[active].tiles_x) * args.table[active].tiles_y) return;
  }
  const CameraParams& cam = args.table[active];
  if (cam.camera < 0) return;
  const int tile_x = static_cast<int>(image_tile % cam.tiles_x);
  const int tile_y = static_cast<int>(image_tile / cam.tiles_x);
  const uint32_t global_tile = cam.tile_base + static_cast<uint32_t>(image_tile);
  const int lane = threadIdx.x;
  const int wave = lane / kWave;
  const int pixel_x = tile_x * kTileSize + Lists::pixel_dx(lane);
  const int pixel_y = tile_y * kTileSize + Lists::pixel_dy(lane);
  const bool valid_pixel = pixel_x < cam.width && pixel_y < cam.height;
  const float px = static_cast<float>(pixel_x) + 0.5f;
  const float py = static_cast<float>(pixel_y) + 0.5f;
  // Pixel-centre rectangle of tile column span and of each wave's rows.
  const float rect_x0 = static_cast<float>(tile_x * kTileSize) + 0.5f;
  const float rect_x1 = rect_x0 + static_cast<float>(kTileSize - 1);
  const float rect_y_base = static_cast<float>(tile_y * kTileSize) + 0.5f;

  float transmittance = 1.0f;
  float red = 0.0f, green = 0.0f, blue = 0.0f, depth_sum = 0.0f;
  float best_weight = -1.0f;
  int best_gaussian = -1;
  bool done = !valid_pixel;
  int evaluations = 0;
  int last_blended = 0;

  const int64_t range_start = args.tile_starts[global_tile];
  const int64_t range_end = args.tile_ends[global_tile];
  if (range_start >= 0 && range_end > range_start) {
    f32x4 next_a[kPerThread] = {}, next_b[kPerThread] = {}, next_c[kPerThread] = {};
    gather_any<kThreads, kBatch, kPerThread, kDegree, kCheckpoint>(args, cam, range_start, range_end, lane, next_a,
                                                                   next_b, next_c);
    for (int64_t batch_start = range_start; batch_start < range_end; batch_start += kBatch) {
      const int count = static_cast<int>(range_end - batch_start < kBatch ? range_end - batch_start : kBatch);
      __syncthreads();  // previous batch fully consumed
      build_wave_lists<kTileSize, kBatch>(s, next_a, next_b, next_c, count, lane, rect_x0, rect_x1, rect_y_base);
      if (batch_start + kBatch < range_end) {
        gather_any<kThreads, kBatch, kPerThread, kDegree, kCheckpoint>(args, cam, batch_start + kBatch, range_end,
                                                                       lane, next_a, next_b, next_c);
      }
      if (!done) {
        const uint32_t list_count = s.count(wave);
        const int position_base = static_cast<int>(batch_start - range_start);
        // Blends one evaluated splat; returns true when the pixel terminates.
        auto blend = [&](int item, const f32x4& sh, float splat_alpha) -> bool {
          if (splat_alpha < VGR_ALPHA_THRESHOLD) return false;
          const float next_transmittance = transmittance * (1.0f - splat_alpha);
          if (next_transmittance <= VGR_TRANSMITTANCE_THRESHOLD) return true;
          const f32x4 col = s.colour[item];
          const float weight = transmittance * splat_alpha;
          red = fmaf(weight, col.x, red);
          green = fmaf(weight, col.y, green);
          blue = fmaf(weight, col.z, blue);
          if (kFull) {
            depth_sum = fmaf(weight, sh.w, depth_sum);
            if (weight > best_weight) {
              best_weight = weight;
              best_gaussian = __float_as_int(col.w);
            }
          }
          if (kCheckpoint) last_blended = position_base + item + 1;
          transmittance = next_transmittance;
          return false;
        };
        uint32_t i = 0;
        // kUnroll list entries per iteration: their powers and exponentials are
        // independent and overlap (the loop is latency-bound); blending stays in
        // list order, so the per-pixel result is unchanged.
        constexpr int kUnroll = 4;
        for (; i + kUnroll <= list_count; i += kUnroll) {
          int items[kUnroll];
          f32x4 shading[kUnroll];
          float alphas[kUnroll];
          bool passes[kUnroll];
#pragma unroll
          for (int u = 0; u < kUnroll; ++u) items[u] = s.list[wave][i + u];
#pragma unroll
          for (int u = 0; u < kUnroll; ++u) {
            const f32x4 g = s.geometry[items[u]];
            shading[u] = s.shading[items[u]];
            const float dx = px - g.x, dy = py - g.y;
            const float power = g.z * dx * dx + g.w * dx * dy + shading[u].x * dy * dy;
            passes[u] = !(power > 0.0f || power < shading[u].z);
            alphas[u] = passes[u] ? fminf(VGR_MAX_ALPHA, shading[u].y * vgr_expf_nonpositive(power)) : 0.0f;
          }
#pragma unroll
          for (int u = 0; u < kUnroll; ++u) {
            if (done) break;
            if (kCount) ++evaluations;
            if (passes[u] && blend(items[u], shading[u], alphas[u])) done = true;
          }
          if (done) break;
        }
        while (!done && i < list_count) {
          const int item = s.list[wave][i++];
          if (kCount) ++evaluations;
          const f32x4 g = s.geometry[item];
          const f32x4 sh = s.shading[item];
          const float delta_x = px - g.x;
          const float delta_y = py - g.y;
          const float power = g.z * delta_x * delta_x + g.w * delta_x * delta_y + sh.x * delta_y * delta_y;
          if (!(power > 0.0f || power < sh.z)) {
            if (blend(item, sh, fminf(VGR_MAX_ALPHA, sh.y * vgr_expf_nonpositive(power)))) done = true;
          }
        }
      }
      if (!__syncthreads_or(done ? 0 : 1)) break;
    }
  }
  if constexpr (kDegree >= 0) {
    if (args.front_lengths != nullptr) {
      // Truncated list with a pixel still open: the rest of the bucket matters;
      // the tile is sorted in full and composited again.
      const int64_t full = args.front_lengths[global_tile];
      const bool truncated = (range_start >= 0 && range_end > range_start ? range_end - range_start : 0) < full;
      if (truncated && __syncthreads_or(done ? 0 : 1) && threadIdx.x == 0) {
        args.open_tiles[atomicAdd(args.open_count, 1u)] = static_cast<int32_t>(global_tile);
      }
    }
  }

  if (kCount) {
    int64_t total = evaluations;
    for (int offset = 32; offset > 0; offset >>= 1) total += __shfl_xor(total, offset, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0 && total != 0) {
      atomicAdd(reinterpret_cast<unsigned long long*>(args.statistics), static_cast<unsigned long long>(total));
    }
  }
  if (!valid_pixel) return;
  const int64_t pixel = cam.pixel_offset + static_cast<int64_t>(pixel_y) * cam.width + pixel_x;
  const float accumulated_alpha = 1.0f - transmittance;
  if (args.output_srgb) {
    red = vgr_linear_to_srgb(red);
    green = vgr_linear_to_srgb(green);
    blue = vgr_linear_to_srgb(blue);
  }
  args.rgb[pixel * 3 + 0] = red;
  args.rgb[pixel * 3 + 1] = green;
  args.rgb[pixel * 3 + 2] = blue;
  if (kFull) {
    args.alpha[pixel] = accumulated_alpha;
    args.depth[pixel] = accumulated_alpha > 1.0e-8f ? depth_sum / accumulated_alpha : __int_as_float(0x7f800000);
    const bool labelled = best_weight >= 0.0f && accumulated_alpha >= args.semantic_min_alpha &&
                          best_gaussian >= 0 && best_gaussian < args.gaussian_count;
    args.semantic[pixel] = labelled ? args.semantic_ids[best_gaussian] : -1;
  }
  if (kCheckpoint) {
    args.final_transmittance[pixel] = transmittance;
    args.last_blended[pixel] = last_blended;
  }
}

// ---- small tiles (1x1, 2x2, 4x4): one thread per pixel ---------------------------
//
// The tile lists are short and private to a few pixels, so there is nothing to
// stage in LDS: each thread walks its tile's list with direct record loads (the
// threads of one tile read the same records, which the caches share).  The
// per-pixel equation is the tiled compositor's: the same alpha test, blend,
// stop rule, depth normalisation and semantic pick (renderer_cuda.cu:2076-2304).
template <bool kFull, bool kCount>
__global__ __launch_bounds__(256) void composite_pixels_kernel(CompositeArgs args, int tile_size,
                                                               int64_t thread_count) {
  const int64_t index = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (index >= thread_count) return;
  const int tile_pixels = tile_size * tile_size;
  const int64_t block = index / tile_pixels;
  const int local = static_cast<int>(index - block * tile_pixels);
  int active;
  int64_t image_tile;
  if (args.uniform_tiles > 0) {
    active = static_cast<int>(block / args.uniform_tiles);
    if (active >= args.active_count) return;
    image_tile = block - static_cast<int64_t>(active) * args.uniform_tiles;
  } else {
    active = find_active_camera(args.table, args.active_count, block);
    image_tile = block - args.table[active].active_tile_begin;
    if (image_tile >= static_cast<int64_t>(args.table[active].tiles_x) * args.table[active].tiles_y) return;
  }
  const CameraParams& cam = args.table[active];
  if (cam.camera < 0) return;
  const int tile_x = static_cast<int>(image_tile % cam.tiles_x);
  const int tile_y = static_cast<int>(image_tile / cam.tiles_x);
  const int pixel_x = tile_x * tile_size + local % tile_size;
  const int pixel_y = tile_y * tile_size + local / tile_size;
  if (pixel_x >= cam.width || pixel_y >= cam.height) return;
  const uint32_t global_tile = cam.tile_base + static_cast<uint32_t>(image_tile);
  const float px = static_cast<float>(pixel_x) + 0.5f;
  const float py = static_cast<float>(pixel_y) + 0.5f;

  float transmittance = 1.0f;
  float red = 0.0f, green = 0.0f, blue = 0.0f, depth_sum = 0.0f;
  float best_weight = -1.0f;
  int best_gaussian = -1;
  int64_t evaluations = 0;
  const int64_t range_start = args.tile_starts[global_tile];
  const int64_t range_end = args.tile_ends[global_tile];
  for (int64_t i = range_start < 0 ? range_end : range_start; i < range_end; ++i) {
    const int32_t record = args.sorted_values[i];
    if (record < 0 || record >= args.visible_capacity) continue;  // only with a corrupted list
    const f32x4* src = reinterpret_cast<const f32x4*>(args.records + record);
    const f32x4 a = src[0];  // mean_x, mean_y, conic_0, conic_1
    const f32x4 b = src[1];  // conic_2, opacity, depth, red
    if (kCount) ++evaluations;
    const float dx = px - a.x;
    const float dy = py - a.y;
    const float power = -0.5f * (a.z * dx * dx + 2.0f * a.w * dx * dy + b.x * dy * dy);
    if (power > 0.0f || power < -20.0f) continue;
    const float alpha = fminf(VGR_MAX_ALPHA, b.y * vgr_expf_nonpositive(power));
    if (alpha < VGR_ALPHA_THRESHOLD) continue;
    const float next_transmittance = transmittance * (1.0f - alpha);
    if (next_transmittance <= VGR_TRANSMITTANCE_THRESHOLD) break;
    const f32x4 c = src[2];  // green, blue, gaussian id bits, camera bits
    const float weight = transmittance * alpha;
    red = fmaf(weight, b.w, red);
    green = fmaf(weight, c.x, green);
    blue = fmaf(weight, c.y, blue);
    if (kFull) {
      depth_sum = fmaf(weight, b.z, depth_sum);
      if (weight > best_weight) {
        best_weight = weight;
        best_gaussian = __float_as_int(c.z);
      }
    }
    transmittance = next_transmittance;
  }
  if (kCount) {
    int64_t total = evaluations;
    for (int offset = 32; offset > 0; offset >>= 1) total += __shfl_xor(total, offset, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0 && total != 0) {
      atomicAdd(reinterpret_cast<unsigned long long*>(args.statistics), static_cast<unsigned long long>(total));
    }
  }
  const int64_t pixel = cam.pixel_offset + static_cast<int64_t>(pixel_y) * cam.width + pixel_x;
  const float accumulated_alpha = 1.0f - transmittance;
  if (args.output_srgb) {
    red = vgr_linear_to_srgb(red);
    green = vgr_linear_to_srgb(green);
    blue = vgr_linear_to_srgb(blue);
  }
  args.rgb[pixel * 3 + 0] = red;
  args.rgb[pixel * 3 + 1] = green;
  args.rgb[pixel * 3 + 2] = blue;
  if (kFull) {
    args.alpha[pixel] = accumulated_alpha;
    args.depth[pixel] = accumulated_alpha > 1.0e-8f ? depth_sum / accumulated_alpha : __int_as_float(0x7f800000);
    const bool labelled = best_weight >= 0.0f && accumulated_alpha >= args.semantic_min_alpha &&
                          best_gaussian >= 0 && best_gaussian < args.gaussian_count;
    args.semantic[pixel] = labelled ? args.semantic_ids[best_gaussian] : -1;
  }
}

// ---- exact-ray compositing ---------------------------------------------------------
//
// One workgroup per kRaster x kRaster raster tile, reading the sorted list of the
// bin that holds it (32x32 bins under 16x16 raster tiles; bins equal raster tiles
// at 8x8).  A batch of list entries is staged in LDS (record and precision rows,
// coalesced), then every thread evaluates its pixel's ray against the batch in
// list order (renderer_cuda.cu:2176-2262): alpha is opacity x exp(-m^2/2) for the
// ray's closest approach m (0 beyond the support, capped at 1), each entry with
// alpha > 0 is blended, and the pixel stops after the entry that takes its
// transmittance to 1e-4.
__device__ __forceinline__ float ray_alpha(float ray_x, float ray_y, const f32x4& a, const f32x4& b,
                                           const f32x4& c, float support_sigma) {
  // a = (P00, P01, P02, P11), b = (P12, P22, Pm0, Pm1), c = (Pm2, mPm, opacity, depth).
  const float along = a.x * ray_x * ray_x + 2.0f * a.y * ray_x * ray_y + 2.0f * a.z * ray_x + a.w * ray_y * ray_y +
                      2.0f * b.x * ray_y + b.y;
  if (!(along > 0.0f)) return 0.0f;
  const float cross = ray_x * b.z + ray_y * b.w + c.x;
  const float distance2 = fmaxf(0.0f, c.y - cross * cross / along);
  if (distance2 > support_sigma * support_sigma) return 0.0f;
  return fminf(1.0f, c.z * vgr_expf(-0.5f * distance2));
}

template <int kRaster, bool kFull>
__global__ __launch_bounds__(kRaster* kRaster) void composite_ray_kernel(CompositeArgs args,
                                                                         const float* __restrict__ ray_records,
                                                                         float support_sigma, int raster_tiles_x,
                                                                         int raster_tiles, int bin_size) {
  constexpr int kThreads = kRaster * kRaster;
  __shared__ f32x4 s_a[kThreads];
  __shared__ f32x4 s_b[kThreads];
  __shared__ f32x4 s_c[kThreads];
  __shared__ f32x4 s_colour[kThreads];  // red, green, blue, gaussian id bits
  // Each wave owns an 8x8 pixel block (16x16 raster tiles: four quadrants) and
  // walks the batch entries whose cull box reaches it, compacted in list order;
  // with 8x8 raster tiles the bin is the tile, whose box every entry reaches.
  constexpr int kWaves = kThreads / kWave;
  constexpr bool kCull = kWaves > 1;
  constexpr int kBlocksX = kRaster / 8;
  __shared__ uint16_t s_list[kWaves][kThreads];
  __shared__ uint32_t s_totals[kWaves][kWaves];  // [source wave][list]
  const int wave = static_cast<int>(threadIdx.x) / kWave;
  const int wave_lane = static_cast<int>(threadIdx.x) & (kWave - 1);
  const int block_x0 = (wave % kBlocksX) * 8, block_y0 = (wave / kBlocksX) * 8;
  const int64_t block = blockIdx.x;
  const int active = static_cast<int>(block / raster_tiles);
  if (active >= args.active_count) return;
  const CameraParams& cam = args.table[active];
  if (cam.camera < 0) return;
  const int raster_tile = static_cast<int>(block - static_cast<int64_t>(active) * raster_tiles);
  const int tile_x = raster_tile % raster_tiles_x;
  const int tile_y = raster_tile / raster_tiles_x;
  const int pixel_x = tile_x * kRaster + block_x0 + (wave_lane & 7);
  const int pixel_y = tile_y * kRaster + block_y0 + (wave_lane >> 3);
  const bool valid_pixel = pixel_x < cam.width && pixel_y < cam.height;
  const int bin_x = tile_x * kRaster / bin_size;
  const int bin_y = tile_y * kRaster / bin_size;
  const uint32_t global_bin = cam.tile_base + static_cast<uint32_t>(bin_y * cam.tiles_x + bin_x);
  const float ray_x = ((static_cast<float>(pixel_x) + 0.5f) - cam.center_x) / cam.focal_x;
  const float ray_y = ((static_cast<float>(pixel_y) + 0.5f) - cam.center_y) / cam.focal_y;

  float transmittance = 1.0f;
  float red = 0.0f, green = 0.0f, blue = 0.0f, depth_sum = 0.0f;
  float best_weight = -1.0f;
  int best_gaussian = -1;
  bool done = !valid_pixel;
  const int64_t range_start = args.tile_starts[global_bin];
  const int64_t range_end = args.tile_ends[global_bin];
  if (range_start >= 0 && range_end > range_start) {
    for (int64_t batch_start = range_start; batch_start < range_end; batch_start += kThreads) {
      const int count = static_cast<int>(range_end - batch_start < kThreads ? range_end - batch_start : kThreads);
      __syncthreads();  // previous batch fully consumed
      uint32_t waves = 0u;
      if (static_cast<int>(threadIdx.x) < count) {
        const int32_t record = args.sorted_values[batch_start + threadIdx.x];
        if (record >= 0 && record < args.visible_capacity) {
          const f32x4* rec = reinterpret_cast<const f32x4*>(args.records + record);
          const f32x4* ray = reinterpret_cast<const f32x4*>(ray_records + static_cast<int64_t>(record) * 12);
          const f32x4 r0 = rec[0], r1 = rec[1], r2 = rec[2], q2 = ray[2];
          s_a[threadIdx.x] = ray[0];
          s_b[threadIdx.x] = ray[1];
          s_c[threadIdx.x] = f32x4{q2.x, q2.y, r1.y, r1.z};
          s_colour[threadIdx.x] = f32x4{r1.w, r2.x, r2.y, r2.z};
          // Each wave's block of pixel centres against the entry's cull box (centre
          // r0.xy, half extents q2.zw): a wave outside it has alpha 0 at every
          // pixel, so the entry is left out of its list.
          waves = kCull ? 0u : 1u;
          if constexpr (kCull) {
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
              const float x0 = static_cast<float>(tile_x * kRaster + (w % kBlocksX) * 8) + 0.5f;
              const float y0 = static_cast<float>(tile_y * kRaster + (w / kBlocksX) * 8) + 0.5f;
              if (fabsf(r0.x - fminf(fmaxf(r0.x, x0), x0 + 7.0f)) <= q2.z &&
                  fabsf(r0.y - fminf(fmaxf(r0.y, y0), y0 + 7.0f)) <= q2.w) {
                waves |= 1u << w;
              }
            }
          }
        } else {
          // Only reachable with a corrupted list: an entry that never blends.
          s_a[threadIdx.x] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
          s_b[threadIdx.x] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
          s_c[threadIdx.x] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
          s_colour[threadIdx.x] = f32x4{0.0f, 0.0f, 0.0f, __int_as_float(-1)};
        }
      }
      // Per-wave lists in batch order: ballot prefix within the source wave, then
      // the totals of the source waves before it.
      int list_count = count;
      if constexpr (kCull) {
        const uint64_t below = wave_lane == 0 ? 0ull : (~0ull >> (kWave - wave_lane));
        uint32_t prefix[kWaves];
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
          const uint64_t vote = __ballot((waves >> w) & 1u);
          prefix[w] = static_cast<uint32_t>(__popcll(vote & below));
          if (wave_lane == 0) s_totals[wave][w] = static_cast<uint32_t>(__popcll(vote));
        }
        __syncthreads();
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
          if ((waves >> w) & 1u) {
            uint32_t position = prefix[w];
            for (int q = 0; q < wave; ++q) position += s_totals[q][w];
            s_list[w][position] = static_cast<uint16_t>(threadIdx.x);
          }
        }
        list_count = 0;
#pragma unroll
        for (int q = 0; q < kWaves; ++q) list_count += static_cast<int>(s_totals[q][wave]);
      }
      __syncthreads();
      for (int i = 0; !done && i < list_count; ++i) {
        const int item = kCull ? static_cast<int>(s_list[wave][i]) : i;
        const f32x4 c = s_c[item];
        const float alpha = ray_alpha(ray_x, ray_y, s_a[item], s_b[item], c, support_sigma);
        if (!(alpha > 0.0f)) continue;
        const float next_transmittance = transmittance * (1.0f - alpha);
        const f32x4 col = s_colour[item];
        const float weight = transmittance * alpha;
        red = fmaf(weight, col.x, red);
        green = fmaf(weight, col.y, green);
        blue = fmaf(weight, col.z, blue);
        if (kFull) {
          depth_sum = fmaf(weight, c.w, depth_sum);
          if (weight > best_weight) {
            best_weight = weight;
            best_gaussian = __float_as_int(col.w);
          }
        }
        transmittance = next_transmittance;
        if (next_transmittance <= VGR_TRANSMITTANCE_THRESHOLD) done = true;
      }
      if (!__syncthreads_or(done ? 0 : 1)) break;
    }
  }
  if (!valid_pixel) return;
  const int64_t pixel = cam.pixel_offset + static_cast<int64_t>(pixel_y) * cam.width + pixel_x;
  const float accumulated_alpha = 1.0f - transmittance;
  if (args.output_srgb) {
    red = vgr_linear_to_srgb(red);
    green = vgr_linear_to_srgb(green);
    blue = vgr_linear_to_srgb(blue);
  }
  args.rgb[pixel * 3 + 0] = red;
  args.rgb[pixel * 3 + 1] = green;
  args.rgb[pixel * 3 + 2] = blue;
  if (kFull) {
    args.alpha[pixel] = accumulated_alpha;
    args.depth[pixel] = accumulated_alpha > 1.0e-8f ? depth_sum / accumulated_alpha : __int_as_float(0x7f800000);
    const bool labelled = best_weight >= 0.0f && accumulated_alpha >= args.semantic_min_alpha &&
                          best_gaussian >= 0 && best_gaussian < args.gaussian_count;
    args.semantic[pixel] = labelled ? args.semantic_ids[best_gaussian] : -1;
  }
}

template <int kRaster>
void launch_ray(const CompositeArgs& args, const vgr_options& options, const float* ray_records, int active_count,
                bool full, hipStream_t stream) {
  const int raster_tiles_x = (options.width + kRaster - 1) / kRaster;
  const int raster_tiles = raster_tiles_x * ((options.height + kRaster - 1) / kRaster);
  const int64_t blocks = static_cast<int64_t>(active_count) * raster_tiles;
  if (blocks <= 0) return;
  const dim3 grid(static_cast<unsigned>(blocks)), block(kRaster * kRaster);
  if (full) {
    hipLaunchKernelGGL((composite_ray_kernel<kRaster, true>), grid, block, 0, stream, args, ray_records,
                       options.gaussian_support_sigma, raster_tiles_x, raster_tiles, options.tile_size);
  } else {
    hipLaunchKernelGGL((composite_ray_kernel<kRaster, false>), grid, block, 0, stream, args, ray_records,
                       options.gaussian_support_sigma, raster_tiles_x, raster_tiles, options.tile_size);
  }
}

// Minimum waves per SIMD for 16x16 tiles (launch-bounds occupancy target): eight
// means at most 64 VGPRs.  The blend loop is latency-bound, so waves in flight
// pay more than registers.  The evaluation-counting and persistent variants
// need a few more registers and take seven and six (no spills).
template <int kTileSize, int kBatch, bool kCount, bool kPersistent = false, int kDegree = -1>
constexpr int kCompositeWavesPerSimd =
    kTileSize != 16 ? 1
                    : (kDegree >= 2 ? 4  // recomputed records with SH2/3 coefficients in registers
                                    : (kDegree >= 0 ? 6
                                                    : (kBatch >= 512 ? 4 : (kPersistent ? 6 : (kCount ? 7 : 8)))));

template <int kTileSize, int kBatch, bool kFull, bool kCount, bool kCheckpoint, int kDegree = -1>
__global__ __launch_bounds__(kTileSize* kTileSize, (kCompositeWavesPerSimd<kTileSize, kBatch, kCount, false, kDegree>)) void
composite_kernel(CompositeArgsFor<kDegree> args) {
  const int64_t block = args.tile_order != nullptr ? args.tile_order[blockIdx.x] : blockIdx.x;
  composite_tile<kTileSize, kBatch, kFull, kCount, kCheckpoint, kDegree>(args, block);
}

// Persistent form: a fixed grid whose workgroups pull (camera, tile) items from
// a device counter until the queue is empty; every workgroup exits once the
// counter passes the item count.
template <int kTileSize, int kBatch, bool kFull, bool kCount, bool kCheckpoint, int kDegree = -1>
__global__ __launch_bounds__(kTileSize* kTileSize, (kCompositeWavesPerSimd<kTileSize, kBatch, kCount, true, kDegree>)) void
composite_persistent_kernel(CompositeArgsFor<kDegree> args) {
  __shared__ int64_t s_item;
  const int64_t total = item_total(args);
  for (;;) {
    if (threadIdx.x == 0) s_item = static_cast<int64_t>(atomicAdd(args.work_counter, 1u));
    __syncthreads();
    const int64_t item = s_item;
    __syncthreads();
    if (item >= total) break;
    composite_tile<kTileSize, kBatch, kFull, kCount, kCheckpoint, kDegree>(
        args, args.tile_order != nullptr ? args.tile_order[item] : item);
  }
}

template <int kTileSize, int kBatch, bool kFull, bool kCount, bool kCheckpoint = false, int kDegree = -1>
void launch_variant(const CompositeArgsFor<kDegree>& args, int64_t blocks, int persistent, hipStream_t stream) {
  const dim3 block(kTileSize * kTileSize);
  if (persistent > 0) {
    const int64_t grid = blocks < persistent ? blocks : persistent;
    hipLaunchKernelGGL((composite_persistent_kernel<kTileSize, kBatch, kFull, kCount, kCheckpoint, kDegree>),
                       dim3(static_cast<unsigned>(grid)), block, 0, stream, args);
  } else {
    hipLaunchKernelGGL((composite_kernel<kTileSize, kBatch, kFull, kCount, kCheckpoint, kDegree>),
                       dim3(static_cast<unsigned>(blocks)), block, 0, stream, args);
  }
}

// Tile-front lists (16x16 tiles, 256-entry batches): records recomputed per SH degree.
template <int kDegree>
void launch_projected_degree(bool full, const FrontCompositeArgs& args, int64_t blocks, int persistent,
                             hipStream_t stream) {
  if (args.final_transmittance != nullptr) {
    // Checkpointed (pose-training) forward: final transmittance and last blended position.
    launch_variant<16, 256, true, false, true, kDegree>(args, blocks, persistent, stream);
  } else if (args.statistics != nullptr) {
    if (full) {
      launch_variant<16, 256, true, true, false, kDegree>(args, blocks, persistent, stream);
    } else {
      launch_variant<16, 256, false, true, false, kDegree>(args, blocks, persistent, stream);
    }
  } else if (full) {
    launch_variant<16, 256, true, false, false, kDegree>(args, blocks, persistent, stream);
  } else {
    launch_variant<16, 256, false, false, false, kDegree>(args, blocks, persistent, stream);
  }
}

void launch_projected(int degree, bool full, const FrontCompositeArgs& args, int64_t blocks, int persistent,
                      hipStream_t stream) {
  switch (degree) {
    case 1:
      launch_projected_degree<1>(full, args, blocks, persistent, stream);
      break;
    case 2:
      launch_projected_degree<2>(full, args, blocks, persistent, stream);
      break;
    case 3:
      launch_projected_degree<3>(full, args, blocks, persistent, stream);
      break;
    default:
      launch_projected_degree<0>(full, args, blocks, persistent, stream);
      break;
  }
}

template <int kTileSize, int kBatch>
void launch_full_or_rgb(bool full, const CompositeArgs& args, int64_t blocks, int persistent, hipStream_t stream) {
  if (args.final_transmittance != nullptr) {
    launch_variant<kTileSize, kBatch, true, false, true>(args, blocks, persistent, stream);
  } else if (args.statistics != nullptr) {
    if (full) {
      launch_variant<kTileSize, kBatch, true, true>(args, blocks, persistent, stream);
    } else {
      launch_variant<kTileSize, kBatch, false, true>(args, blocks, persistent, stream);
    }
  } else if (full) {
    launch_variant<kTileSize, kBatch, true, false>(args, blocks, persistent, stream);
  } else {
    launch_variant<kTileSize, kBatch, false, false>(args, blocks, persistent, stream);
  }
}

template <int kTileSize>
void launch_batch(int batch, bool full, const CompositeArgs& args, int64_t blocks, int persistent,
                  hipStream_t stream) {
  switch (batch) {
    case 64:
      launch_full_or_rgb<kTileSize, 64>(full, args, blocks, persistent, stream);
      break;
    case 128:
      launch_full_or_rgb<kTileSize, 128>(full, args, blocks, persistent, stream);
      break;
    case 512:
      launch_full_or_rgb<kTileSize, 512>(full, args, blocks, persistent, stream);
      break;
    default:
      launch_full_or_rgb<kTileSize, 256>(full, args, blocks, persistent, stream);
      break;
  }
}

}  // namespace

int launch_tile_order(const CameraParams* table, const int32_t* tile_starts, const int32_t* tile_ends,
                      int active_count, int uniform_tiles, int64_t blocks, int32_t* order, hipStream_t stream) {
  if (blocks <= 0 || active_count <= 0) return 0;
  hipLaunchKernelGGL(tile_order_kernel, dim3(1), dim3(kOrderThreads), 0, stream, table, tile_starts, tile_ends,
                     active_count, uniform_tiles, blocks, order);
  const hipError_t error = hipGetLastError();
  if (error != hipSuccess) {
    set_error("tile order launch failed: %s", hipGetErrorString(error));
    return VGR_E_LAUNCH;
  }
  return 0;
}

int launch_composite(const vgr_scene& scene, const vgr_cameras& cameras, const vgr_outputs& outputs,
                     const vgr_options& options, const vgr_workspace& ws, const ScratchLayout& layout,
                     const char* scratch, bool sorted_in_out, hipStream_t stream) {
  if (cameras.active_count <= 0) return 0;
  CompositeArgs args;
  args.table = reinterpret_cast<const CameraParams*>(scratch + layout.camera_table);
  args.tile_starts = ws.tile_starts;
  args.tile_ends = ws.tile_ends;
  args.sorted_values = sorted_in_out ? ws.values_out : ws.values_in;
  args.records = ws.records;
  args.semantic_ids = scene.semantic_ids;
  args.rgb = outputs.rgb;
  args.depth = outputs.depth;
  args.alpha = outputs.alpha;
  args.semantic = outputs.semantic;
  args.statistics = options.statistics;
  args.visible_capacity = ws.visible_capacity;
  args.gaussian_count = scene.gaussian_count;
  args.active_count = cameras.active_count;
  args.output_srgb = options.output_srgb;
  args.semantic_min_alpha = options.semantic_min_alpha;
  args.final_transmittance = outputs.final_transmittance;
  args.last_blended = outputs.last_blended;
  args.slot_records = ws.slot_records;
  const bool front = options.ordering == VGR_ORDERING_TILE_FRONT;
  if (args.final_transmittance != nullptr && args.slot_records == nullptr && !front) {
    set_error("training forwards (transmittance checkpoints) need workspace.slot_records");
    return VGR_E_INVALID_ARGUMENT;
  }
  if (front && options.tile_size != 16) {
    set_error("tile-front ordering composites 16x16 tiles");
    return VGR_E_INVALID_ARGUMENT;
  }
  if (args.final_transmittance != nullptr &&
      (args.last_blended == nullptr || options.full_sensor_output == 0 || options.output_srgb != 0)) {
    set_error("transmittance checkpoints need last_blended, full sensor output and linear RGB");
    return VGR_E_INVALID_ARGUMENT;
  }
  int64_t blocks;
  if (cameras.widths == nullptr) {
    const int tiles_x = (options.width + options.tile_size - 1) / options.tile_size;
    const int tiles_y = (options.height + options.tile_size - 1) / options.tile_size;
    args.uniform_tiles = tiles_x * tiles_y;
    blocks = static_cast<int64_t>(cameras.active_count) * args.uniform_tiles;
  } else {
    args.uniform_tiles = 0;
    blocks = cameras.max_global_tiles;
  }
  if (blocks <= 0) return 0;
  if (blocks > 0x7fffffffLL) {
    set_error("composite grid of %lld tiles exceeds the launch limit", static_cast<long long>(blocks));
    return VGR_E_INVALID_ARGUMENT;
  }
  const bool full = options.full_sensor_output != 0;
  if (options.ray_evaluation != 0) {
    // options.tile_size is the bin size here (32 bins under 16x16 raster tiles).
    if (options.tile_size == 8) {
      launch_ray<8>(args, options, ws.ray_records, cameras.active_count, full, stream);
    } else {
      launch_ray<16>(args, options, ws.ray_records, cameras.active_count, full, stream);
    }
    const hipError_t error = hipGetLastError();
    if (error != hipSuccess) {
      set_error("composite launch failed: %s", hipGetErrorString(error));
      return VGR_E_LAUNCH;
    }
    return 0;
  }
  if (options.tile_size < 8) {
    const int64_t threads = blocks * options.tile_size * options.tile_size;
    const dim3 grid(static_cast<unsigned>((threads + 255) / 256)), block(256);
    const bool count = options.statistics != nullptr;
    if (full && count) {
      hipLaunchKernelGGL((composite_pixels_kernel<true, true>), grid, block, 0, stream, args, options.tile_size, threads);
    } else if (full) {
      hipLaunchKernelGGL((composite_pixels_kernel<true, false>), grid, block, 0, stream, args, options.tile_size,
                         threads);
    } else if (count) {
      hipLaunchKernelGGL((composite_pixels_kernel<false, true>), grid, block, 0, stream, args, options.tile_size,
                         threads);
    } else {
      hipLaunchKernelGGL((composite_pixels_kernel<false, false>), grid, block, 0, stream, args, options.tile_size,
                         threads);
    }
    const hipError_t error = hipGetLastError();
    if (error != hipSuccess) {
      set_error("composite launch failed: %s", hipGetErrorString(error));
      return VGR_E_LAUNCH;
    }
    return 0;
  }
  args.total_items = blocks;
  args.work_counter = reinterpret_cast<uint32_t*>(const_cast<char*>(scratch) + layout.work_counter);
  // Longest tiles first pays when the grid is a few waves deep per SIMD; on deep
  // grids the tail is short already.
  args.tile_order = nullptr;
  if (blocks <= kOrderedCompositeBlocks) {
    int32_t* order = reinterpret_cast<int32_t*>(const_cast<char*>(scratch) + layout.tile_order);
    const int order_status = launch_tile_order(args.table, args.tile_starts, args.tile_ends, args.active_count,
                                               args.uniform_tiles, blocks, order, stream);
    if (order_status != 0) return order_status;
    args.tile_order = order;
  }
  const int persistent = options.persistent_composite;
  if (persistent > 0) launch_fill_u32(args.work_counter, 1, 0u, stream);
  if (front) {
    FrontCompositeArgs front_args;
    static_cast<CompositeArgs&>(front_args) = args;
    make_projection_args(scene, cameras, options, layout, scratch, front_args.projection);
    uint32_t* open_state = reinterpret_cast<uint32_t*>(const_cast<char*>(scratch) + layout.front_open_state);
    int32_t* open_tiles = reinterpret_cast<int32_t*>(const_cast<char*>(scratch) + layout.front_open_tiles);
    front_args.front_lengths = reinterpret_cast<const int32_t*>(scratch + layout.front_lengths);
    front_args.open_tiles = open_tiles;
    front_args.open_count = open_state + 0;
    front_args.item_count = nullptr;
    launch_fill_u32(open_state, 4, 0u, stream);
    launch_projected(scene.sh_degree, full, front_args, blocks, persistent, stream);
    // Tiles left open at the end of a truncated list: full sort, then composite
    // them again from their complete lists (fixed grids that exit at once when
    // no tile is open).
    constexpr int kFallbackWorkgroups = 256;
    const int status = launch_front_fallback_sort(scene, ws, layout, const_cast<char*>(scratch), kFallbackWorkgroups,
                                                  stream);
    if (status != 0) return status;
    FrontCompositeArgs again = front_args;
    again.open_tiles = nullptr;
    again.front_lengths = nullptr;
    again.tile_order = open_tiles;
    again.item_count = open_state + 0;
    again.work_counter = open_state + 2;
    launch_projected(scene.sh_degree, full, again, blocks, kFallbackWorkgroups, stream);
    if (options.statistics != nullptr) launch_copy_u32_to_i64(open_state + 0, options.statistics + 3, stream);
  } else if (options.tile_size == 8) {
    launch_batch<8>(options.raster_batch, full, args, blocks, persistent, stream);
  } else {
    launch_batch<16>(options.raster_batch, full, args, blocks, persistent, stream);
  }
  const hipError_t error = hipGetLastError();
  if (error != hipSuccess) {
    set_error("composite launch failed: %s", hipGetErrorString(error));
    return VGR_E_LAUNCH;
  }
  return 0;
}

}  // namespace vgr
