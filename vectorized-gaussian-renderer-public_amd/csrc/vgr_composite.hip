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
#include "vgr_composite_device.h"

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
constexpr int64_t kOrderedCompositeBlocks = 32768;  // forward grids up to this many tiles are ordered

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
};

// Renders one (active camera, tile) work item with the whole workgroup.  With
// kCheckpoint the final transmittance and the list position one past the last
// blended splat are stored per pixel for the backward pass.
template <int kTileSize, int kBatch, bool kFull, bool kCount, bool kCheckpoint>
__device__ __forceinline__ void composite_tile(const CompositeArgs& args, int64_t block) {
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
    gather_batch<kThreads, kBatch, kPerThread>(args.sorted_values, args.records, args.visible_capacity, range_start,
                                               range_end, lane, next_a, next_b, next_c);
    for (int64_t batch_start = range_start; batch_start < range_end; batch_start += kBatch) {
      const int count = static_cast<int>(range_end - batch_start < kBatch ? range_end - batch_start : kBatch);
      __syncthreads();  // previous batch fully consumed
      build_wave_lists<kTileSize, kBatch>(s, next_a, next_b, next_c, count, lane, rect_x0, rect_x1, rect_y_base);
      if (batch_start + kBatch < range_end) {
        gather_batch<kThreads, kBatch, kPerThread>(args.sorted_values, args.records, args.visible_capacity,
                                                   batch_start + kBatch, range_end, lane, next_a, next_b, next_c);
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

// Minimum waves per SIMD for 16x16 tiles (launch-bounds occupancy target): eight
// means at most 64 VGPRs.  The blend loop is latency-bound, so waves in flight
// pay more than registers.  The evaluation-counting and persistent variants
// need a few more registers and take seven and six (no spills).
template <int kTileSize, int kBatch, bool kCount, bool kPersistent = false>
constexpr int kCompositeWavesPerSimd =
    kTileSize != 16 ? 1 : (kBatch >= 512 ? 4 : (kPersistent ? 6 : (kCount ? 7 : 8)));  // 512: LDS-limited

template <int kTileSize, int kBatch, bool kFull, bool kCount, bool kCheckpoint>
__global__ __launch_bounds__(kTileSize* kTileSize, (kCompositeWavesPerSimd<kTileSize, kBatch, kCount>)) void composite_kernel(
    CompositeArgs args) {
  const int64_t block = args.tile_order != nullptr ? args.tile_order[blockIdx.x] : blockIdx.x;
  composite_tile<kTileSize, kBatch, kFull, kCount, kCheckpoint>(args, block);
}

// Persistent form: a fixed grid whose workgroups pull (camera, tile) items from
// a device counter until the queue is empty; every workgroup exits once the
// counter passes the item count.
template <int kTileSize, int kBatch, bool kFull, bool kCount, bool kCheckpoint>
__global__ __launch_bounds__(kTileSize* kTileSize, (kCompositeWavesPerSimd<kTileSize, kBatch, kCount, true>)) void
composite_persistent_kernel(CompositeArgs args) {
  __shared__ int64_t s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = static_cast<int64_t>(atomicAdd(args.work_counter, 1u));
    __syncthreads();
    const int64_t item = s_item;
    __syncthreads();
    if (item >= args.total_items) break;
    composite_tile<kTileSize, kBatch, kFull, kCount, kCheckpoint>(args,
                                                                  args.tile_order != nullptr ? args.tile_order[item]
                                                                                             : item);
  }
}

template <int kTileSize, int kBatch, bool kFull, bool kCount, bool kCheckpoint = false>
void launch_variant(const CompositeArgs& args, int64_t blocks, int persistent, hipStream_t stream) {
  const dim3 block(kTileSize * kTileSize);
  if (persistent > 0) {
    const int64_t grid = blocks < persistent ? blocks : persistent;
    hipLaunchKernelGGL((composite_persistent_kernel<kTileSize, kBatch, kFull, kCount, kCheckpoint>),
                       dim3(static_cast<unsigned>(grid)), block, 0, stream, args);
  } else {
    hipLaunchKernelGGL((composite_kernel<kTileSize, kBatch, kFull, kCount, kCheckpoint>), dim3(static_cast<unsigned>(blocks)),
                       block, 0, stream, args);
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
  if (options.tile_size == 8) {
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
