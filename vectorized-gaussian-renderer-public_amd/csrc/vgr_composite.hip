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
#include "vgr_internal.h"

namespace vgr {
namespace {

// Native 4-wide vector (first-class in LLVM, so per-thread arrays of it stay in
// registers; HIP's float4 wrapper struct defeats scalar replacement).
typedef float f32x4 __attribute__((ext_vector_type(4)));

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
  int64_t total_items;     // (camera, tile) items in the launch
  int active_count;
  int uniform_tiles;     // tiles per camera when every camera has the same resolution, else 0
  int output_srgb;
  float semantic_min_alpha;
};

__device__ __forceinline__ int find_active_camera(const CameraParams* table, int active_count, int64_t block) {
  int lo = 0, hi = active_count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (table[mid].active_tile_begin <= block) {
      lo = mid;
    } else {
      hi = mid - 1;
    }
  }
  return lo;
}

// Loads the records of one batch of the sorted list into registers.
template <int kThreads, int kBatch, int kPerThread>
__device__ __forceinline__ void gather_batch(const CompositeArgs& args, int64_t batch_start, int64_t range_end,
                                             int lane, f32x4 (&a)[kPerThread], f32x4 (&b)[kPerThread],
                                             f32x4 (&c)[kPerThread]) {
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) {
    const int64_t index = batch_start + k * kThreads + lane;
    if (k * kThreads + lane < kBatch && index < range_end) {
      const int32_t record = args.sorted_values[index];
      if (record >= 0 && record < args.visible_capacity) {
        const f32x4* src = reinterpret_cast<const f32x4*>(args.records + record);
        a[k] = src[0];
        b[k] = src[1];
        c[k] = src[2];
      } else {
        // Only reachable with a corrupted list: contributes nothing (opacity 0).
        a[k] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        b[k] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        c[k] = f32x4{0.0f, 0.0f, __int_as_float(-1), 0.0f};
      }
    }
  }
}

// Margin subtracted from the per-Gaussian log-threshold: far larger than the
// error of vgr_logf/vgr_expf and of the alpha product, so "power below the
// threshold" implies the exact alpha test fails.
constexpr float kThresholdMargin = 1.0e-4f;

// Staged form of one record.  The power of the reference,
//   -0.5 * (c0*dx*dx + 2*c1*dx*dy + c2*dy*dy),
// is evaluated as (A*dx*dx + B*dx*dy) + C*dy*dy with A = -0.5*c0, B = -c1,
// C = -0.5*c2: scaling by powers of two commutes with rounding, so the value is
// bit-identical and two multiplies cheaper.
struct StagedSplat {
  f32x4 geometry;  // mean_x, mean_y, A, B
  f32x4 shading;   // C, opacity, power threshold, depth
  f32x4 colour;    // red, green, blue, gaussian id bits
};

__device__ __forceinline__ StagedSplat stage_record(const f32x4& a, const f32x4& b, const f32x4& c) {
  StagedSplat s;
  const float opacity = b.y;
  // power < threshold  =>  min(0.99, opacity * exp(power)) < 1/255 exactly.
  const float threshold = opacity > 0.0f ? vgr_logf(VGR_ALPHA_THRESHOLD / opacity) - kThresholdMargin
                                         : __int_as_float(0x7f800000);
  s.geometry = f32x4{a.x, a.y, -0.5f * a.z, -a.w};
  s.shading = f32x4{-0.5f * b.x, opacity, threshold, b.z};
  s.colour = f32x4{b.w, c.x, c.y, c.z};
  return s;
}

// Half extents of a box containing every pixel offset whose float-evaluated
// power can reach the threshold: the ellipse d^T C d <= R^2 (R^2 = -2 threshold)
// inflated by a bound on the rounding error of the power expression, which
// grows with the conic's condition.  Returns false when no finite box is
// certain (non-positive-definite or extremely anisotropic conic).
__device__ __forceinline__ bool splat_extent(const f32x4& a, const f32x4& b, float threshold, float& half_x,
                                             float& half_y) {
  const double c0 = a.z, c1 = a.w, c2 = b.x;
  const double det = c0 * c2 - c1 * c1;
  if (!(det > 0.0) || !(threshold < 0.0f)) return false;
  const double sxx = c2 / det;
  const double syy = c0 / det;
  const double condition = (fabs(c0) + 2.0 * fabs(c1) + fabs(c2)) * (sxx + syy);
  if (!(condition < 1.0e5)) return false;
  const double r2 = -2.0 * static_cast<double>(threshold) * (1.0 + 1.0e-6 * condition) * 1.001 + 1.0e-6;
  half_x = static_cast<float>(sqrt(r2 * sxx)) + 1.0e-3f;
  half_y = static_cast<float>(sqrt(r2 * syy)) + 1.0e-3f;
  return true;
}

// Renders one (active camera, tile) work item with the whole workgroup.
template <int kTileSize, int kBatch, bool kFull, bool kCount>
__device__ __forceinline__ void composite_tile(const CompositeArgs& args, int64_t block) {
  constexpr int kThreads = kTileSize * kTileSize;
  constexpr int kWaves = kThreads / kWave;
  constexpr int kRowsPerWave = kWave / kTileSize;
  constexpr int kPerThread = (kBatch + kThreads - 1) / kThreads;
  static_assert(kThreads % kWave == 0, "tile must be a whole number of waves");
  __shared__ f32x4 s_geometry[kBatch];
  __shared__ f32x4 s_shading[kBatch];
  __shared__ f32x4 s_colour[kBatch];
  __shared__ uint16_t s_list[kWaves][kBatch];  // per-wave ordered sub-lists of batch slots
  __shared__ uint32_t s_totals[kPerThread][kWaves][kWaves];  // [slice][source wave][list]

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
  const int wave_lane = lane & (kWave - 1);
  const int pixel_x = tile_x * kTileSize + lane % kTileSize;
  const int pixel_y = tile_y * kTileSize + lane / kTileSize;
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

  const int64_t range_start = args.tile_starts[global_tile];
  const int64_t range_end = args.tile_ends[global_tile];
  if (range_start >= 0 && range_end > range_start) {
    f32x4 next_a[kPerThread] = {}, next_b[kPerThread] = {}, next_c[kPerThread] = {};
    gather_batch<kThreads, kBatch, kPerThread>(args, range_start, range_end, lane, next_a, next_b, next_c);
    for (int64_t batch_start = range_start; batch_start < range_end; batch_start += kBatch) {
      const int count = static_cast<int>(range_end - batch_start < kBatch ? range_end - batch_start : kBatch);
      __syncthreads();  // previous batch fully consumed
      uint32_t wave_bits[kPerThread];
#pragma unroll
      for (int k = 0; k < kPerThread; ++k) {
        const int slot = k * kThreads + lane;
        wave_bits[k] = 0u;
        if (slot < count) {
          const StagedSplat staged = stage_record(next_a[k], next_b[k], next_c[k]);
          s_geometry[slot] = staged.geometry;
          s_shading[slot] = staged.shading;
          s_colour[slot] = staged.colour;
          float half_x, half_y;
          if (splat_extent(next_a[k], next_b[k], staged.shading.z, half_x, half_y)) {
            const float mx = next_a[k].x, my = next_a[k].y;
            const float gap_x = fabsf(mx - fminf(fmaxf(mx, rect_x0), rect_x1));
            if (gap_x <= half_x) {
#pragma unroll
              for (int w = 0; w < kWaves; ++w) {
                const float y0 = rect_y_base + static_cast<float>(w * kRowsPerWave);
                const float y1 = y0 + static_cast<float>(kRowsPerWave - 1);
                const float gap_y = fabsf(my - fminf(fmaxf(my, y0), y1));
                if (gap_y <= half_y) wave_bits[k] |= 1u << w;
              }
            }
          } else {
            wave_bits[k] = (1u << kWaves) - 1u;  // no certain box: every wave evaluates it
          }
        }
      }
      // Ordered compaction of each wave's sub-list.  Slots are ordered by
      // (slice k, source wave, lane), which is batch-slot order = depth order.
      const uint64_t below = wave_lane == 0 ? 0ull : (~0ull >> (kWave - wave_lane));
      uint32_t prefix[kPerThread][kWaves];
#pragma unroll
      for (int k = 0; k < kPerThread; ++k) {
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
          const uint64_t vote = __ballot((wave_bits[k] >> w) & 1u);
          prefix[k][w] = static_cast<uint32_t>(__popcll(vote & below));
          if (wave_lane == 0) s_totals[k][wave][w] = static_cast<uint32_t>(__popcll(vote));
        }
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < kPerThread; ++k) {
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
          if ((wave_bits[k] >> w) & 1u) {
            uint32_t position = prefix[k][w];
            for (int kk = 0; kk < k; ++kk)
              for (int q = 0; q < kWaves; ++q) position += s_totals[kk][q][w];
            for (int q = 0; q < wave; ++q) position += s_totals[k][q][w];
            s_list[w][position] = static_cast<uint16_t>(k * kThreads + lane);
          }
        }
      }
      __syncthreads();
      if (batch_start + kBatch < range_end) {
        gather_batch<kThreads, kBatch, kPerThread>(args, batch_start + kBatch, range_end, lane, next_a, next_b,
                                                   next_c);
      }
      if (!done) {
        uint32_t list_count = 0;
#pragma unroll
        for (int k = 0; k < kPerThread; ++k)
#pragma unroll
          for (int q = 0; q < kWaves; ++q) list_count += s_totals[k][q][wave];
        for (uint32_t i = 0; i < list_count; ++i) {
          const int item = s_list[wave][i];
          if (kCount) ++evaluations;
          const f32x4 g = s_geometry[item];
          const float delta_x = px - g.x;
          const float delta_y = py - g.y;
          const f32x4 sh = s_shading[item];
          const float power = g.z * delta_x * delta_x + g.w * delta_x * delta_y + sh.x * delta_y * delta_y;
          if (power > 0.0f || power < sh.z) continue;
          const float splat_alpha = fminf(VGR_MAX_ALPHA, sh.y * vgr_expf(power));
          if (splat_alpha < VGR_ALPHA_THRESHOLD) continue;
          const float next_transmittance = transmittance * (1.0f - splat_alpha);
          if (next_transmittance <= VGR_TRANSMITTANCE_THRESHOLD) {
            done = true;
            break;
          }
          const f32x4 col = s_colour[item];
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
          transmittance = next_transmittance;
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
}

template <int kTileSize, int kBatch, bool kFull, bool kCount>
__global__ __launch_bounds__(kTileSize* kTileSize) void composite_kernel(CompositeArgs args) {
  composite_tile<kTileSize, kBatch, kFull, kCount>(args, blockIdx.x);
}

// Persistent form: a fixed grid whose workgroups pull (camera, tile) items from
// a device counter until the queue is empty; every workgroup exits once the
// counter passes the item count.
template <int kTileSize, int kBatch, bool kFull, bool kCount>
__global__ __launch_bounds__(kTileSize* kTileSize) void composite_persistent_kernel(CompositeArgs args) {
  __shared__ int64_t s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = static_cast<int64_t>(atomicAdd(args.work_counter, 1u));
    __syncthreads();
    const int64_t item = s_item;
    __syncthreads();
    if (item >= args.total_items) break;
    composite_tile<kTileSize, kBatch, kFull, kCount>(args, item);
  }
}

template <int kTileSize, int kBatch, bool kFull, bool kCount>
void launch_variant(const CompositeArgs& args, int64_t blocks, int persistent, hipStream_t stream) {
  const dim3 block(kTileSize * kTileSize);
  if (persistent > 0) {
    const int64_t grid = blocks < persistent ? blocks : persistent;
    hipLaunchKernelGGL((composite_persistent_kernel<kTileSize, kBatch, kFull, kCount>),
                       dim3(static_cast<unsigned>(grid)), block, 0, stream, args);
  } else {
    hipLaunchKernelGGL((composite_kernel<kTileSize, kBatch, kFull, kCount>), dim3(static_cast<unsigned>(blocks)),
                       block, 0, stream, args);
  }
}

template <int kTileSize, int kBatch>
void launch_full_or_rgb(bool full, const CompositeArgs& args, int64_t blocks, int persistent, hipStream_t stream) {
  if (args.statistics != nullptr) {
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
