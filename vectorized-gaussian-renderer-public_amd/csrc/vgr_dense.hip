// Dense per-pixel path (tile size 1): the optical-thickness depth cutoff.
//
// Reference behaviour (renderer_cuda.cu:1402-1690, orchestration 2888-3144):
// every pixel sums, depth bucket by depth bucket, a conservative optical
// thickness of the splats that reach it, and pixel intersections in buckets
// behind the first bucket where the sum reaches 9.2103 (transmittance 1e-4)
// are not emitted.  The compositor stops before such entries anyway (the
// thickness is a lower bound of -log T), so the cutoff only shrinks the
// intersection count; images do not change.
//
// This build's form: the cutoff of a pixel is the first bucket where the
// bucket-ordered prefix sum of its thickness reaches the threshold, and integer
// sums (1/65536 fixed point, 64-bit) do not depend on the order the GPU adds
// them in, so any grouping of the buckets gives the same cutoffs (the reference
// sums floats with atomics, group by group).  Two levels instead of one pass per
// group: the records pass writes the visible records and their radii; a coarse
// pass adds every (pixel, record) thickness into [C x pixels] columns of F
// buckets each; a per-pixel scan finds the column where the running sum
// reaches the threshold; a fine pass adds only the thickness that falls in that
// column, per bucket, into [F x pixels]; a second scan finds the bucket.  The
// records are read twice and no per-group table is swept (the per-group form
// read every record and swept a [pixels x group] table once per group of
// buckets).  The emission pass then drops pixel intersections whose bucket
// lies behind the pixel's cutoff.
#include <hip/hip_runtime.h>

#include "vgr_internal.h"
#include "vgr_project_device.h"

namespace vgr {
namespace {

using dev::depth_bucket;
using dev::kTauThresholdFixed;
using dev::optical_thickness_fixed;
using dev::pixel_alpha;
using dev::Projected;
using dev::tile_rect;
using dev::TileRect;

constexpr int kDenseThreads = 256;

// camera index -> active index (cameras outside the active set stay -1).
__global__ void dense_camera_map_kernel(const CameraParams* __restrict__ table, int active_count,
                                        int32_t* __restrict__ map, int batch) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= active_count) return;
  const int camera = table[a].camera;
  if (camera >= 0 && camera < batch) map[camera] = a;
}

// One visible record per lane, rebuilt for the cooperative pixel walk: its
// camera's tile base and width, the splat, its pixel rectangle and depth bucket.
struct RecordSplat {
  bool valid;
  uint32_t tile_base;
  int tiles_x;
  int bucket;
  Projected p;
  TileRect rect;
};

__device__ __forceinline__ RecordSplat load_record_splat(const vgr_record* __restrict__ records,
                                                         const int32_t* __restrict__ radii, int64_t r,
                                                         int64_t visible, const CameraParams* __restrict__ table,
                                                         const int32_t* __restrict__ camera_map, int batch,
                                                         float near_plane, float far_plane, int buckets) {
  RecordSplat s;
  s.valid = false;
  s.tile_base = 0;
  s.tiles_x = 0;
  s.bucket = 0;
  if (r >= visible) return s;
  const float4* row = reinterpret_cast<const float4*>(records + r);
  const float4 a = row[0], b = row[1], c = row[2];
  const int camera = __float_as_int(c.w);
  if (camera < 0 || camera >= batch) return s;
  const int active = camera_map[camera];
  if (active < 0) return s;
  const CameraParams& cam = table[active];
  s.p.mean_x = a.x;
  s.p.mean_y = a.y;
  s.p.conic_0 = a.z;
  s.p.conic_1 = a.w;
  s.p.conic_2 = b.x;
  s.p.opacity = b.y;
  s.p.depth = b.z;
  s.p.radius_x = radii[2 * r + 0];
  s.p.radius_y = radii[2 * r + 1];
  s.rect = tile_rect(s.p, 1, cam.tiles_x, cam.tiles_y);
  s.tile_base = cam.tile_base;
  s.tiles_x = cam.tiles_x;
  s.bucket = depth_bucket(b.z, near_plane, far_plane, buckets);
  s.valid = true;
  return s;
}

__device__ __forceinline__ int64_t visible_records(const int64_t* counters, int64_t visible_capacity) {
  return counters[VGR_COUNTER_VISIBLE] < visible_capacity ? counters[VGR_COUNTER_VISIBLE] : visible_capacity;
}

// Coarse level: thickness per (column of F buckets, pixel), column-major so the
// lanes of a wave, which walk neighbouring pixels of one record (one column),
// add into neighbouring words of few cache lines.  Each wave takes 64 records at
// a time and walks their pixels together.
__global__ __launch_bounds__(kDenseThreads) void dense_coarse_kernel(
    const vgr_record* __restrict__ records, const int32_t* __restrict__ radii, const int64_t* __restrict__ counters,
    int64_t visible_capacity, const CameraParams* __restrict__ table, const int32_t* __restrict__ camera_map,
    int batch, float near_plane, float far_plane, int buckets, int fine, int64_t pixels,
    unsigned long long* __restrict__ coarse) {
  const int64_t visible = visible_records(counters, visible_capacity);
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t first = static_cast<int64_t>(blockIdx.x) * blockDim.x + (threadIdx.x - lane); first < visible;
       first += stride) {  // uniform over the wave
    const RecordSplat s = load_record_splat(records, radii, first + lane, visible, table, camera_map, batch,
                                            near_plane, far_plane, buckets);
    dense_wave_pixels(s.valid, s.p, s.rect, s.tile_base, s.tiles_x, s.bucket, nullptr,
                      [&](bool keep, int, int owner, uint32_t pixel, float alpha) {
                        const int column = __shfl(s.bucket, owner, kWave) / fine;
                        if (keep) {
                          atomicAdd(&coarse[static_cast<int64_t>(column) * pixels + pixel],
                                    static_cast<unsigned long long>(optical_thickness_fixed(alpha)));
                        }
                      });
  }
}

// Per pixel: the column where the running sum reaches the threshold and the sum
// before it; a pixel that never reaches it keeps every bucket.
__global__ __launch_bounds__(kDenseThreads) void dense_coarse_scan_kernel(
    const unsigned long long* __restrict__ coarse, int64_t pixels, int columns, int buckets,
    int32_t* __restrict__ crossing, uint64_t* __restrict__ before, int32_t* __restrict__ cutoff) {
  const int64_t pixel = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (pixel >= pixels) return;
  uint64_t sum = 0;
  int found = -1;
  for (int c = 0; c < columns; ++c) {
    const uint64_t next = sum + coarse[static_cast<int64_t>(c) * pixels + pixel];
    if (next >= kTauThresholdFixed) {
      found = c;
      break;
    }
    sum = next;
  }
  crossing[pixel] = found;
  before[pixel] = sum;
  cutoff[pixel] = buckets - 1;  // refined by the fine scan where a column crosses
}

// Per 8x8 pixel block of every active camera: the largest cutoff of its pixels
// (kCrossing false), or the smallest and largest crossing column of its pixels
// that cross (kCrossing true; a block with none keeps smallest > largest).  One
// thread per block, found from a global pixel index: the camera whose pixel
// range holds it (binary search over the ascending first pixels), then the
// block at that offset if the camera has that many blocks.
template <bool kCrossing>
__global__ __launch_bounds__(kDenseThreads) void dense_block_bounds_kernel(
    const CameraParams* __restrict__ table, int active_count, int64_t pixels, const int32_t* __restrict__ per_pixel,
    int32_t* __restrict__ bounds) {
  const int64_t index = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (index >= pixels || active_count <= 0) return;
  int lo = 0, hi = active_count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (static_cast<int64_t>(table[mid].tile_base) <= index) {
      lo = mid;
    } else {
      hi = mid - 1;
    }
  }
  const CameraParams& cam = table[lo];
  const int64_t local = index - static_cast<int64_t>(cam.tile_base);
  const int blocks_x = dev::dense_blocks_x(cam.tiles_x);
  const int blocks_y = (cam.tiles_y + 7) >> 3;
  if (cam.camera < 0 || local < 0 || local >= static_cast<int64_t>(blocks_x) * blocks_y) return;
  const int bx = static_cast<int>(local % blocks_x), by = static_cast<int>(local / blocks_x);
  int smallest = 0x7fffffff, largest = -1;
  for (int y = by * 8; y < min(by * 8 + 8, cam.tiles_y); ++y) {
    for (int x = bx * 8; x < min(bx * 8 + 8, cam.tiles_x); ++x) {
      const int v = per_pixel[cam.tile_base + static_cast<uint32_t>(y * cam.tiles_x + x)];
      if (kCrossing && v < 0) continue;  // the pixel never reaches the threshold
      smallest = min(smallest, v);
      largest = max(largest, v);
    }
  }
  if (kCrossing) {
    bounds[2 * index + 0] = smallest;
    bounds[2 * index + 1] = largest;
  } else {
    bounds[index] = largest;
  }
}

// Fine level: thickness per (bucket of the pixel's crossing column, pixel),
// column-major like the coarse level.
__global__ __launch_bounds__(kDenseThreads) void dense_fine_kernel(
    const vgr_record* __restrict__ records, const int32_t* __restrict__ radii, const int64_t* __restrict__ counters,
    int64_t visible_capacity, const CameraParams* __restrict__ table, const int32_t* __restrict__ camera_map,
    int batch, float near_plane, float far_plane, int buckets, int fine, int64_t pixels,
    const int32_t* __restrict__ crossing, const int32_t* __restrict__ block_crossing,
    unsigned long long* __restrict__ cells) {
  const int64_t visible = visible_records(counters, visible_capacity);
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t first = static_cast<int64_t>(blockIdx.x) * blockDim.x + (threadIdx.x - lane); first < visible;
       first += stride) {  // uniform over the wave
    const RecordSplat s = load_record_splat(records, radii, first + lane, visible, table, camera_map, batch,
                                            near_plane, far_plane, buckets);
    // Splats whose column is no crossing column of any block they reach add
    // nothing here and are not walked.
    bool walk = false;
    if (s.valid) {
      const int column = s.bucket / fine;
      const int blocks_x = dev::dense_blocks_x(s.tiles_x);
      for (int by = s.rect.min_y >> 3; by <= (s.rect.max_y - 1) >> 3 && !walk; ++by) {
        for (int bx = s.rect.min_x >> 3; bx <= (s.rect.max_x - 1) >> 3; ++bx) {
          const int64_t b = static_cast<int64_t>(s.tile_base) + by * blocks_x + bx;
          if (block_crossing[2 * b] <= column && column <= block_crossing[2 * b + 1]) {
            walk = true;
            break;
          }
        }
      }
    }
    dense_wave_pixels(walk, s.p, s.rect, s.tile_base, s.tiles_x, s.bucket, nullptr,
                      [&](bool keep, int, int owner, uint32_t pixel, float alpha) {
                        const int bucket = __shfl(s.bucket, owner, kWave);
                        const int column = bucket / fine;
                        if (keep && crossing[pixel] == column) {
                          atomicAdd(&cells[static_cast<int64_t>(bucket - column * fine) * pixels + pixel],
                                    static_cast<unsigned long long>(optical_thickness_fixed(alpha)));
                        }
                      });
  }
}

__global__ __launch_bounds__(kDenseThreads) void dense_fine_scan_kernel(
    const unsigned long long* __restrict__ cells, int64_t pixels, int fine, int buckets,
    const int32_t* __restrict__ crossing, const uint64_t* __restrict__ before, int32_t* __restrict__ cutoff) {
  const int64_t pixel = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (pixel >= pixels) return;
  const int column = crossing[pixel];
  if (column < 0) return;
  uint64_t sum = before[pixel];
  for (int b = 0; b < fine && column * fine + b < buckets; ++b) {
    sum += cells[static_cast<int64_t>(b) * pixels + pixel];
    if (sum >= kTauThresholdFixed) {
      cutoff[pixel] = column * fine + b;
      return;
    }
  }
}

}  // namespace

int launch_dense_cutoff(const vgr_cameras& cameras, const vgr_options& options, const vgr_workspace& ws,
                        const ScratchLayout& layout, char* scratch, hipStream_t stream) {
  const CameraParams* table = reinterpret_cast<const CameraParams*>(scratch + layout.camera_table);
  int32_t* camera_map = reinterpret_cast<int32_t*>(scratch + layout.dense_camera_map);
  int32_t* cutoff = reinterpret_cast<int32_t*>(scratch + layout.dense_cutoff);
  uint64_t* before = reinterpret_cast<uint64_t*>(scratch + layout.dense_accumulated);
  int32_t* crossing = reinterpret_cast<int32_t*>(scratch + layout.dense_crossing);
  unsigned long long* coarse = reinterpret_cast<unsigned long long*>(scratch + layout.dense_tau);
  const int64_t pixels = layout.dense_pixels;
  const int buckets = options.depth_bucket_count;
  const int fine = layout.dense_fine;
  const int columns = layout.dense_columns;
  unsigned long long* cells = coarse + pixels * columns;
  int32_t* block_cutoff = reinterpret_cast<int32_t*>(scratch + layout.dense_block_cutoff);
  int32_t* block_crossing = reinterpret_cast<int32_t*>(scratch + layout.dense_block_crossing);
  if (ws.radii == nullptr) {
    set_error("the depth cutoff needs workspace.radii");
    return VGR_E_INVALID_ARGUMENT;
  }
  int status = launch_fill_u32(reinterpret_cast<uint32_t*>(camera_map), cameras.batch, 0xffffffffu, stream);
  if (status == 0) status = launch_fill_i64(reinterpret_cast<int64_t*>(coarse), pixels * (columns + fine), 0, stream);
  if (status != 0) {
    set_error("depth cutoff fills failed");
    return VGR_E_LAUNCH;
  }
  if (cameras.active_count > 0) {
    hipLaunchKernelGGL(dense_camera_map_kernel, dim3(static_cast<unsigned>((cameras.active_count + 63) / 64)),
                       dim3(64), 0, stream, table, cameras.active_count, camera_map, cameras.batch);
  }
  // Records are walked with a grid sized for the visible capacity (the count is on the device).
  int64_t record_blocks = (ws.visible_capacity + kDenseThreads - 1) / kDenseThreads;
  record_blocks = record_blocks < 1 ? 1 : (record_blocks > 65536 ? 65536 : record_blocks);
  const int64_t pixel_blocks = (pixels + kDenseThreads - 1) / kDenseThreads;
  if (pixel_blocks > 0) {
    hipLaunchKernelGGL(dense_coarse_kernel, dim3(static_cast<unsigned>(record_blocks)), dim3(kDenseThreads), 0,
                       stream, ws.records, ws.radii, ws.counters, ws.visible_capacity, table, camera_map,
                       cameras.batch, options.near_plane, options.far_plane, buckets, fine, pixels, coarse);
    hipLaunchKernelGGL(dense_coarse_scan_kernel, dim3(static_cast<unsigned>(pixel_blocks)), dim3(kDenseThreads), 0,
                       stream, coarse, pixels, columns, buckets, crossing, before, cutoff);
    hipLaunchKernelGGL(dense_block_bounds_kernel<true>, dim3(static_cast<unsigned>(pixel_blocks)),
                       dim3(kDenseThreads), 0, stream, table, cameras.active_count, pixels, crossing, block_crossing);
    hipLaunchKernelGGL(dense_fine_kernel, dim3(static_cast<unsigned>(record_blocks)), dim3(kDenseThreads), 0, stream,
                       ws.records, ws.radii, ws.counters, ws.visible_capacity, table, camera_map, cameras.batch,
                       options.near_plane, options.far_plane, buckets, fine, pixels, crossing, block_crossing, cells);
    hipLaunchKernelGGL(dense_fine_scan_kernel, dim3(static_cast<unsigned>(pixel_blocks)), dim3(kDenseThreads), 0,
                       stream, cells, pixels, fine, buckets, crossing, before, cutoff);
    hipLaunchKernelGGL(dense_block_bounds_kernel<false>, dim3(static_cast<unsigned>(pixel_blocks)),
                       dim3(kDenseThreads), 0, stream, table, cameras.active_count, pixels, cutoff, block_cutoff);
  }
  const hipError_t error = hipGetLastError();
  if (error != hipSuccess) {
    set_error("depth cutoff launch failed: %s", hipGetErrorString(error));
    return VGR_E_LAUNCH;
  }
  return 0;
}

}  // namespace vgr
