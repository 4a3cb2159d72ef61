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
// This build's form: the records pass writes the visible records and their
// radii; then, for each group of depth_bucket_group_size buckets, one kernel
// walks the records of the group's buckets and adds their per-pixel thickness
// into a [pixels x group] table of 1/65536 fixed-point integers (integer adds
// commute, so the sums and hence the cutoffs do not depend on the order in
// which the GPU performs them; the reference sums floats with atomics), and a
// per-pixel kernel advances each open pixel's running sum over the group's
// buckets, records its cutoff bucket and clears its row for the next group.
// Pixels that have a cutoff are skipped by later groups.  The emission pass
// then drops pixel intersections whose bucket lies behind the pixel's cutoff.
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

// Thickness of the records whose depth bucket lies in [group_start, group_start + group_size).
__global__ __launch_bounds__(kDenseThreads) void dense_tau_kernel(
    const vgr_record* __restrict__ records, const int32_t* __restrict__ radii, const int64_t* __restrict__ counters,
    int64_t visible_capacity, const CameraParams* __restrict__ table, const int32_t* __restrict__ camera_map,
    int batch, float near_plane, float far_plane, int buckets, int group_start, int group_size,
    const int32_t* __restrict__ cutoff, uint32_t* __restrict__ tau) {
  const int64_t visible = counters[VGR_COUNTER_VISIBLE] < visible_capacity ? counters[VGR_COUNTER_VISIBLE]
                                                                          : visible_capacity;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; r < visible; r += stride) {
    const float4* row = reinterpret_cast<const float4*>(records + r);
    const float4 a = row[0], b = row[1], c = row[2];
    const int local = depth_bucket(b.z, near_plane, far_plane, buckets) - group_start;
    if (local < 0 || local >= group_size) continue;
    const int camera = __float_as_int(c.w);
    if (camera < 0 || camera >= batch) continue;
    const int active = camera_map[camera];
    if (active < 0) continue;
    const CameraParams& cam = table[active];
    Projected p;
    p.mean_x = a.x;
    p.mean_y = a.y;
    p.conic_0 = a.z;
    p.conic_1 = a.w;
    p.conic_2 = b.x;
    p.opacity = b.y;
    p.depth = b.z;
    p.radius_x = radii[2 * r + 0];
    p.radius_y = radii[2 * r + 1];
    const TileRect rect = tile_rect(p, 1, cam.tiles_x, cam.tiles_y);
    for (int y = rect.min_y; y < rect.max_y; ++y) {
      for (int x = rect.min_x; x < rect.max_x; ++x) {
        const int64_t pixel = static_cast<int64_t>(cam.tile_base) + static_cast<int64_t>(y) * cam.tiles_x + x;
        if (cutoff[pixel] >= 0) continue;
        const float alpha = pixel_alpha(p, x, y);
        if (alpha == 0.0f) continue;
        atomicAdd(&tau[pixel * group_size + local], optical_thickness_fixed(alpha));
      }
    }
  }
}

// Running sums of the open pixels over the group's buckets; the first bucket that
// reaches the threshold becomes the pixel's cutoff (the last bucket when none
// does by the end).  The group rows are cleared for the next group.
__global__ __launch_bounds__(kDenseThreads) void dense_advance_kernel(uint32_t* __restrict__ tau,
                                                                      uint64_t* __restrict__ accumulated,
                                                                      int32_t* __restrict__ cutoff, int64_t pixels,
                                                                      int group_start, int group_count,
                                                                      int group_size, int buckets) {
  const int64_t pixel = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (pixel >= pixels || cutoff[pixel] >= 0) return;
  uint64_t sum = accumulated[pixel];
  int found = -1;
  uint32_t* row = tau + pixel * group_size;
  for (int b = 0; b < group_count; ++b) {
    sum += row[b];
    row[b] = 0u;
    if (found < 0 && sum >= kTauThresholdFixed) found = group_start + b;
  }
  if (found < 0 && group_start + group_count >= buckets) found = buckets - 1;
  accumulated[pixel] = sum;
  cutoff[pixel] = found;
}

}  // namespace

int launch_dense_cutoff(const vgr_cameras& cameras, const vgr_options& options, const vgr_workspace& ws,
                        const ScratchLayout& layout, char* scratch, hipStream_t stream) {
  const CameraParams* table = reinterpret_cast<const CameraParams*>(scratch + layout.camera_table);
  int32_t* camera_map = reinterpret_cast<int32_t*>(scratch + layout.dense_camera_map);
  int32_t* cutoff = reinterpret_cast<int32_t*>(scratch + layout.dense_cutoff);
  uint64_t* accumulated = reinterpret_cast<uint64_t*>(scratch + layout.dense_accumulated);
  uint32_t* tau = reinterpret_cast<uint32_t*>(scratch + layout.dense_tau);
  const int64_t pixels = layout.dense_pixels;
  const int buckets = options.depth_bucket_count;
  const int group_size = layout.dense_group_size;
  if (ws.radii == nullptr) {
    set_error("the depth cutoff needs workspace.radii");
    return VGR_E_INVALID_ARGUMENT;
  }
  int status = launch_fill_u32(reinterpret_cast<uint32_t*>(camera_map), cameras.batch, 0xffffffffu, stream);
  if (status == 0) status = launch_fill_u32(reinterpret_cast<uint32_t*>(cutoff), pixels, 0xffffffffu, stream);
  if (status == 0) status = launch_fill_i64(reinterpret_cast<int64_t*>(accumulated), pixels, 0, stream);
  if (status == 0) status = launch_fill_u32(tau, pixels * group_size, 0u, stream);
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
  for (int start = 0; start < buckets; start += group_size) {
    const int count = buckets - start < group_size ? buckets - start : group_size;
    hipLaunchKernelGGL(dense_tau_kernel, dim3(static_cast<unsigned>(record_blocks)), dim3(kDenseThreads), 0, stream,
                       ws.records, ws.radii, ws.counters, ws.visible_capacity, table, camera_map, cameras.batch,
                       options.near_plane, options.far_plane, buckets, start, group_size, cutoff, tau);
    if (pixel_blocks > 0) {
      hipLaunchKernelGGL(dense_advance_kernel, dim3(static_cast<unsigned>(pixel_blocks)), dim3(kDenseThreads), 0,
                         stream, tau, accumulated, cutoff, pixels, start, count, group_size, buckets);
    }
  }
  const hipError_t error = hipGetLastError();
  if (error != hipSuccess) {
    set_error("depth cutoff launch failed: %s", hipGetErrorString(error));
    return VGR_E_LAUNCH;
  }
  return 0;
}

}  // namespace vgr
