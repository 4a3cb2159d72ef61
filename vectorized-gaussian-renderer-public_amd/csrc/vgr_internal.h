// Internal declarations shared by the HIP translation units of libvgr_hip.
#ifndef VGR_INTERNAL_H
#define VGR_INTERNAL_H

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/vgr.h"
#include "vgr_math.h"

namespace vgr {

constexpr int kBlockThreads = 256;
constexpr int kWave = 64;
constexpr uint64_t kSentinelKey = ~0ull;

// One row of the device camera table (256 bytes; read with uniform indices, so
// the compiler keeps it in scalar registers).  Built on the device by
// build_camera_table_kernel from the caller's camera tensors.
struct alignas(16) CameraParams {
  float env[12];        // rows 0..2 of the scene->world affine transform
  float view[12];       // rows 0..2 of the world->camera affine transform
  float rotation[9];    // R_view * R_env (reference multiply_rotation_3x3 order)
  float focal_x, focal_y, center_x, center_y;
  float limit_x_pos, limit_x_neg, limit_y_pos, limit_y_neg;
  float origin_x, origin_y, origin_z;  // camera centre in scene coordinates (SH view direction)
  int32_t camera;        // index into the batch
  int32_t scene_slot;    // -1 when the camera's scene is not registered
  int32_t width, height;
  int32_t tiles_x, tiles_y;
  uint32_t tile_base;    // global tile index of this camera's tile (0, 0)
  int32_t pad0;
  int64_t pixel_offset;  // first output pixel of this camera
  int64_t gaussian_begin, gaussian_end;  // packed Gaussian range of its scene
  int64_t active_tile_begin;  // tiles of the active cameras before this one
  int32_t pad1[2];
};
static_assert(sizeof(CameraParams) == 256, "camera table row must be 256 bytes");

// Scratch layout inside workspace.scratch (all offsets 256-byte aligned).
struct ScratchLayout {
  size_t camera_table;     // CameraParams[active_count]
  size_t block_counts;     // int64[2 * projection_waves]: (visible, intersections) per projection wave
  size_t block_bases;      // int64[2 * (projection_waves + scan partials)]: exclusive bases per wave
  size_t wave_masks;       // uint64[projection_waves]: group cameras with a visible pair in the wave
  size_t sort_histogram;   // uint32[passes * 256]
  size_t sort_digit_base;  // uint32[passes * 256]
  size_t sort_tickets;     // uint32[passes]
  size_t sort_flags;       // uint32[passes * sort_tiles * 256]
  size_t cull_counts;      // int64[2 * (projection_blocks + 1)]   (chunk culling only)
  size_t cull_bases;       // int64[2 * (projection_blocks + scan_blocks + 1)]
  size_t cull_masks;       // uint64[projection_blocks]: camera mask per candidate block
  size_t item_blocks;      // int32[projection_blocks]: kept block ids in order
  size_t item_masks;       // uint64[projection_blocks]
  size_t item_totals;      // int64[2]: kept item count
  size_t work_counter;     // uint32[4]: persistent compositor queue
  size_t tile_order;       // int32[global_tiles]: backward tile launch order (longest lists first)
  // Tile-front ordering only (zero-sized otherwise).
  size_t front_counts;       // int64[2 * global_tiles]: (entries, 0) per tile
  size_t front_bases;        // int64[2 * (global_tiles + scan partials)]: exclusive bases per tile
  size_t front_totals;       // int64[2]
  size_t front_cursors;      // uint64[global_tiles]: next free slot of each bucket
  size_t front_lengths;      // int32[global_tiles]: bucket sizes (clamped to the capacity)
  size_t front_block_masks;  // uint64[projection_blocks]: group cameras seen by each workgroup
  size_t front_wave_masks;   // uint64[projection_blocks * waves]: group cameras seen by each wave
  size_t front_block_visible;  // int64[projection_blocks]: visible pairs of each workgroup
  size_t front_open_state;   // uint32[4]: open-tile count, fallback-sort and recomposite queues
  size_t front_open_tiles;   // int32[global_tiles]: tiles to sort in full and composite again
  // Dense (tile size 1) depth cutoff only (zero-sized otherwise).
  size_t dense_camera_map;   // int32[batch]: camera -> active index
  size_t dense_cutoff;       // int32[pixels]: last depth bucket emitted per pixel (-1 while open)
  size_t dense_accumulated;  // uint64[pixels]: thickness of the buckets before the crossing column
  size_t dense_crossing;     // int32[pixels]: column where the running thickness reaches the threshold
  size_t dense_block_cutoff;    // int32 per 8x8 pixel block (indexed from the camera's first pixel): largest cutoff
  size_t dense_block_crossing;  // int32[2] per 8x8 pixel block: smallest and largest crossing column of its pixels
  size_t dense_tau;          // uint64[(columns + fine) * pixels]: coarse column sums, then fine bucket sums, column-major
  int64_t dense_pixels;
  int32_t dense_fine;        // buckets per coarse column
  int32_t dense_columns;     // coarse columns: ceil(buckets / dense_fine)
  size_t total;
  bool culling;
  bool front;
  int64_t front_group_tiles;   // LDS histogram size of the count kernel
  int64_t front_camera_tiles;  // LDS tables of the write kernel
  int64_t projection_blocks;
  int64_t projection_chunks;
  int32_t projection_groups;
  int32_t cameras_per_block;
  int64_t sort_tiles;
  int32_t sort_tile;        // keys per sort tile
  int32_t sort_passes;
  int32_t sort_digit_bits;
  int32_t sort_key_bits;
  int32_t tile_bits;
  int32_t depth_bits;
  uint32_t depth_base;     // float bits of the near plane
};

constexpr int kProjectionWaves = kBlockThreads / kWave;  // waves of a projection workgroup
// Compositing grids up to this many tiles run in a longest-list-first order
// (ScratchLayout::tile_order), which the backward of a checkpointed forward reuses.
constexpr int64_t kOrderedCompositeBlocks = 32768;
constexpr int kSortThreads = 512;      // threads per onesweep workgroup (8 waves)
constexpr int kDefaultSortItems = 16;  // keys per thread per sort tile (tile = kSortThreads * items keys)

// Host helpers implemented in vgr_capi.hip.
void set_error(const char* fmt, ...);

// Launch wrappers (each returns 0 or sets the error and returns nonzero).
int launch_camera_table(const vgr_scene& scene, const vgr_cameras& cameras,
                        const vgr_options& options, CameraParams* table,
                        hipStream_t stream);
int launch_projection(const vgr_scene& scene, const vgr_cameras& cameras,
                      const vgr_options& options, const vgr_workspace& ws,
                      const ScratchLayout& layout, char* scratch,
                      hipStream_t stream);
int launch_sort(const vgr_workspace& ws, const ScratchLayout& layout, char* scratch, uint32_t spin_limit,
                hipStream_t stream, int* result_in_out_buffer, bool restore_id_order);
int launch_ranges(const vgr_workspace& ws, int64_t global_tiles, bool sorted_in_out,
                  hipStream_t stream);
int launch_composite(const vgr_scene& scene, const vgr_cameras& cameras,
                     const vgr_outputs& outputs, const vgr_options& options,
                     const vgr_workspace& ws, const ScratchLayout& layout,
                     const char* scratch, bool sorted_in_out, hipStream_t stream);
int launch_backward(const vgr_scene& scene, const vgr_cameras& cameras, const vgr_outputs& outputs,
                    const vgr_options& options, const vgr_workspace& ws, const ScratchLayout& layout,
                    const char* scratch, const vgr_gradients& grads, hipStream_t stream);
int launch_covariance_backward(const float* scales, const float* rotations, const float* grad_covariances,
                               int64_t count, float* grad_scales, float* grad_rotations, hipStream_t stream);
// Writes the (camera, tile) compositing blocks [0, blocks) into `order` sorted by
// descending tile-list length, so the longest tiles are launched first.
int launch_tile_order(const CameraParams* table, const int32_t* tile_starts, const int32_t* tile_ends,
                      int active_count, int uniform_tiles, int64_t blocks, int32_t* order, hipStream_t stream);
// Tile-front ordering (vgr_front.hip): bucketed emission, front sort, fallback full sort.
int launch_front_projection(const vgr_scene& scene, const vgr_cameras& cameras, const vgr_options& options,
                            const vgr_workspace& ws, const ScratchLayout& layout, char* scratch, int64_t global_tiles,
                            hipStream_t stream);
int launch_front_sort(const vgr_scene& scene, const vgr_workspace& ws, const ScratchLayout& layout, char* scratch,
                      int64_t global_tiles, hipStream_t stream);
int launch_front_fallback_sort(const vgr_scene& scene, const vgr_workspace& ws, const ScratchLayout& layout,
                               char* scratch, int workgroups, hipStream_t stream);
// Device-wide exclusive scan of n int64 pairs (vgr_project.hip); partials hold
// 2 * ceil(n / 2048) entries.  Totals go to `totals` and/or the counter block.
void scan_pairs(const int64_t* counts, int64_t n, int64_t* bases, int64_t* partials, int64_t visible_capacity,
                int64_t intersection_capacity, int64_t* counters, int64_t* totals, hipStream_t stream);
int launch_fill_i64(int64_t* data, int64_t n, int64_t value, hipStream_t stream);
int launch_dense_cutoff(const vgr_cameras& cameras, const vgr_options& options, const vgr_workspace& ws,
                        const ScratchLayout& layout, char* scratch, hipStream_t stream);
int launch_pose_compose(const float* twists, const float* base, int64_t count, float* views, hipStream_t stream);
int launch_pose_compose_backward(const float* twists, const float* base, const float* grad_views, int64_t count,
                                 float* grad_twists, float* grad_base, hipStream_t stream);
int launch_copy_u32_to_i64(const uint32_t* source, int64_t* target, hipStream_t stream);
int launch_fill_u32(uint32_t* data, int64_t n, uint32_t value, hipStream_t stream);
int launch_fill_f32(float* data, int64_t n, float value, hipStream_t stream);
int launch_copy_i64(const int64_t* source, int64_t* target, hipStream_t stream);

}  // namespace vgr

#endif  // VGR_INTERNAL_H
