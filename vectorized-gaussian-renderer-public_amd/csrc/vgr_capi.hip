// C ABI entry points of libvgr_hip: argument validation, scratch planning and
// the launch sequence of one render submission.
//
// Launch order mirrors render_cuda (renderer_cuda.cu:2536-3456) for the
// general 16x16 / 8x8 tile path: camera table -> projection count -> block
// scan (writes the counter block) -> projection write + emission -> radix
// ordering -> tile ranges -> fused compositing.  Unlike the reference there
// is no device-to-host count copy: every stage reads the counts it needs from
// the device, so one submission never synchronizes the host.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "vgr_internal.h"

namespace vgr {

namespace {
thread_local char g_error[512] = "";

size_t align_up(size_t value) { return (value + 255u) & ~static_cast<size_t>(255u); }

int bit_length_u64(uint64_t value) {
  int bits = 0;
  while (value != 0) {
    ++bits;
    value >>= 1;
  }
  return bits;
}

int64_t global_tile_count(const vgr_cameras& cameras, const vgr_options& options) {
  if (cameras.widths != nullptr || cameras.heights != nullptr) return cameras.max_global_tiles;
  const int64_t tiles_x = (options.width + options.tile_size - 1) / options.tile_size;
  const int64_t tiles_y = (options.height + options.tile_size - 1) / options.tile_size;
  return static_cast<int64_t>(cameras.batch) * tiles_x * tiles_y;
}

// Default digit width: 8 bits, or 9 where that saves a radix pass (a 44-bit key:
// 5 passes instead of 6; an 8192-key tile still averages 16 keys per digit).
int resolve_digit_bits(const vgr_options& options, int key_bits) {
  if (options.sort_digit_bits != 0) return options.sort_digit_bits;
  return (key_bits + 8) / 9 < (key_bits + 7) / 8 ? 9 : 8;
}

// Default camera group: 32 cameras per projection workgroup.  Each group re-reads
// the Gaussians' means, covariances and SH coefficients, so larger groups read
// less; 64 cameras of 640x480 over 1M SH3 Gaussians project in 1.79 ms with 32
// per group against 2.00 ms with 8 (64 gives no further gain).
int resolve_cameras_per_block(const vgr_options& options) {
  return options.cameras_per_block == 0 ? 32 : options.cameras_per_block;
}

// Exact-ray renders with 16x16 raster tiles bin their intersections in 32x32
// tiles (four raster tiles share one sorted list, cuda_backend.py:1973-1985):
// every tile count, range and capacity of such a render is in bins.
vgr_options binning_options(const vgr_options& options) {
  vgr_options binned = options;
  if (options.ray_evaluation != 0 && options.tile_size == 16) binned.tile_size = 32;
  return binned;
}

// Tile-front ordering: LDS budgets of the bucketed emission (count kernel
// histogram over a camera group's tiles; write kernel tables over one camera's
// tiles).
constexpr int64_t kFrontMaxGroupTiles = 8192;
constexpr int64_t kFrontMaxCameraTiles = 2048;

int64_t max_camera_tiles(const vgr_cameras& cameras, const vgr_options& options) {
  if (cameras.max_camera_tiles > 0) return cameras.max_camera_tiles;
  if (cameras.widths != nullptr || cameras.heights != nullptr) return cameras.max_global_tiles;
  const int64_t tiles_x = (options.width + options.tile_size - 1) / options.tile_size;
  const int64_t tiles_y = (options.height + options.tile_size - 1) / options.tile_size;
  return tiles_x * tiles_y;
}

ScratchLayout compute_layout(int32_t active_count, int64_t gaussian_count, const vgr_options& options,
                             int64_t global_tiles, int64_t intersection_capacity, int64_t camera_tiles,
                             int32_t batch) {
  ScratchLayout layout{};
  layout.cameras_per_block = resolve_cameras_per_block(options);
  layout.front = options.ordering == VGR_ORDERING_TILE_FRONT;
  if (layout.front) {
    // A camera group's tiles must fit the count kernel's LDS histogram.
    const int64_t per_camera = camera_tiles > 0 ? camera_tiles : 1;
    int64_t group = kFrontMaxGroupTiles / per_camera;
    group = group < 1 ? 1 : group;
    if (layout.cameras_per_block > group) layout.cameras_per_block = static_cast<int32_t>(group);
    layout.front_camera_tiles = per_camera < kFrontMaxCameraTiles ? per_camera : kFrontMaxCameraTiles;
    const int64_t group_tiles = per_camera * layout.cameras_per_block;
    layout.front_group_tiles = group_tiles < kFrontMaxGroupTiles ? group_tiles : kFrontMaxGroupTiles;
  }
  layout.projection_chunks = (gaussian_count + kBlockThreads - 1) / kBlockThreads;
  layout.projection_groups = (active_count + layout.cameras_per_block - 1) / layout.cameras_per_block;
  layout.projection_blocks = layout.projection_chunks * layout.projection_groups;
  const int64_t scan_blocks = (layout.projection_blocks + kBlockThreads * 8 - 1) / (kBlockThreads * 8);
  // Record and intersection bases are scanned per projection wave.
  const int64_t projection_waves = layout.projection_blocks * kProjectionWaves;
  const int64_t wave_scan_blocks = (projection_waves + kBlockThreads * 8 - 1) / (kBlockThreads * 8);

  const uint32_t near_bits = vgr_float_bits(options.near_plane);
  const uint32_t far_bits = vgr_float_bits(options.far_plane);
  layout.depth_base = near_bits;
  layout.depth_bits = bit_length_u64(static_cast<uint64_t>(far_bits - near_bits) + 1u);
  layout.tile_bits = global_tiles > 1 ? bit_length_u64(static_cast<uint64_t>(global_tiles - 1)) : 0;
  layout.sort_key_bits = layout.tile_bits + layout.depth_bits;
  layout.sort_digit_bits = resolve_digit_bits(options, layout.sort_key_bits);
  layout.sort_passes = (layout.sort_key_bits + layout.sort_digit_bits - 1) / layout.sort_digit_bits;
  layout.sort_tile = kSortThreads * (options.sort_tile_items == 0 ? kDefaultSortItems : options.sort_tile_items);
  layout.sort_tiles = (intersection_capacity + layout.sort_tile - 1) / layout.sort_tile;
  const int64_t radix = 1 << layout.sort_digit_bits;
  // Per-pass sort arrays hold passes x radix slots.  With the default digit width
  // they are sized for the larger of the 8- and 9-bit choices, which grows with the
  // key width, so scratch sized for the largest tile count fits every smaller render.
  int64_t pass_slots = static_cast<int64_t>(layout.sort_passes) * radix;
  int64_t pass_count = layout.sort_passes;
  if (options.sort_digit_bits == 0) {
    const int64_t passes8 = (layout.sort_key_bits + 7) / 8, passes9 = (layout.sort_key_bits + 8) / 9;
    pass_slots = passes8 * 256 > passes9 * 512 ? passes8 * 256 : passes9 * 512;
    pass_count = passes8;
  }

  size_t offset = 0;
  layout.camera_table = offset;
  offset = align_up(offset + sizeof(CameraParams) * static_cast<size_t>(active_count > 0 ? active_count : 1));
  layout.block_counts = offset;
  offset = align_up(offset + sizeof(int64_t) * 2 * static_cast<size_t>(projection_waves + 1));
  layout.block_bases = offset;
  offset = align_up(offset + sizeof(int64_t) * 2 * static_cast<size_t>(projection_waves + wave_scan_blocks + 1));
  layout.wave_masks = offset;
  offset = align_up(offset + sizeof(uint64_t) * static_cast<size_t>(projection_waves + 1));
  layout.sort_histogram = offset;
  offset = align_up(offset + sizeof(uint32_t) * static_cast<size_t>(pass_slots));
  layout.sort_digit_base = offset;
  offset = align_up(offset + sizeof(uint32_t) * static_cast<size_t>(pass_slots));
  layout.sort_tickets = offset;
  offset = align_up(offset + sizeof(uint32_t) * static_cast<size_t>(pass_count));
  layout.sort_flags = offset;
  const size_t flag_tiles = layout.front ? 0u : static_cast<size_t>(layout.sort_tiles);  // no radix pass
  offset = align_up(offset + sizeof(uint32_t) * static_cast<size_t>(pass_slots) * flag_tiles);
  layout.work_counter = offset;
  offset = align_up(offset + sizeof(uint32_t) * 4);
  layout.tile_order = offset;
  offset = align_up(offset + sizeof(int32_t) * static_cast<size_t>(global_tiles > 0 ? global_tiles : 1));
  layout.culling = options.chunk_culling != 0;
  const size_t blocks = static_cast<size_t>(layout.projection_blocks);
  const size_t cull_blocks = layout.culling ? blocks : 0;
  layout.cull_counts = offset;
  offset = align_up(offset + sizeof(int64_t) * 2 * (cull_blocks + 1));
  layout.cull_bases = offset;
  offset = align_up(offset + sizeof(int64_t) * 2 * (cull_blocks + (layout.culling ? scan_blocks : 0) + 1));
  layout.cull_masks = offset;
  offset = align_up(offset + sizeof(uint64_t) * (cull_blocks + 1));
  layout.item_blocks = offset;
  offset = align_up(offset + sizeof(int32_t) * (cull_blocks + 1));
  layout.item_masks = offset;
  offset = align_up(offset + sizeof(uint64_t) * (cull_blocks + 1));
  layout.item_totals = offset;
  offset = align_up(offset + sizeof(int64_t) * 2);
  const size_t front_tiles = layout.front ? static_cast<size_t>(global_tiles > 0 ? global_tiles : 1) : 0u;
  const size_t front_scan = layout.front ? static_cast<size_t>((global_tiles + kBlockThreads * 8 - 1) /
                                                               (kBlockThreads * 8)) : 0u;
  layout.front_counts = offset;
  offset = align_up(offset + sizeof(int64_t) * 2 * front_tiles);
  layout.front_bases = offset;
  offset = align_up(offset + sizeof(int64_t) * 2 * (front_tiles + front_scan + 1));
  layout.front_totals = offset;
  offset = align_up(offset + sizeof(int64_t) * 2);
  layout.front_cursors = offset;
  offset = align_up(offset + sizeof(uint64_t) * front_tiles);
  layout.front_lengths = offset;
  offset = align_up(offset + sizeof(int32_t) * front_tiles);
  layout.front_block_masks = offset;
  offset = align_up(offset + sizeof(uint64_t) * (layout.front ? blocks + 1 : 0u));
  layout.front_wave_masks = offset;
  offset = align_up(offset + sizeof(uint64_t) * (layout.front ? (blocks + 1) * (kBlockThreads / kWave) : 0u));
  layout.front_block_visible = offset;
  offset = align_up(offset + sizeof(int64_t) * (layout.front ? blocks + 1 : 0u));
  layout.front_open_state = offset;
  offset = align_up(offset + sizeof(uint32_t) * 4);
  layout.front_open_tiles = offset;
  offset = align_up(offset + sizeof(int32_t) * front_tiles);
  const bool cutoff = options.tile_size == 1 && options.depth_bucket_count > 0;
  layout.dense_pixels = cutoff ? global_tiles : 0;
  // Two-level cutoff: columns of up to 64 buckets.  The cutoffs do not depend on
  // this split (nor on depth_bucket_group_size): they are prefix sums of integers.
  layout.dense_fine = cutoff ? (options.depth_bucket_count < 64 ? options.depth_bucket_count : 64) : 0;
  layout.dense_columns = cutoff ? (options.depth_bucket_count + layout.dense_fine - 1) / layout.dense_fine : 0;
  const size_t dense_pixels = static_cast<size_t>(layout.dense_pixels);
  layout.dense_camera_map = offset;
  offset = align_up(offset + sizeof(int32_t) * (cutoff ? static_cast<size_t>(batch > 0 ? batch : 1) : 0u));
  layout.dense_cutoff = offset;
  offset = align_up(offset + sizeof(int32_t) * dense_pixels);
  layout.dense_accumulated = offset;
  offset = align_up(offset + sizeof(uint64_t) * dense_pixels);
  layout.dense_crossing = offset;
  offset = align_up(offset + sizeof(int32_t) * dense_pixels);
  layout.dense_block_cutoff = offset;
  offset = align_up(offset + sizeof(int32_t) * dense_pixels);
  layout.dense_block_crossing = offset;
  offset = align_up(offset + sizeof(int32_t) * 2 * dense_pixels);
  layout.dense_tau = offset;
  offset = align_up(offset + sizeof(uint64_t) * dense_pixels *
                                 static_cast<size_t>(layout.dense_columns + layout.dense_fine));
  layout.total = offset;
  return layout;
}

int validate_options(const vgr_options& options, const vgr_cameras& cameras) {
  if (options.tile_size <= 0 || options.tile_size > 16 || (options.tile_size & (options.tile_size - 1)) != 0) {
    set_error("tile_size must be a power of two in [1, 16]; got %d", options.tile_size);
    return VGR_E_INVALID_ARGUMENT;
  }
  const bool uniform = cameras.widths == nullptr && cameras.heights == nullptr;
  if (uniform && (options.width <= 0 || options.height <= 0)) {
    set_error("width and height must be positive; got %dx%d", options.width, options.height);
    return VGR_E_INVALID_ARGUMENT;
  }
  if (!uniform && (cameras.widths == nullptr || cameras.heights == nullptr || cameras.max_global_tiles <= 0)) {
    set_error("per-camera resolutions need widths, heights and max_global_tiles > 0");
    return VGR_E_INVALID_ARGUMENT;
  }
  if (!(options.near_plane > 0.0f) || !(options.far_plane > options.near_plane)) {
    set_error("expected 0 < near_plane < far_plane; got %g, %g", options.near_plane, options.far_plane);
    return VGR_E_INVALID_ARGUMENT;
  }
  if (!(options.gaussian_support_sigma > 0.0f) || !(options.covariance_epsilon >= 0.0f)) {
    set_error("gaussian_support_sigma must be positive and covariance_epsilon non-negative");
    return VGR_E_INVALID_ARGUMENT;
  }
  if (options.sort_digit_bits != 0 && (options.sort_digit_bits < 4 || options.sort_digit_bits > 9)) {
    set_error("sort_digit_bits must be 0 (default) or in [4, 9]; got %d", options.sort_digit_bits);
    return VGR_E_INVALID_ARGUMENT;
  }
  const int batch = options.raster_batch == 0 ? 256 : options.raster_batch;
  if (batch != 64 && batch != 128 && batch != 256 && batch != 512) {
    set_error("raster_batch must be 64, 128, 256 or 512; got %d", options.raster_batch);
    return VGR_E_INVALID_ARGUMENT;
  }
  if (options.persistent_composite < 0 || options.persistent_composite > (1 << 20)) {
    set_error("persistent_composite must be in [0, 2^20]; got %d", options.persistent_composite);
    return VGR_E_INVALID_ARGUMENT;
  }
  if (options.lookback_spin_limit < -1) {
    set_error("lookback_spin_limit must be -1 (forced give-up, tests), 0 (default) or positive; got %d",
              options.lookback_spin_limit);
    return VGR_E_INVALID_ARGUMENT;
  }
  if (options.sort_tile_items != 0 && options.sort_tile_items != 8 && options.sort_tile_items != 12 &&
      options.sort_tile_items != 16) {
    set_error("sort_tile_items must be 0 (default), 8, 12 or 16; got %d", options.sort_tile_items);
    return VGR_E_INVALID_ARGUMENT;
  }
  const int group = resolve_cameras_per_block(options);
  if (group < 1 || group > 64) {
    set_error("cameras_per_block must be in [1, 64]; got %d", options.cameras_per_block);
    return VGR_E_INVALID_ARGUMENT;
  }
  if (options.depth_bucket_count < 0 || options.depth_bucket_group_size < 0 ||
      (options.depth_bucket_count > 0 && options.tile_size != 1)) {
    set_error("depth_bucket_count (>= 0) and depth_bucket_group_size (>= 0) apply to tile_size 1; got %d, %d at "
              "tile_size %d",
              options.depth_bucket_count, options.depth_bucket_group_size, options.tile_size);
    return VGR_E_INVALID_ARGUMENT;
  }
  if (options.depth_bucket_count > (1 << 24)) {
    set_error("depth_bucket_count must be at most 2^24; got %d", options.depth_bucket_count);
    return VGR_E_INVALID_ARGUMENT;
  }
  if (options.ray_evaluation != 0) {
    if (options.tile_size != 8 && options.tile_size != 16) {
      set_error("exact-ray evaluation renders 8x8 or 16x16 raster tiles; got tile_size %d", options.tile_size);
      return VGR_E_INVALID_ARGUMENT;
    }
    if (options.antialiased != 0) {
      set_error("antialiased rasterization is incompatible with exact-ray evaluation (the ray model has no "
                "screen-space covariance dilation to compensate)");
      return VGR_E_INVALID_ARGUMENT;
    }
    if (!uniform) {
      set_error("exact-ray evaluation renders batches of one resolution (cameras.widths/heights must be NULL)");
      return VGR_E_INVALID_ARGUMENT;
    }
    if (options.ordering != VGR_ORDERING_GLOBAL_RADIX || options.depth_bucket_count != 0) {
      set_error("exact-ray evaluation uses the global radix ordering without a depth cutoff");
      return VGR_E_INVALID_ARGUMENT;
    }
  }
  if (options.ordering != VGR_ORDERING_GLOBAL_RADIX && options.ordering != VGR_ORDERING_TILE_FRONT) {
    set_error("ordering must be VGR_ORDERING_GLOBAL_RADIX (0) or VGR_ORDERING_TILE_FRONT (1); got %d",
              options.ordering);
    return VGR_E_INVALID_ARGUMENT;
  }
  if (options.ordering == VGR_ORDERING_TILE_FRONT) {
    if (options.tile_size != 16) {
      set_error("tile-front ordering needs 16x16 tiles; got %d", options.tile_size);
      return VGR_E_INVALID_ARGUMENT;
    }
    if (cameras.active_ids != nullptr) {
      set_error("tile-front ordering renders every camera of the batch (active_ids must be NULL)");
      return VGR_E_INVALID_ARGUMENT;
    }
    const int64_t camera_tiles = max_camera_tiles(cameras, options);
    if (camera_tiles > kFrontMaxCameraTiles) {
      set_error("tile-front ordering supports cameras of at most %lld tiles; this batch has %lld",
                static_cast<long long>(kFrontMaxCameraTiles), static_cast<long long>(camera_tiles));
      return VGR_E_INVALID_ARGUMENT;
    }
  }
  return 0;
}

}  // namespace

void set_error(const char* fmt, ...) {
  va_list args;
  va_start(args, fmt);
  vsnprintf(g_error, sizeof(g_error), fmt, args);
  va_end(args);
}

}  // namespace vgr

using namespace vgr;

extern "C" int32_t vgr_abi_version(void) { return VGR_ABI_VERSION; }

extern "C" const char* vgr_last_error(void) { return g_error; }

extern "C" int32_t vgr_emitted_sort_key_end_bit(int64_t global_tile_capacity) {
  if (global_tile_capacity <= 0 || global_tile_capacity > 0x7fffffffLL) return -1;
  return 32 + bit_length_u64(static_cast<uint64_t>(global_tile_capacity - 1));
}

extern "C" int32_t vgr_workspace_sizes_query(const vgr_cameras* cameras, const vgr_options* options,
                                             int64_t scene_gaussians, int64_t intersection_capacity,
                                             vgr_workspace_sizes* out) {
  if (cameras == nullptr || options == nullptr || out == nullptr || scene_gaussians < 0 ||
      intersection_capacity < 0) {
    set_error("vgr_workspace_sizes_query: null or negative argument");
    return VGR_E_INVALID_ARGUMENT;
  }
  const int status = validate_options(*options, *cameras);
  if (status != 0) return status;
  const vgr_options binned = binning_options(*options);
  const int64_t tiles = global_tile_count(*cameras, binned);
  const ScratchLayout layout = compute_layout(cameras->active_count, scene_gaussians, binned, tiles,
                                              intersection_capacity, max_camera_tiles(*cameras, binned),
                                              cameras->batch);
  out->scratch_bytes = static_cast<int64_t>(layout.total);
  out->tile_capacity = tiles;
  out->sort_key_bits = layout.sort_key_bits;
  out->sort_passes = layout.sort_passes;
  return 0;
}

extern "C" int64_t vgr_sort_temp_bytes(int64_t n_items, int64_t n_segments) {
  if (n_items < 0 || n_segments <= 0) return -1;
  vgr_options options{};
  options.tile_size = 16;
  options.near_plane = 0.01f;
  options.far_plane = 100.0f;
  const ScratchLayout layout = compute_layout(0, 0, options, n_segments, n_items, 0, 0);
  return static_cast<int64_t>(layout.total - layout.sort_histogram);
}

namespace vgr {
namespace {

// Validates one render's arguments and resolves its scratch layout and options.
int prepare_render(const vgr_scene& sc, const vgr_cameras& cams, const vgr_outputs& out, const vgr_workspace& ws,
                   const vgr_options& raster_opt, ScratchLayout& layout_out, int64_t& tiles_out,
                   vgr_options& resolved_out) {
  int status = validate_options(raster_opt, cams);
  if (status != 0) return status;
  if (raster_opt.ray_evaluation != 0 &&
      (ws.ray_records == nullptr || out.final_transmittance != nullptr || ws.slot_records != nullptr)) {
    set_error("exact-ray evaluation needs workspace.ray_records and takes no training checkpoints");
    return VGR_E_INVALID_ARGUMENT;
  }
  // Tile geometry from here on is the binning geometry.
  const vgr_options opt = binning_options(raster_opt);
  if (sc.gaussian_count < 0 || sc.gaussian_count > 0x7fffffffLL) {
    set_error("gaussian_count must be in [0, INT32_MAX]; got %lld", static_cast<long long>(sc.gaussian_count));
    return VGR_E_INVALID_ARGUMENT;
  }
  if (sc.gaussian_count > 0 &&
      (sc.means == nullptr || sc.covariances == nullptr || sc.opacities == nullptr || sc.features == nullptr ||
       sc.semantic_ids == nullptr)) {
    set_error("scene tensors must be non-null");
    return VGR_E_INVALID_ARGUMENT;
  }
  if (sc.scene_count < 0 || (sc.scene_count > 0 && (sc.registered_scene_ids == nullptr || sc.scene_offsets == nullptr))) {
    set_error("registered scene ids and offsets must be non-null");
    return VGR_E_INVALID_ARGUMENT;
  }
  if (sc.feature_width < 3) {
    set_error("features must provide at least RGB channels; got width %d", sc.feature_width);
    return VGR_E_INVALID_ARGUMENT;
  }
  if (sc.sh_degree < 0 || sc.sh_degree > 3 || 3 * (sc.sh_degree + 1) * (sc.sh_degree + 1) > sc.feature_width) {
    set_error("sh_degree %d needs %d feature channels; features have %d", sc.sh_degree,
              3 * (sc.sh_degree + 1) * (sc.sh_degree + 1), sc.feature_width);
    return VGR_E_INVALID_ARGUMENT;
  }
  if (cams.batch < 0 || cams.active_count < 0 || (cams.active_ids == nullptr && cams.active_count != cams.batch)) {
    set_error("active_count must equal batch when active_ids is NULL");
    return VGR_E_INVALID_ARGUMENT;
  }
  if (cams.active_count > 0 && cams.batch <= 0) {
    set_error("active cameras given for an empty batch");
    return VGR_E_INVALID_ARGUMENT;
  }
  if (cams.batch > 0 &&
      (cams.viewmats == nullptr || cams.intrinsics == nullptr || cams.env_xforms == nullptr ||
       cams.scene_ids == nullptr)) {
    set_error("camera tensors must be non-null");
    return VGR_E_INVALID_ARGUMENT;
  }
  if (out.rgb == nullptr ||
      (opt.full_sensor_output && (out.depth == nullptr || out.alpha == nullptr || out.semantic == nullptr))) {
    set_error("output tensors must be non-null for the requested output set");
    return VGR_E_INVALID_ARGUMENT;
  }
  if (ws.records == nullptr || ws.keys_in == nullptr || ws.keys_out == nullptr || ws.values_in == nullptr ||
      ws.values_out == nullptr || ws.tile_starts == nullptr || ws.tile_ends == nullptr ||
      ws.counters == nullptr || ws.scratch == nullptr) {
    set_error("workspace buffers must be non-null");
    return VGR_E_INVALID_ARGUMENT;
  }
  if (ws.visible_capacity <= 0 || ws.visible_capacity > 0x7fffffffLL) {
    set_error("visible capacity must be in [1, INT32_MAX]; got %lld", static_cast<long long>(ws.visible_capacity));
    return VGR_E_INVALID_ARGUMENT;
  }
  // Both orderings address intersections with int32 positions (tile starts and
  // ends, the radix sort's 31-bit look-back prefixes): at most 2^31 - 1.
  const int64_t intersection_limit = 0x7fffffffLL;
  if (ws.intersection_capacity <= 0 || ws.intersection_capacity > intersection_limit) {
    set_error("intersection capacity must be in [1, %lld]; got %lld", static_cast<long long>(intersection_limit),
              static_cast<long long>(ws.intersection_capacity));
    return VGR_E_INVALID_ARGUMENT;
  }
  const int64_t tiles = global_tile_count(cams, opt);
  if (tiles > 0x7fffffffLL) {
    set_error("global tile count %lld exceeds the int32/uint32 key range", static_cast<long long>(tiles));
    return VGR_E_INVALID_ARGUMENT;
  }
  if (opt.tile_size < 8 && (out.final_transmittance != nullptr || ws.slot_records != nullptr)) {
    set_error("training checkpoints need 8x8 or 16x16 tiles; got tile_size %d", opt.tile_size);
    return VGR_E_INVALID_ARGUMENT;
  }
  if (opt.tile_size == 1 && opt.depth_bucket_count > 0 && ws.radii == nullptr) {
    set_error("the depth cutoff (tile_size 1, depth_bucket_count > 0) needs workspace.radii");
    return VGR_E_INVALID_ARGUMENT;
  }
  if (ws.slot_records != nullptr && opt.ordering == VGR_ORDERING_TILE_FRONT) {
    set_error("tile-front lists hold Gaussians, not intersection slots: workspace.slot_records must be NULL "
              "(its checkpointed forward feeds the pose-only backward)");
    return VGR_E_INVALID_ARGUMENT;
  }
  if (ws.slot_records != nullptr && sc.gaussian_ids != nullptr) {
    set_error("slot-sorted (training) forwards take the scene in registration order (gaussian_ids must be NULL)");
    return VGR_E_INVALID_ARGUMENT;
  }
  if (ws.tile_capacity < tiles) {
    set_error("tile range workspace holds %lld tiles; the batch needs %lld", static_cast<long long>(ws.tile_capacity),
              static_cast<long long>(tiles));
    return VGR_E_INVALID_ARGUMENT;
  }
  const ScratchLayout layout = compute_layout(cams.active_count, sc.gaussian_count, opt, tiles,
                                              ws.intersection_capacity, max_camera_tiles(cams, opt), cams.batch);
  if (ws.scratch_bytes < static_cast<int64_t>(layout.total)) {
    set_error("scratch holds %lld bytes; the render needs %lld", static_cast<long long>(ws.scratch_bytes),
              static_cast<long long>(layout.total));
    return VGR_E_INVALID_ARGUMENT;
  }
  if (layout.projection_blocks > 0x7fffffffLL || layout.sort_tiles > 0x7fffffffLL) {
    set_error("projection or sort grid exceeds the launch limit");
    return VGR_E_INVALID_ARGUMENT;
  }
  vgr_options resolved = opt;
  resolved.sort_digit_bits = layout.sort_digit_bits;
  resolved.cameras_per_block = layout.cameras_per_block;
  if (resolved.raster_batch == 0) resolved.raster_batch = 256;

  layout_out = layout;
  tiles_out = tiles;
  resolved_out = resolved;
  return 0;
}

}  // namespace
}  // namespace vgr

namespace vgr {
namespace {

// The launch sequence of one forward submission (validated arguments).
int render_submission(const vgr_scene& sc, const vgr_cameras& cams, const vgr_outputs& out, const vgr_workspace& ws,
                      const vgr_options& opt, const ScratchLayout& layout, int64_t tiles,
                      const vgr_options& resolved, hipStream_t stream, int64_t* selector_out) {
  int status = 0;

  if (opt.stage_event_count != 0 && (opt.stage_event_count != 5 || opt.stage_events == nullptr)) {
    set_error("stage_events must hold exactly 5 events");
    return VGR_E_INVALID_ARGUMENT;
  }
  auto mark = [&](int stage) -> int {
    if (opt.stage_event_count == 0) return 0;
    const hipError_t error = hipEventRecord(static_cast<hipEvent_t>(opt.stage_events[stage]), stream);
    if (error != hipSuccess) {
      set_error("hipEventRecord(stage %d) failed: %s", stage, hipGetErrorString(error));
      return VGR_E_LAUNCH;
    }
    return 0;
  };

  char* scratch_base = static_cast<char*>(ws.scratch);
  CameraParams* table = reinterpret_cast<CameraParams*>(scratch_base + layout.camera_table);
  if ((status = mark(0)) != 0) return status;
  if (opt.statistics != nullptr && (status = launch_fill_i64(opt.statistics, 4, 0, stream)) != 0) return status;
  if ((status = launch_camera_table(sc, cams, resolved, table, stream)) != 0) return status;
  int sorted_in_out = 1;
  if (!opt.reuse_projection && layout.front) {
    if ((status = launch_front_projection(sc, cams, resolved, ws, layout, scratch_base, tiles, stream)) != 0)
      return status;
    if ((status = mark(1)) != 0) return status;
    if ((status = launch_front_sort(sc, ws, layout, scratch_base, tiles, stream)) != 0) return status;
    if ((status = mark(2)) != 0 || (status = mark(3)) != 0) return status;
  } else if (!opt.reuse_projection) {
    if ((status = launch_projection(sc, cams, resolved, ws, layout, scratch_base, stream)) != 0) return status;
    if ((status = mark(1)) != 0) return status;
    if ((status = launch_sort(ws, layout, scratch_base, static_cast<uint32_t>(opt.lookback_spin_limit), stream,
                              &sorted_in_out, sc.gaussian_ids != nullptr)) != 0)
      return status;
    if ((status = mark(2)) != 0) return status;
    if ((status = launch_ranges(ws, tiles, sorted_in_out != 0, stream)) != 0) return status;
    if ((status = mark(3)) != 0) return status;
  } else {
    if ((status = mark(1)) != 0 || (status = mark(2)) != 0 || (status = mark(3)) != 0) return status;
  }
  if ((status = launch_composite(sc, cams, out, resolved, ws, layout, scratch_base, sorted_in_out != 0, stream)) != 0)
    return status;
  if ((status = mark(4)) != 0) return status;
  if (opt.statistics != nullptr && !opt.reuse_projection) {
    if ((status = launch_copy_i64(ws.counters + VGR_COUNTER_SORT_FAILURES, opt.statistics + 1, stream)) != 0)
      return status;
  }
  if (selector_out != nullptr) *selector_out = sorted_in_out;
  return 0;
}

}  // namespace
}  // namespace vgr

extern "C" int32_t vgr_render(const vgr_scene* scene, const vgr_cameras* cameras, const vgr_outputs* outputs,
                              const vgr_workspace* workspace, const vgr_options* options, void* stream_handle,
                              int64_t* selector_out) {
  if (scene == nullptr || cameras == nullptr || outputs == nullptr || workspace == nullptr || options == nullptr) {
    set_error("vgr_render: null argument struct");
    return VGR_E_INVALID_ARGUMENT;
  }
  ScratchLayout layout;
  int64_t tiles = 0;
  vgr_options resolved;
  const int status = prepare_render(*scene, *cameras, *outputs, *workspace, *options, layout, tiles, resolved);
  if (status != 0) return status;
  return render_submission(*scene, *cameras, *outputs, *workspace, *options, layout, tiles, resolved,
                           static_cast<hipStream_t>(stream_handle), selector_out);
}

// ---- allocation-free variant: scratch in a library-owned device array ----
//
// vgr_render_static_scratch renders like vgr_render but takes its scratch
// (camera table, projection counts and bases, sort histograms and look-back
// flags, queue counter, tile order) from `scratch` below instead of
// workspace->scratch, for callers that cannot allocate device memory.  Each
// device holds its own copy of the array for the life of the process; host
// code reaches it only through the address hipGetSymbolAddress returns for the
// current device (the array's name is a host-side stand-in).  Every scratch
// word a render reads is written or cleared by a kernel of the same render, so
// leftovers of an earlier call are harmless; calls that overlap on one device
// share the array and race.
static __device__ __align__(256) unsigned char scratch[VGR_STATIC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_static_scratch_mutex;
void* g_static_scratch[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per device.
int static_scratch_address(void** address) {
  int device = 0;
  hipError_t error = hipGetDevice(&device);
  if (error != hipSuccess) {
    set_error("vgr_render_static_scratch: hipGetDevice failed: %s", hipGetErrorString(error));
    return VGR_E_LAUNCH;
  }
  if (device < 0 || device >= kMaxDevices) {
    set_error("vgr_render_static_scratch: device %d is beyond the %d devices supported", device, kMaxDevices);
    return VGR_E_INVALID_ARGUMENT;
  }
  std::lock_guard<std::mutex> lock(g_static_scratch_mutex);
  if (g_static_scratch[device] == nullptr) {
    void* pointer = nullptr;
    error = hipGetSymbolAddress(&pointer, HIP_SYMBOL(scratch));
    if (error != hipSuccess || pointer == nullptr) {
      set_error("vgr_render_static_scratch: hipGetSymbolAddress failed on device %d: %s", device,
                hipGetErrorString(error));
      return VGR_E_LAUNCH;
    }
    g_static_scratch[device] = pointer;
  }
  *address = g_static_scratch[device];
  return 0;
}

}  // namespace

extern "C" int64_t vgr_static_scratch_bytes(void) { return VGR_STATIC_SCRATCH_BYTES; }

extern "C" int32_t vgr_render_static_scratch(const vgr_scene* scene, const vgr_cameras* cameras,
                                             const vgr_outputs* outputs, const vgr_workspace* workspace,
                                             const vgr_options* options, void* stream_handle,
                                             int64_t* selector_out) {
  if (scene == nullptr || cameras == nullptr || outputs == nullptr || workspace == nullptr || options == nullptr) {
    set_error("vgr_render_static_scratch: null argument struct");
    return VGR_E_INVALID_ARGUMENT;
  }
  if (outputs->final_transmittance != nullptr || outputs->last_blended != nullptr ||
      workspace->slot_records != nullptr) {
    // vgr_render_backward re-reads the forward's scratch from workspace->scratch.
    set_error("vgr_render_static_scratch: checkpointed (training) forwards go through vgr_render, whose scratch "
              "the backward pass re-reads");
    return VGR_E_INVALID_ARGUMENT;
  }
  void* address = nullptr;
  int status = static_scratch_address(&address);
  if (status != 0) return status;
  vgr_workspace ws = *workspace;
  ws.scratch = address;
  ws.scratch_bytes = VGR_STATIC_SCRATCH_BYTES;
  vgr_workspace_sizes sizes{};
  status = vgr_workspace_sizes_query(cameras, options, scene->gaussian_count, ws.intersection_capacity, &sizes);
  if (status != 0) return status;
  if (sizes.scratch_bytes > VGR_STATIC_SCRATCH_BYTES) {
    set_error("vgr_render_static_scratch: this render needs %lld scratch bytes; the static scratch holds %lld "
              "(use vgr_render with a caller-owned scratch buffer)",
              static_cast<long long>(sizes.scratch_bytes), static_cast<long long>(VGR_STATIC_SCRATCH_BYTES));
    return VGR_E_CAPACITY;
  }
  ScratchLayout layout;
  int64_t tiles = 0;
  vgr_options resolved;
  status = prepare_render(*scene, *cameras, *outputs, ws, *options, layout, tiles, resolved);
  if (status != 0) return status;
  return render_submission(*scene, *cameras, *outputs, ws, *options, layout, tiles, resolved,
                           static_cast<hipStream_t>(stream_handle), selector_out);
}

extern "C" int32_t vgr_render_backward(const vgr_scene* scene, const vgr_cameras* cameras,
                                       const vgr_outputs* outputs, const vgr_workspace* workspace,
                                       const vgr_options* options, const vgr_gradients* gradients,
                                       void* stream_handle) {
  if (scene == nullptr || cameras == nullptr || outputs == nullptr || workspace == nullptr || options == nullptr ||
      gradients == nullptr) {
    set_error("vgr_render_backward: null argument struct");
    return VGR_E_INVALID_ARGUMENT;
  }
  const vgr_scene& sc = *scene;
  const vgr_cameras& cams = *cameras;
  const vgr_outputs& out = *outputs;
  const vgr_workspace& ws = *workspace;
  const vgr_gradients& grads = *gradients;
  ScratchLayout layout;
  int64_t tiles = 0;
  vgr_options resolved;
  int status = prepare_render(sc, cams, out, ws, *options, layout, tiles, resolved);
  if (status != 0) return status;
  if (out.final_transmittance == nullptr || out.last_blended == nullptr || !resolved.full_sensor_output ||
      resolved.output_srgb) {
    set_error("backward needs the forward's transmittance checkpoints, full sensor output and linear RGB");
    return VGR_E_INVALID_ARGUMENT;
  }
  const bool front = resolved.ordering == VGR_ORDERING_TILE_FRONT;
  const bool scene_grads = grads.grad_means != nullptr || grads.grad_covariances != nullptr ||
                           grads.grad_opacities != nullptr || grads.grad_features != nullptr;
  if (front) {
    // Tile-front lists carry no record slots: the pass computes pose gradients only.
    if (grads.grad_rgb == nullptr || grads.grad_viewmats == nullptr || scene_grads) {
      set_error("tile-front backward computes camera-pose gradients only: it needs grad_rgb and grad_viewmats, "
                "and scene gradients must be NULL");
      return VGR_E_INVALID_ARGUMENT;
    }
  } else {
    if (grads.grad_rgb == nullptr || grads.record_grads == nullptr) {
      set_error("backward needs grad_rgb and the record_grads workspace");
      return VGR_E_INVALID_ARGUMENT;
    }
    if (ws.slot_records == nullptr) {
      set_error("backward needs the forward's workspace.slot_records (slot-sorted training forward)");
      return VGR_E_INVALID_ARGUMENT;
    }
  }
  if (scene_grads && (grads.grad_means == nullptr || grads.grad_covariances == nullptr ||
                      grads.grad_opacities == nullptr || grads.grad_features == nullptr)) {
    set_error("scene gradients are all given or all NULL");
    return VGR_E_INVALID_ARGUMENT;
  }
  if (grads.grad_viewmats != nullptr && grads.camera_grads == nullptr) {
    set_error("pose gradients need the camera_grads workspace");
    return VGR_E_INVALID_ARGUMENT;
  }
  if (resolved.reuse_projection) {
    set_error("backward replays the forward's own projection; reuse_projection must be 0");
    return VGR_E_INVALID_ARGUMENT;
  }
  return launch_backward(sc, cams, out, resolved, ws, layout, static_cast<const char*>(ws.scratch), grads,
                         static_cast<hipStream_t>(stream_handle));
}

extern "C" int32_t vgr_covariance_backward(const float* scales, const float* rotations, const float* grad_covariances,
                                           int64_t count, float* grad_scales, float* grad_rotations,
                                           void* stream_handle) {
  if (count < 0 || (count > 0 && (scales == nullptr || rotations == nullptr || grad_covariances == nullptr ||
                                  grad_scales == nullptr || grad_rotations == nullptr))) {
    set_error("vgr_covariance_backward: null tensor or negative count");
    return VGR_E_INVALID_ARGUMENT;
  }
  return launch_covariance_backward(scales, rotations, grad_covariances, count, grad_scales, grad_rotations,
                                    static_cast<hipStream_t>(stream_handle));
}

extern "C" int32_t vgr_pose_compose(const float* twists, const float* base_views, int64_t count, float* views,
                                    void* stream_handle) {
  if (count < 0 || (count > 0 && (twists == nullptr || base_views == nullptr || views == nullptr))) {
    set_error("vgr_pose_compose: null tensor or negative count");
    return VGR_E_INVALID_ARGUMENT;
  }
  return launch_pose_compose(twists, base_views, count, views, static_cast<hipStream_t>(stream_handle));
}

extern "C" int32_t vgr_pose_compose_backward(const float* twists, const float* base_views, const float* grad_views,
                                             int64_t count, float* grad_twists, float* grad_base_views,
                                             void* stream_handle) {
  if (count < 0 || (count > 0 && (twists == nullptr || base_views == nullptr || grad_views == nullptr ||
                                  grad_twists == nullptr))) {
    set_error("vgr_pose_compose_backward: null tensor or negative count");
    return VGR_E_INVALID_ARGUMENT;
  }
  return launch_pose_compose_backward(twists, base_views, grad_views, count, grad_twists, grad_base_views,
                                      static_cast<hipStream_t>(stream_handle));
}

namespace vgr {
namespace {

__global__ void mark_logical_active_ids_kernel(const int64_t* __restrict__ active_ids, int64_t active_count,
                                               int32_t* __restrict__ active_mask, int64_t logical_batch) {
  const int64_t index = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (index >= active_count) return;
  const int64_t camera = active_ids[index];
  if (camera >= 0 && camera < logical_batch) active_mask[camera] = 1;
}

__global__ void build_chunk_active_ids_kernel(const int32_t* __restrict__ active_mask, int64_t logical_batch,
                                              int64_t physical_batch, int64_t slot_count,
                                              int64_t* __restrict__ chunk_active_ids) {
  const int64_t slot = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (slot >= slot_count) return;
  chunk_active_ids[slot] = (slot < logical_batch && active_mask[slot] != 0) ? slot % physical_batch : -1;
}

__global__ void aggregate_chunk_counters_kernel(const int64_t* __restrict__ chunk_counters, int64_t stride,
                                                int64_t chunk_count, int64_t* __restrict__ counters,
                                                int64_t* __restrict__ maxima) {
  const int counter = threadIdx.x;
  if (counter >= VGR_COUNTER_COUNT) return;
  int64_t total = 0, maximum = 0;
  for (int64_t chunk = 0; chunk < chunk_count; ++chunk) {
    const int64_t value = chunk_counters[chunk * stride + counter];
    total += value;
    maximum = value > maximum ? value : maximum;
  }
  counters[counter] = total;
  maxima[counter] = maximum;
}

__global__ void clear_mask_kernel(int32_t* __restrict__ mask, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) mask[i] = 0;
}

}  // namespace
}  // namespace vgr

extern "C" int32_t vgr_prepare_chunked_active_ids(const int64_t* active_ids, int64_t active_count,
                                                  int32_t* active_mask, int64_t* chunk_active_ids,
                                                  int64_t logical_batch, int64_t physical_batch, void* stream_handle) {
  if (logical_batch <= 0 || physical_batch <= 0 || active_mask == nullptr || chunk_active_ids == nullptr ||
      (active_count > 0 && active_ids == nullptr)) {
    set_error("vgr_prepare_chunked_active_ids: invalid arguments");
    return VGR_E_INVALID_ARGUMENT;
  }
  hipStream_t stream = static_cast<hipStream_t>(stream_handle);
  const int64_t chunks = (logical_batch + physical_batch - 1) / physical_batch;
  const int64_t slots = chunks * physical_batch;
  hipLaunchKernelGGL(clear_mask_kernel, dim3(static_cast<unsigned>((logical_batch + 255) / 256)), dim3(256), 0,
                     stream, active_mask, logical_batch);
  if (active_count > 0) {
    hipLaunchKernelGGL(mark_logical_active_ids_kernel, dim3(static_cast<unsigned>((active_count + 255) / 256)),
                       dim3(256), 0, stream, active_ids, active_count, active_mask, logical_batch);
  }
  hipLaunchKernelGGL(build_chunk_active_ids_kernel, dim3(static_cast<unsigned>((slots + 255) / 256)), dim3(256), 0,
                     stream, active_mask, logical_batch, physical_batch, slots, chunk_active_ids);
  const hipError_t error = hipGetLastError();
  if (error != hipSuccess) {
    set_error("chunk active-id launch failed: %s", hipGetErrorString(error));
    return VGR_E_LAUNCH;
  }
  return 0;
}

extern "C" int32_t vgr_aggregate_chunk_counters(const int64_t* chunk_counters, int64_t counter_stride,
                                                int64_t chunk_count, int64_t* counters,
                                                int64_t* physical_capacity_counters, void* stream_handle) {
  if (chunk_counters == nullptr || counters == nullptr || physical_capacity_counters == nullptr ||
      chunk_count <= 0 || counter_stride < VGR_COUNTER_COUNT) {
    set_error("vgr_aggregate_chunk_counters: invalid arguments");
    return VGR_E_INVALID_ARGUMENT;
  }
  hipLaunchKernelGGL(aggregate_chunk_counters_kernel, dim3(1), dim3(64), 0, static_cast<hipStream_t>(stream_handle),
                     chunk_counters, counter_stride, chunk_count, counters, physical_capacity_counters);
  const hipError_t error = hipGetLastError();
  if (error != hipSuccess) {
    set_error("counter aggregation launch failed: %s", hipGetErrorString(error));
    return VGR_E_LAUNCH;
  }
  return 0;
}

extern "C" int32_t vgr_fill_i64(int64_t* data, int64_t n, int64_t value, void* stream) {
  if (n < 0 || (n > 0 && data == nullptr)) {
    set_error("vgr_fill_i64: invalid arguments");
    return VGR_E_INVALID_ARGUMENT;
  }
  return launch_fill_i64(data, n, value, static_cast<hipStream_t>(stream));
}

extern "C" int32_t vgr_fill_f32(float* data, int64_t n, float value, void* stream) {
  if (n < 0 || (n > 0 && data == nullptr)) {
    set_error("vgr_fill_f32: invalid arguments");
    return VGR_E_INVALID_ARGUMENT;
  }
  return vgr::launch_fill_f32(data, n, value, static_cast<hipStream_t>(stream));
}
