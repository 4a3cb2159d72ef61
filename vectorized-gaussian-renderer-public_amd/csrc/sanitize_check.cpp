// Sanitizer driver for the host side of libvgr_hip: built with the library by
// `make sanitize` (host code under -fsanitize=address,undefined through
// -Xarch_host, device code unchanged) and run on a machine without a GPU by
// tests/test_sanitizers.py.  It exercises every host path that does not launch
// a kernel: the end-bit table, scratch planning for both orderings, culling
// and per-camera resolutions, and the argument validation of every entry point
// (each rejected call must return its error code and a message before any
// device work).
#include <cstdio>
#include <cstring>
#include <initializer_list>

#include "../../include/vgr.h"

namespace {

int failures = 0;

void expect(bool condition, const char* what) {
  if (!condition) {
    std::fprintf(stderr, "sanitize_check: FAILED %s (last error: %s)\n", what, vgr_last_error());
    ++failures;
  }
}

vgr_options default_options() {
  vgr_options options;
  std::memset(&options, 0, sizeof(options));
  options.height = 480;
  options.width = 640;
  options.near_plane = 0.01f;
  options.far_plane = 100.0f;
  options.gaussian_support_sigma = 3.33f;
  options.covariance_epsilon = 0.3f;
  options.semantic_min_alpha = 0.01f;
  options.tile_size = 16;
  options.full_sensor_output = 1;
  return options;
}

vgr_cameras uniform_cameras(int batch) {
  vgr_cameras cameras;
  std::memset(&cameras, 0, sizeof(cameras));
  cameras.batch = batch;
  cameras.active_count = batch;
  return cameras;
}

void planning() {
  expect(vgr_abi_version() == VGR_ABI_VERSION, "abi version");
  const long long tiles[] = {1, 2, 3, 4, 1 << 20, (1 << 20) + 1, 2147483647LL};
  const int bits[] = {32, 33, 34, 34, 52, 53, 63};
  for (int i = 0; i < 7; ++i) expect(vgr_emitted_sort_key_end_bit(tiles[i]) == bits[i], "end-bit table");
  expect(vgr_emitted_sort_key_end_bit(0) == -1 && vgr_emitted_sort_key_end_bit(2147483648LL) == -1,
         "end-bit rejects");
  expect(vgr_sort_temp_bytes(1000, 64) > 0 && vgr_sort_temp_bytes(-1, 1) < 0, "sort temp bytes");
  int64_t previous = 0;
  for (int batch : {1, 2, 5, 16, 64, 200}) {
    vgr_cameras cameras = uniform_cameras(batch);
    vgr_options options = default_options();
    options.chunk_culling = 1;
    vgr_workspace_sizes sizes;
    expect(vgr_workspace_sizes_query(&cameras, &options, 100000, 5000000, &sizes) == 0, "sizes query");
    expect(sizes.scratch_bytes >= previous && sizes.sort_passes > 0, "scratch grows with the batch");
    previous = sizes.scratch_bytes;
  }
  // Tile-front ordering with per-camera resolutions.
  vgr_cameras cameras = uniform_cameras(1024);
  cameras.widths = reinterpret_cast<const int32_t*>(1);  // only the tile totals are read when planning
  cameras.heights = reinterpret_cast<const int32_t*>(1);
  cameras.max_global_tiles = 1024 * 80;
  cameras.max_camera_tiles = 80;
  vgr_options options = default_options();
  options.ordering = VGR_ORDERING_TILE_FRONT;
  options.chunk_culling = 1;
  vgr_workspace_sizes sizes;
  expect(vgr_workspace_sizes_query(&cameras, &options, 5000000, 1500000000LL, &sizes) == 0, "front sizes");
  options.ordering = VGR_ORDERING_GLOBAL_RADIX;
  vgr_workspace_sizes global;
  expect(vgr_workspace_sizes_query(&cameras, &options, 5000000, 300000000LL, &global) == 0, "global sizes");
  cameras.max_camera_tiles = 4096;
  options.ordering = VGR_ORDERING_TILE_FRONT;
  expect(vgr_workspace_sizes_query(&cameras, &options, 5000000, 1000, &sizes) == VGR_E_INVALID_ARGUMENT,
         "front rejects large cameras");
  // Dense path: per-pixel tiles with and without the depth cutoff; the cutoff only at tile size 1.
  vgr_cameras dense = uniform_cameras(3);
  vgr_options dense_options = default_options();
  dense_options.tile_size = 1;
  vgr_workspace_sizes plain_dense, cut_dense;
  expect(vgr_workspace_sizes_query(&dense, &dense_options, 100000, 4000000, &plain_dense) == 0, "dense sizes");
  expect(plain_dense.tile_capacity == 3LL * 640 * 480, "dense tiles are pixels");
  dense_options.depth_bucket_count = 4096;
  dense_options.depth_bucket_group_size = 64;
  expect(vgr_workspace_sizes_query(&dense, &dense_options, 100000, 4000000, &cut_dense) == 0, "cutoff sizes");
  expect(cut_dense.scratch_bytes >= plain_dense.scratch_bytes + 3LL * 640 * 480 * (64 * 4 + 8 + 4),
         "cutoff tables in scratch");
  dense_options.tile_size = 16;
  expect(vgr_workspace_sizes_query(&dense, &dense_options, 100000, 4000000, &sizes) == VGR_E_INVALID_ARGUMENT,
         "cutoff needs tile size 1");
  dense_options.tile_size = 3;
  dense_options.depth_bucket_count = 0;
  expect(vgr_workspace_sizes_query(&dense, &dense_options, 100000, 4000000, &sizes) == VGR_E_INVALID_ARGUMENT,
         "tile size must be a power of two");
}

void validation() {
  vgr_scene scene;
  std::memset(&scene, 0, sizeof(scene));
  scene.feature_width = 3;
  vgr_cameras cameras = uniform_cameras(1);
  vgr_outputs outputs;
  std::memset(&outputs, 0, sizeof(outputs));
  vgr_workspace workspace;
  std::memset(&workspace, 0, sizeof(workspace));
  vgr_options options = default_options();
  int64_t selector = 0;
  expect(vgr_render(nullptr, &cameras, &outputs, &workspace, &options, nullptr, &selector) ==
             VGR_E_INVALID_ARGUMENT, "null scene");
  expect(std::strlen(vgr_last_error()) > 0, "message");
  options.tile_size = 4;
  expect(vgr_render(&scene, &cameras, &outputs, &workspace, &options, nullptr, &selector) ==
             VGR_E_INVALID_ARGUMENT, "tile size");
  options = default_options();
  options.sort_digit_bits = 12;
  expect(vgr_render(&scene, &cameras, &outputs, &workspace, &options, nullptr, &selector) ==
             VGR_E_INVALID_ARGUMENT, "digit bits");
  options = default_options();
  options.ordering = 7;
  expect(vgr_render(&scene, &cameras, &outputs, &workspace, &options, nullptr, &selector) ==
             VGR_E_INVALID_ARGUMENT, "ordering");
  options = default_options();
  // Non-null tensors so validation reaches the workspace checks.
  float dummy_f[64] = {};
  int64_t dummy_i[64] = {};
  uint64_t dummy_k[64] = {};
  int32_t dummy_v[64] = {};
  scene.means = scene.covariances = scene.opacities = scene.features = dummy_f;
  scene.semantic_ids = dummy_i;
  scene.registered_scene_ids = dummy_i;
  scene.scene_offsets = dummy_i;
  scene.gaussian_count = 1;
  scene.scene_count = 1;
  cameras.viewmats = cameras.intrinsics = cameras.env_xforms = dummy_f;
  cameras.scene_ids = dummy_i;
  outputs.rgb = outputs.depth = outputs.alpha = dummy_f;
  outputs.semantic = dummy_i;
  workspace.records = reinterpret_cast<vgr_record*>(dummy_f);
  workspace.keys_in = workspace.keys_out = dummy_k;
  workspace.values_in = workspace.values_out = workspace.tile_starts = workspace.tile_ends = dummy_v;
  workspace.counters = dummy_i;
  workspace.scratch = dummy_f;
  workspace.visible_capacity = 1;
  workspace.intersection_capacity = 1LL << 31;  // beyond the int32 position limit
  workspace.tile_capacity = 1200;
  workspace.scratch_bytes = 1 << 30;
  expect(vgr_render(&scene, &cameras, &outputs, &workspace, &options, nullptr, &selector) ==
             VGR_E_INVALID_ARGUMENT, "intersection capacity");
  workspace.intersection_capacity = 64;
  workspace.tile_capacity = 10;
  expect(vgr_render(&scene, &cameras, &outputs, &workspace, &options, nullptr, &selector) ==
             VGR_E_INVALID_ARGUMENT, "tile capacity");
  workspace.tile_capacity = 1200;
  workspace.scratch_bytes = 16;
  expect(vgr_render(&scene, &cameras, &outputs, &workspace, &options, nullptr, &selector) ==
             VGR_E_INVALID_ARGUMENT, "scratch size");
  scene.sh_degree = 3;  // needs 48 feature channels
  expect(vgr_render(&scene, &cameras, &outputs, &workspace, &options, nullptr, &selector) ==
             VGR_E_INVALID_ARGUMENT, "sh degree");
  scene.sh_degree = 0;
  int64_t active[1] = {0};
  cameras.active_ids = active;
  options.ordering = VGR_ORDERING_TILE_FRONT;
  expect(vgr_render(&scene, &cameras, &outputs, &workspace, &options, nullptr, &selector) ==
             VGR_E_INVALID_ARGUMENT, "front needs all cameras");
  cameras.active_ids = nullptr;
  options = default_options();
  options.tile_size = 1;
  options.depth_bucket_count = 64;
  expect(vgr_render(&scene, &cameras, &outputs, &workspace, &options, nullptr, &selector) ==
             VGR_E_INVALID_ARGUMENT, "cutoff needs radii");
  options = default_options();
  vgr_gradients gradients;
  std::memset(&gradients, 0, sizeof(gradients));
  expect(vgr_render_backward(&scene, &cameras, &outputs, &workspace, &options, &gradients, nullptr) ==
             VGR_E_INVALID_ARGUMENT, "backward without checkpoints");
  // Tile-front lists: checkpointed forwards take no slot records, and the backward is pose-only.
  options.ordering = VGR_ORDERING_TILE_FRONT;
  outputs.final_transmittance = dummy_f;
  outputs.last_blended = dummy_v;
  workspace.slot_records = dummy_v;
  expect(vgr_render(&scene, &cameras, &outputs, &workspace, &options, nullptr, &selector) ==
             VGR_E_INVALID_ARGUMENT, "front forward rejects slot records");
  workspace.slot_records = nullptr;
  gradients.grad_rgb = dummy_f;
  gradients.grad_viewmats = dummy_f;
  gradients.camera_grads = dummy_f;
  gradients.grad_means = dummy_f;
  expect(vgr_render_backward(&scene, &cameras, &outputs, &workspace, &options, &gradients, nullptr) ==
             VGR_E_INVALID_ARGUMENT, "front backward is pose-only");
  std::memset(&gradients, 0, sizeof(gradients));
  outputs.final_transmittance = nullptr;
  outputs.last_blended = nullptr;
  options = default_options();
  expect(vgr_covariance_backward(nullptr, nullptr, nullptr, 4, nullptr, nullptr, nullptr) ==
             VGR_E_INVALID_ARGUMENT, "covariance backward nulls");
  expect(vgr_prepare_chunked_active_ids(nullptr, 0, nullptr, nullptr, 0, 0, nullptr) == VGR_E_INVALID_ARGUMENT,
         "chunked ids");
  expect(vgr_aggregate_chunk_counters(nullptr, 6, 1, nullptr, nullptr, nullptr) == VGR_E_INVALID_ARGUMENT,
         "aggregate counters");
  expect(vgr_fill_i64(nullptr, 4, 0, nullptr) == VGR_E_INVALID_ARGUMENT, "fill");
  expect(vgr_static_scratch_bytes() == VGR_STATIC_SCRATCH_BYTES, "static scratch size");
  outputs.final_transmittance = dummy_f;
  outputs.last_blended = dummy_v;
  expect(vgr_render_static_scratch(&scene, &cameras, &outputs, &workspace, &options, nullptr, &selector) ==
             VGR_E_INVALID_ARGUMENT, "static scratch rejects training checkpoints");
  outputs.final_transmittance = nullptr;
  outputs.last_blended = nullptr;
  expect(vgr_pose_compose(nullptr, nullptr, 2, nullptr, nullptr) == VGR_E_INVALID_ARGUMENT, "pose compose nulls");
  expect(vgr_pose_compose_backward(nullptr, nullptr, nullptr, -1, nullptr, nullptr, nullptr) ==
             VGR_E_INVALID_ARGUMENT, "pose compose backward count");
}

}  // namespace

int main() {
  planning();
  validation();
  if (failures != 0) return 1;
  std::printf("sanitize_check: library host code OK\n");
  return 0;
}
