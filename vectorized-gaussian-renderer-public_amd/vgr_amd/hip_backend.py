"""gfx950 renderer backend: the drop-in for the reference's ``CustomCudaBackend``.

Same constructor knobs, validation messages, workspace lifecycle, bounded
adaptive capacity (grow x1.25 with x1.25 headroom, release-then-allocate with
restoration on allocator failure, fail-closed with NaN/-1 poisoned outputs),
logical-batch chunking, projection cache and ``capacity_stats`` telemetry as
``src/isaacsim_gaussian_renderer/cuda_backend.py:63-2231``.  The native
module is ``libvgr_hip`` through :mod:`vgr_amd.native`.

Differences from the reference, by design:
  * one render submission never synchronizes the host: the radix sort reads
    the intersection count on the device, so ``fixed_capacity_sort`` changes
    nothing and the non-deterministic mode is deterministic as well;
  * intersection order is always (tile, depth, Gaussian ID);
  * SH bands 1..3 in ``features`` are evaluated (``sh_degree``; the reference
    composites ``features[:, :3]`` only, which ``sh_degree=0`` reproduces);
  * the dense 1x1 compact path and exact-ray evaluation are not part of this
    build and are rejected at construction.
"""

from __future__ import annotations

import copy
import math
from collections import OrderedDict
from collections.abc import Callable
from typing import Any, Literal, NoReturn

import torch

from .backend import RenderRequest, RenderStats
from .native import RECORD_FLOATS, WORKSPACE_TENSORS, float_bits, infer_sh_degree, load_native_module
from .scene import GaussianScene

FULL_SENSOR_OUTPUTS = ("rgb", "depth", "alpha", "semantic_id")
RGB_ONLY_OUTPUTS = ("rgb",)
DEFAULT_GAUSSIAN_SUPPORT_SIGMA = 3.33
COUNTER_NAMES = (
    "visible_gaussians",
    "tile_intersections",
    "visible_overflow",
    "intersection_overflow",
    "active_tiles",
)
_INT32_MAX = 2_147_483_647
_SORT_CAPACITY_LIMIT = (1 << 30) - 1  # look-back flags carry 30-bit prefixes
_RECORD_BYTES = 4 * RECORD_FLOATS
_DTYPE_BYTES = {torch.float32: 4, torch.int32: 4, torch.int64: 8, torch.uint8: 1, torch.uint64: 8}


def _emitted_sort_key_end_bit(global_tile_capacity: int) -> int:
    """32 + bit length of the largest tile index (reference cuda_backend.py:40-47)."""
    if global_tile_capacity <= 0:
        raise ValueError("Global tile capacity must be positive.")
    if global_tile_capacity > _INT32_MAX:
        raise ValueError("Global tile capacity exceeds the int32 implementation limit.")
    return 32 + (global_tile_capacity - 1).bit_length()


def default_sort_digit_bits(key_bits: int) -> int:
    """The library's default radix digit width: 8, or 9 where that saves a pass."""
    return 9 if -(-key_bits // 9) < -(-key_bits // 8) else 8


def compressed_sort_key_bits(global_tiles: int, near_plane: float, far_plane: float) -> int:
    """Significant bits of the order-preserving sort-key image used on the device."""
    tile_bits = (global_tiles - 1).bit_length() if global_tiles > 1 else 0
    depth_bits = (float_bits(far_plane) - float_bits(near_plane) + 1).bit_length()
    return tile_bits + depth_bits


class RendererCapacityError(RuntimeError):
    """Raised when bounded capacity handling cannot produce a valid render."""

    def __init__(self, message: str, *, telemetry: dict[str, Any] | None = None) -> None:
        super().__init__(message)
        self.telemetry = copy.deepcopy(telemetry)


class CustomHipBackend:
    """Bounded adaptive Gaussian renderer on gfx950 with device-only submissions."""

    def __init__(
        self,
        *,
        max_visible_records: int | None = None,
        max_intersections: int | None = None,
        visible_capacity_per_view: int = 100_000,
        intersections_per_visible: float = 4.0,
        near_plane: float = 0.01,
        far_plane: float = 100.0,
        gaussian_support_sigma: float = DEFAULT_GAUSSIAN_SUPPORT_SIGMA,
        covariance_epsilon: float = 0.3,
        rasterize_mode: Literal["classic", "antialiased"] = "classic",
        semantic_min_alpha: float = 0.01,
        ray_gaussian_evaluation: bool = False,
        tile_size: int = 16,
        depth_bucket_count: int = 4096,
        depth_bucket_group_size: int = 64,
        compact_projection_cache: bool = False,
        materialize_projected_records: bool = False,
        enable_projection_cache: bool = False,
        output_srgb: bool = True,
        deterministic: bool = False,
        fixed_capacity_sort: bool = False,
        max_physical_views: int | None = None,
        adaptive_capacity: bool = True,
        capacity_growth_factor: float = 1.25,
        capacity_headroom: float = 1.25,
        max_capacity_retries: int = 4,
        max_workspace_bytes: int | None = None,
        sh_degree: int | None = None,
        sort_digit_bits: int = 0,
        raster_batch: int = 0,
        cameras_per_block: int = 0,
        sort_tile_items: int = 0,
        native_module: Any | None = None,
        native_loader: Callable[[], Any] = load_native_module,
        allow_cpu_for_tests: bool = False,
    ) -> None:
        if max_visible_records is not None and max_visible_records <= 0:
            raise ValueError("max_visible_records must be positive.")
        if max_intersections is not None and max_intersections <= 0:
            raise ValueError("max_intersections must be positive.")
        if visible_capacity_per_view <= 0:
            raise ValueError("visible_capacity_per_view must be positive.")
        if intersections_per_visible <= 0:
            raise ValueError("intersections_per_visible must be positive.")
        if near_plane <= 0 or far_plane <= near_plane:
            raise ValueError("Expected 0 < near_plane < far_plane.")
        if gaussian_support_sigma <= 0:
            raise ValueError("gaussian_support_sigma must be positive.")
        if not math.isfinite(covariance_epsilon) or covariance_epsilon < 0:
            raise ValueError("covariance_epsilon must be finite and non-negative.")
        if rasterize_mode not in ("classic", "antialiased"):
            raise ValueError("rasterize_mode must be 'classic' or 'antialiased'.")
        if not math.isfinite(semantic_min_alpha) or semantic_min_alpha < 0 or semantic_min_alpha > 1:
            raise ValueError("semantic_min_alpha must be finite and in [0, 1].")
        if tile_size <= 0 or tile_size > 16 or tile_size & (tile_size - 1):
            raise ValueError("tile_size must be a power of two in [1, 16].")
        if tile_size not in (8, 16):
            raise ValueError("tile_size must be 8 or 16 on the gfx950 tile path (the dense 1x1 path is not built).")
        if depth_bucket_count <= 0:
            raise ValueError("depth_bucket_count must be positive.")
        if depth_bucket_group_size <= 0:
            raise ValueError("depth_bucket_group_size must be positive.")
        if compact_projection_cache or materialize_projected_records:
            raise ValueError("compact_projection_cache / materialize_projected_records (the dense 1x1 path) "
                             "are not available in the gfx950 build.")
        if ray_gaussian_evaluation:
            raise ValueError("ray_gaussian_evaluation is not available in the gfx950 build.")
        if fixed_capacity_sort and deterministic:
            raise ValueError("fixed_capacity_sort is only meaningful for the nondeterministic global-radix path.")
        if max_physical_views is not None and max_physical_views <= 0:
            raise ValueError("max_physical_views must be positive when set.")
        if max_physical_views is not None and not fixed_capacity_sort:
            raise ValueError("max_physical_views requires fixed_capacity_sort=True so logical chunks never "
                             "synchronize on an emitted count.")
        if max_physical_views is not None and adaptive_capacity:
            raise ValueError("max_physical_views requires adaptive_capacity=False; physical capacities must be "
                             "calibrated and installed before steady-state logical-batch rendering.")
        if max_physical_views is not None and enable_projection_cache:
            raise ValueError("max_physical_views currently requires enable_projection_cache=False.")
        if not math.isfinite(capacity_growth_factor) or capacity_growth_factor <= 1.0:
            raise ValueError("capacity_growth_factor must be finite and greater than 1.")
        if not math.isfinite(capacity_headroom) or capacity_headroom < 1.0:
            raise ValueError("capacity_headroom must be finite and at least 1.")
        if isinstance(max_capacity_retries, bool) or not isinstance(max_capacity_retries, int) \
                or max_capacity_retries < 0:
            raise ValueError("max_capacity_retries must be non-negative.")
        if max_workspace_bytes is not None and max_workspace_bytes <= 0:
            raise ValueError("max_workspace_bytes must be positive when set.")
        if sh_degree is not None and not 0 <= sh_degree <= 3:
            raise ValueError("sh_degree must be in [0, 3] or None (infer from the feature width).")
        if sort_digit_bits != 0 and not 4 <= sort_digit_bits <= 9:
            raise ValueError("sort_digit_bits must be 0 (default) or in [4, 9].")
        if raster_batch not in (0, 64, 128, 256, 512):
            raise ValueError("raster_batch must be 0 (default), 64, 128, 256 or 512.")
        if not 0 <= cameras_per_block <= 64:
            raise ValueError("cameras_per_block must be 0 (default) or in [1, 64].")
        if sort_tile_items not in (0, 8, 12, 16):
            raise ValueError("sort_tile_items must be 0 (default), 8, 12 or 16.")

        self.max_visible_records = max_visible_records
        self.max_intersections = max_intersections
        self.visible_capacity_per_view = visible_capacity_per_view
        self.intersections_per_visible = intersections_per_visible
        self.near_plane = near_plane
        self.far_plane = far_plane
        self.gaussian_support_sigma = gaussian_support_sigma
        self.covariance_epsilon = covariance_epsilon
        self._rasterize_mode: Literal["classic", "antialiased"] = rasterize_mode
        self.semantic_min_alpha = semantic_min_alpha
        self.ray_gaussian_evaluation = False
        self.tile_size = tile_size
        self.depth_bucket_count = depth_bucket_count
        self.depth_bucket_group_size = min(depth_bucket_group_size, depth_bucket_count)
        self.compact_projection_cache = False
        self.materialize_projected_records = False
        self.enable_projection_cache = enable_projection_cache
        self.output_srgb = output_srgb
        self.deterministic = deterministic
        self.fixed_capacity_sort = fixed_capacity_sort
        self.max_physical_views = max_physical_views
        self.adaptive_capacity = adaptive_capacity
        self.capacity_growth_factor = capacity_growth_factor
        self.capacity_headroom = capacity_headroom
        self.max_capacity_retries = max_capacity_retries
        self.max_workspace_bytes = max_workspace_bytes
        self.sh_degree = sh_degree
        self.sort_digit_bits = sort_digit_bits
        self.raster_batch = raster_batch
        self.cameras_per_block = cameras_per_block
        self.sort_tile_items = sort_tile_items
        self._native = native_module
        self._native_loader = native_loader
        # Optional instrumentation forwarded to the native render: five recorded
        # torch events (stage boundaries) and an int64[4] statistics tensor.
        self.stage_events: list[torch.cuda.Event] | None = None
        self.statistics: torch.Tensor | None = None
        self.allow_cpu_for_tests = allow_cpu_for_tests

        self.device: torch.device | None = None
        self.stage: Any | None = None
        self.workspace: dict[str, torch.Tensor] = {}
        self._scenes: OrderedDict[int, GaussianScene] = OrderedDict()
        self._packed_scene: dict[str, torch.Tensor] = {}
        self._max_scene_gaussians = 0
        self._resolved_sh_degree = 0
        self._env_scene_ids: torch.Tensor | None = None
        self._env_transforms: torch.Tensor | None = None
        self._max_views = 0
        self._physical_max_views = 0
        self._max_logical_chunks = 1
        self._height = 0
        self._width = 0
        self._scene_revision = 0
        self._projection_cache_token: tuple[Any, ...] | None = None
        self._projection_cache_tensor_refs: tuple[torch.Tensor, ...] = ()
        self._projection_cache_hits = 0
        self._projection_cache_misses = 0
        self._allocated_visible_capacity = 0
        self._allocated_intersection_capacity = 0
        self._allocated_scratch_gaussians = 0
        self._initial_visible_capacity = 0
        self._initial_intersection_capacity = 0
        self._workspace_config: tuple[int, int, int, torch.device, torch.dtype] | None = None
        self._capacity_validation_token: tuple[Any, ...] | None = None
        self._capacity_validation_tensor_refs: tuple[torch.Tensor, ...] = ()
        self._capacity_validation_counters: dict[str, int] | None = None
        self._capacity_validations = 0
        self._capacity_validation_reuses = 0
        self._capacity_growth_events = 0
        self._capacity_total_retries = 0
        self._capacity_last_render: dict[str, Any] | None = None
        self._capacity_last_growth: dict[str, Any] | None = None
        self._capacity_last_allocation: dict[str, Any] | None = None

    # ------------------------------------------------------------------ lifecycle
    def initialize(self, stage: Any, device: torch.device) -> None:
        if device.type != "cuda" and not self.allow_cpu_for_tests:
            raise ValueError(f"CustomHipBackend requires a GPU device; got {device}.")
        self.stage = stage
        self.device = device
        if self._native is None:
            self._native = self._native_loader()
        self.invalidate_projection_cache()
        self._projection_cache_hits = 0
        self._projection_cache_misses = 0
        self._capacity_validations = 0
        self._capacity_validation_reuses = 0
        self._capacity_growth_events = 0
        self._capacity_total_retries = 0
        self._capacity_last_render = None
        self._capacity_last_growth = None
        self._capacity_last_allocation = None

    def register_scene(self, scene_id: int, scene: GaussianScene, offset: int) -> None:
        self._require_initialized()
        expected_offset = sum(registered.count for registered in self._scenes.values())
        if offset != expected_offset:
            raise ValueError(f"Scene {scene_id} offset {offset} does not match packed offset {expected_offset}.")
        if scene.features.shape[1] < 3:
            raise ValueError("CustomHipBackend requires at least three RGB feature channels.")
        if self._scenes:
            feature_width = next(iter(self._scenes.values())).features.shape[1]
            if scene.features.shape[1] != feature_width:
                raise ValueError("All packed scenes must use the same feature width; "
                                 f"expected {feature_width}, got {scene.features.shape[1]}.")
        if self.sh_degree is not None and 3 * (self.sh_degree + 1) ** 2 > scene.features.shape[1]:
            raise ValueError(f"sh_degree={self.sh_degree} needs {3 * (self.sh_degree + 1) ** 2} feature channels; "
                             f"scene {scene_id} has {scene.features.shape[1]}.")
        self._scenes[scene_id] = scene
        self._rebuild_packed_scene()
        if self.workspace and self._scratch_required() > int(self.workspace["scratch"].numel()):
            self._install_workspace(visible_capacity=self._allocated_visible_capacity,
                                    intersection_capacity=self._allocated_intersection_capacity,
                                    reason="scene-registration")

    def allocate_workspace(self, *, max_views: int, height: int, width: int, device: torch.device,
                           dtype: torch.dtype) -> dict[str, torch.Tensor]:
        self._require_initialized()
        if dtype != torch.float32:
            raise ValueError("The gfx950 backend workspace uses float32.")
        if max_views <= 0:
            raise ValueError("max_views must be positive.")
        physical_max_views = min(max_views, self.max_physical_views or max_views)
        visible_capacity = self.max_visible_records or physical_max_views * self.visible_capacity_per_view
        intersection_capacity = self.max_intersections or math.ceil(visible_capacity * self.intersections_per_visible)
        intersection_capacity = min(intersection_capacity, _SORT_CAPACITY_LIMIT) \
            if self.max_intersections is None else intersection_capacity
        self._max_views = max_views
        self._physical_max_views = physical_max_views
        self._max_logical_chunks = math.ceil(max_views / physical_max_views)
        self._workspace_config = (physical_max_views, height, width, device, dtype)
        self._install_workspace(visible_capacity=visible_capacity, intersection_capacity=intersection_capacity,
                                reason="initial-allocation")
        self._initial_visible_capacity = visible_capacity
        self._initial_intersection_capacity = intersection_capacity
        return self.workspace

    def update_scene_transforms(self, scene_ids: torch.Tensor, transforms: torch.Tensor) -> None:
        self._env_scene_ids = scene_ids
        self._env_transforms = transforms
        self.invalidate_projection_cache()

    def reserve_workspace_capacity(self, *, visible_records: int, tile_intersections: int) -> None:
        """Install an explicit capacity outside measured work (synchronizes)."""
        self._require_initialized()
        if not self.workspace:
            raise RuntimeError("allocate_workspace() must be called before reservation.")
        self.synchronize()
        self._install_workspace(visible_capacity=visible_records, intersection_capacity=tile_intersections,
                                reason="explicit-benchmark-reservation")

    def synchronize(self) -> None:
        self._require_initialized()
        if self.device is not None and self.device.type == "cuda":
            torch.cuda.synchronize(self.device)

    def shutdown(self) -> None:
        self.invalidate_projection_cache()
        self.workspace = {}
        self._scenes.clear()
        self._packed_scene = {}
        self._env_scene_ids = None
        self._env_transforms = None
        self._workspace_config = None
        self._allocated_visible_capacity = 0
        self._allocated_intersection_capacity = 0
        self._allocated_scratch_gaussians = 0
        self._initial_visible_capacity = 0
        self._initial_intersection_capacity = 0
        self.stage = None
        self.device = None

    # --------------------------------------------------------------------- render
    def render(self, request: RenderRequest) -> RenderStats:
        self._require_initialized()
        if not self.workspace:
            raise RuntimeError("allocate_workspace() must be called before render().")
        if not self._packed_scene:
            raise RuntimeError("At least one Gaussian scene must be registered.")
        output_names = tuple(request.outputs)
        if output_names not in (RGB_ONLY_OUTPUTS, FULL_SENSOR_OUTPUTS):
            raise ValueError("CustomHipBackend supports exactly the RGB-only or fused RGB/depth/alpha/semantic_id "
                             f"output contract; got {output_names}.")
        full_sensor_output = output_names == FULL_SENSOR_OUTPUTS
        batch = int(request.camera_transforms.shape[0])
        if batch > self._max_views:
            raise ValueError(f"Render batch {batch} exceeds workspace max_views {self._max_views}.")
        if request.outputs["rgb"].shape[1:3] != (self._height, self._width):
            raise ValueError("Output resolution does not match the allocated workspace.")

        attempts: list[dict[str, Any]] = []
        for retry_count in range(self.max_capacity_retries + 1):
            if batch > self._physical_max_views:
                stats, input_token, tensor_refs = self._submit_chunked_render(request,
                                                                               full_sensor_output=full_sensor_output)
            else:
                stats, input_token, tensor_refs = self._submit_render(request, full_sensor_output=full_sensor_output)
                if self.max_physical_views is not None:
                    self.workspace["physical_capacity_counters"].copy_(self.workspace["counters"])
            if not self.adaptive_capacity:
                self._capacity_last_render = {
                    "status": "submitted-unvalidated",
                    "validation": "explicit-manual-capacity-mode",
                    "synchronized": False,
                    "retry_count": 0,
                    "attempts": [],
                    "final_capacity": self._capacity_snapshot(),
                    "final_counters": None,
                    "headroom": None,
                }
                return stats
            if input_token == self._capacity_validation_token and self._capacity_validation_counters is not None:
                counters = dict(self._capacity_validation_counters)
                self._capacity_validation_reuses += 1
                attempt = self._capacity_attempt_record(retry_count=retry_count, counters=counters,
                                                        counter_source="validated-identical-input",
                                                        synchronized=False)
                self._capacity_last_render = {
                    "status": "success",
                    "validation": "validated-identical-input",
                    "synchronized": False,
                    "retry_count": retry_count,
                    "attempts": [attempt],
                    "final_capacity": self._capacity_snapshot(),
                    "final_counters": counters,
                    "headroom": self._capacity_headroom(counters),
                }
                return RenderStats(rendered_views=stats.rendered_views,
                                   visible_gaussians=counters["visible_gaussians"],
                                   tile_intersections=counters["tile_intersections"])
            counters = self.read_counters(synchronize=True)
            attempt = self._capacity_attempt_record(retry_count=retry_count, counters=counters,
                                                    counter_source="device-counters", synchronized=True)
            attempts.append(attempt)
            overflow = bool(counters["visible_overflow"] or counters["intersection_overflow"])
            if not overflow:
                self._capacity_validation_token = input_token
                self._capacity_validation_tensor_refs = tensor_refs
                self._capacity_validation_counters = dict(counters)
                self._capacity_validations += 1
                self._capacity_total_retries += retry_count
                record = {
                    "status": "success",
                    "validation": "device-counters",
                    "synchronized": True,
                    "retry_count": retry_count,
                    "attempts": attempts,
                    "final_capacity": self._capacity_snapshot(),
                    "final_counters": counters,
                    "headroom": self._capacity_headroom(counters),
                }
                self._capacity_last_render = record
                if retry_count:
                    self._capacity_last_growth = copy.deepcopy(record)
                return RenderStats(rendered_views=stats.rendered_views,
                                   visible_gaussians=counters["visible_gaussians"],
                                   tile_intersections=counters["tile_intersections"])
            if retry_count >= self.max_capacity_retries:
                self._capacity_total_retries += retry_count
                self._raise_render_capacity_error(
                    request,
                    message=f"Custom renderer capacity remained insufficient after {retry_count} bounded retries.",
                    attempts=attempts, reason="retry-limit")
            proposed_visible, proposed_intersections, ceiling_adjustment = self._propose_capacity_growth(counters)
            attempt["growth_request"] = {"visible_records": proposed_visible,
                                         "tile_intersections": proposed_intersections}
            if ceiling_adjustment is not None:
                attempt["workspace_ceiling_adjustment"] = ceiling_adjustment
            try:
                self._install_workspace(visible_capacity=proposed_visible,
                                        intersection_capacity=proposed_intersections, reason="overflow-retry")
            except RendererCapacityError as error:
                attempt["growth_failure"] = copy.deepcopy(error.telemetry)
                self._capacity_total_retries += retry_count
                self._raise_render_capacity_error(request, message=str(error), attempts=attempts,
                                                  reason="growth-allocation-failed", cause=error)
            self._capacity_growth_events += 1
        raise AssertionError("Unreachable capacity retry state.")

    def _scene_tensor_list(self) -> list[torch.Tensor]:
        packed = self._packed_scene
        return [packed["means"], packed["covariances"], packed["opacities"], packed["features"],
                packed["semantic_ids"], packed["registered_scene_ids"], packed["scene_offsets"]]

    def _workspace_tensor_list(self, counters: torch.Tensor | None = None) -> list[torch.Tensor]:
        tensors = [self.workspace[name] for name in WORKSPACE_TENSORS]
        if counters is not None:
            tensors[WORKSPACE_TENSORS.index("counters")] = counters
        return tensors

    def _native_render(self, scene, cameras, outputs, workspace, *, reuse_projection: bool,
                       full_sensor_output: bool) -> int:
        assert self._native is not None
        instrumentation: dict[str, Any] = {}
        if self.stage_events is not None:
            instrumentation["stage_events"] = self.stage_events
        if self.statistics is not None:
            instrumentation["statistics"] = self.statistics
        selector = int(self._native.render(
            scene, cameras, outputs, workspace, self._height, self._width, self._max_scene_gaussians,
            self.near_plane, self.far_plane, self.gaussian_support_sigma, self.covariance_epsilon,
            self.rasterize_mode == "antialiased", self.semantic_min_alpha, False, self.tile_size,
            self.depth_bucket_count, self.depth_bucket_group_size, False, False, reuse_projection,
            self.output_srgb, self.deterministic, full_sensor_output, self.fixed_capacity_sort,
            sh_degree=self._resolved_sh_degree, sort_digit_bits=self.sort_digit_bits,
            raster_batch=self.raster_batch, cameras_per_block=self.cameras_per_block,
            sort_tile_items=self.sort_tile_items, **instrumentation))
        if selector not in (0, 1):
            raise RuntimeError(f"Native renderer returned an invalid sorted-buffer selector: {selector}.")
        if selector == 0:
            # Keep values_out as the sorted buffer without copying it.
            self._swap_sorted_buffers()
        return selector

    def _swap_sorted_buffers(self) -> None:
        ws = self.workspace
        ws["keys_in"], ws["keys_out"] = ws["keys_out"], ws["keys_in"]
        ws["values_in"], ws["values_out"] = ws["values_out"], ws["values_in"]

    def _submit_render(self, request: RenderRequest, *, full_sensor_output: bool):
        batch = int(request.camera_transforms.shape[0])
        if self._env_transforms is None:
            env_xforms = self.workspace["identity_env_xforms"][:batch]
        else:
            if self._env_transforms.shape[0] != batch:
                raise ValueError("update_scene_transforms() must provide one transform per render camera; "
                                 f"got {self._env_transforms.shape[0]} transforms for batch {batch}.")
            env_xforms = self._env_transforms
        active_ids = (request.active_camera_ids if request.active_camera_ids is not None
                      else self.workspace["empty_active_ids"])
        token: tuple[Any, ...] | None = None
        reuse_projection = False
        if self.enable_projection_cache or self.adaptive_capacity:
            token = self._projection_input_token(batch=batch, request=request, env_xforms=env_xforms,
                                                 active_ids=active_ids)
        if self.enable_projection_cache:
            reuse_projection = token == self._projection_cache_token
        cameras = [request.camera_transforms, request.intrinsics, env_xforms, request.scene_ids, active_ids]
        outputs = [
            request.outputs["rgb"],
            request.outputs["depth"] if full_sensor_output else self.workspace["dummy_depth"],
            request.outputs["alpha"] if full_sensor_output else self.workspace["dummy_alpha"],
            request.outputs["semantic_id"] if full_sensor_output else self.workspace["dummy_semantic"],
        ]
        if request.active_camera_ids is not None:
            # Kernels visit active cameras only; define every inactive output.
            request.outputs["rgb"].zero_()
            if full_sensor_output:
                request.outputs["depth"].fill_(float("inf"))
                request.outputs["alpha"].zero_()
                request.outputs["semantic_id"].fill_(-1)
        self._native_render(self._scene_tensor_list(), cameras, outputs, self._workspace_tensor_list(),
                            reuse_projection=reuse_projection, full_sensor_output=full_sensor_output)
        if self.enable_projection_cache:
            self._projection_cache_token = token
            # Hold the tensors whose storage identities form the token so the
            # allocator cannot recycle their addresses into a false hit.
            self._projection_cache_tensor_refs = tuple(cameras)
            if reuse_projection:
                self._projection_cache_hits += 1
            else:
                self._projection_cache_misses += 1
        rendered = int(request.active_camera_ids.numel()) if request.active_camera_ids is not None else batch
        return RenderStats(rendered_views=rendered), token, tuple(cameras)

    def _submit_chunked_render(self, request: RenderRequest, *, full_sensor_output: bool):
        """Render one logical batch as bounded physical submissions (no host readback)."""
        batch = int(request.camera_transforms.shape[0])
        if self.max_physical_views is None:
            raise AssertionError("Chunked submission requires max_physical_views.")
        if not self.fixed_capacity_sort or self.adaptive_capacity:
            raise AssertionError("Chunked submission requires fixed, preflighted capacities.")
        if self.enable_projection_cache:
            raise AssertionError("Chunked projection-cache reuse is unsupported.")
        chunk_count = math.ceil(batch / self._physical_max_views)
        if chunk_count > self._max_logical_chunks:
            raise ValueError(f"Logical batch requires {chunk_count} chunks but only {self._max_logical_chunks} "
                             "were allocated.")
        env_xforms = None
        if self._env_transforms is not None:
            if self._env_transforms.shape[0] != batch:
                raise ValueError("update_scene_transforms() must provide one transform per logical render camera; "
                                 f"got {self._env_transforms.shape[0]} transforms for batch {batch}.")
            env_xforms = self._env_transforms
        logical_active_ids = request.active_camera_ids
        assert self._native is not None
        if logical_active_ids is not None:
            request.outputs["rgb"].zero_()
            if full_sensor_output:
                request.outputs["depth"].fill_(float("inf"))
                request.outputs["alpha"].zero_()
                request.outputs["semantic_id"].fill_(-1)
            self._native.prepare_chunked_active_ids(logical_active_ids, self.workspace["logical_active_mask"],
                                                    self.workspace["logical_chunk_active_ids"], batch,
                                                    self._physical_max_views)
        chunk_counters = self.workspace["logical_chunk_counters"]
        scene = self._scene_tensor_list()
        for chunk in range(chunk_count):
            start = chunk * self._physical_max_views
            count = min(self._physical_max_views, batch - start)
            cameras = [
                request.camera_transforms.narrow(0, start, count),
                request.intrinsics.narrow(0, start, count),
                (self.workspace["identity_env_xforms"].narrow(0, 0, count) if env_xforms is None
                 else env_xforms.narrow(0, start, count)),
                request.scene_ids.narrow(0, start, count),
                (self.workspace["empty_active_ids"] if logical_active_ids is None
                 else self.workspace["logical_chunk_active_ids"][chunk].narrow(0, 0, count)),
            ]
            outputs = [
                request.outputs["rgb"].narrow(0, start, count),
                request.outputs["depth"].narrow(0, start, count) if full_sensor_output
                else self.workspace["dummy_depth"],
                request.outputs["alpha"].narrow(0, start, count) if full_sensor_output
                else self.workspace["dummy_alpha"],
                request.outputs["semantic_id"].narrow(0, start, count) if full_sensor_output
                else self.workspace["dummy_semantic"],
            ]
            self._native_render(scene, cameras, outputs, self._workspace_tensor_list(chunk_counters[chunk]),
                                reuse_projection=False, full_sensor_output=full_sensor_output)
        self._native.aggregate_chunk_counters(chunk_counters, self.workspace["counters"],
                                              self.workspace["physical_capacity_counters"], chunk_count)
        rendered = int(logical_active_ids.numel()) if logical_active_ids is not None else batch
        refs = (request.camera_transforms, request.intrinsics, request.scene_ids,
                *(() if env_xforms is None else (env_xforms,)),
                *(() if logical_active_ids is None else (logical_active_ids,)))
        return RenderStats(rendered_views=rendered), None, refs

    # ------------------------------------------------------------------- counters
    def read_counters(self, *, synchronize: bool = True) -> dict[str, int]:
        """Read the device counter block (outside steady-state submission)."""
        if synchronize:
            self.synchronize()
        values = self.workspace["counters"].detach().cpu().tolist()
        return dict(zip(COUNTER_NAMES, (int(v) for v in values)))

    def read_physical_capacity_counters(self, *, synchronize: bool = True) -> dict[str, int]:
        """Maximum per-chunk demand of the last chunked render."""
        if "physical_capacity_counters" not in self.workspace:
            return self.read_counters(synchronize=synchronize)
        if synchronize:
            self.synchronize()
        values = self.workspace["physical_capacity_counters"].detach().cpu().tolist()
        return dict(zip(COUNTER_NAMES, (int(v) for v in values)))

    def check_capacity(self, *, synchronize: bool = True) -> dict[str, int]:
        counters = self.read_counters(synchronize=synchronize)
        if counters["visible_overflow"] or counters["intersection_overflow"]:
            raise RendererCapacityError(
                "Custom renderer workspace overflow: "
                f"visible={counters['visible_overflow']} intersections={counters['intersection_overflow']}.",
                telemetry={"schema_version": "renderer-capacity-error/v1", "reason": "explicit-check-overflow",
                           "counters": counters, "capacity": self.capacity_stats})
        return counters

    # ------------------------------------------------------------------ workspace
    def _global_tiles(self, views: int) -> int:
        tiles_x = math.ceil(self._workspace_config[2] / self.tile_size) if self._workspace_config else 0
        tiles_y = math.ceil(self._workspace_config[1] / self.tile_size) if self._workspace_config else 0
        return views * tiles_x * tiles_y

    def _scratch_required(self, intersection_capacity: int | None = None) -> int:
        assert self._workspace_config is not None and self._native is not None
        views, height, width, _device, _dtype = self._workspace_config
        capacity = self._allocated_intersection_capacity if intersection_capacity is None else intersection_capacity
        gaussians = int(self._packed_scene["means"].shape[0]) if self._packed_scene else 0
        query = getattr(self._native, "scratch_bytes", None)
        if query is None:
            return int(self._native.sort_temp_bytes(capacity, self._global_tiles(views)))
        return int(query(max_views=views, height=height, width=width, gaussians=gaussians,
                         intersection_capacity=capacity, tile_size=self.tile_size, near_plane=self.near_plane,
                         far_plane=self.far_plane, sort_digit_bits=self.sort_digit_bits,
                         cameras_per_block=self.cameras_per_block, sort_tile_items=self.sort_tile_items))

    def _workspace_plan(self, *, visible_capacity: int, intersection_capacity: int) -> dict[str, Any]:
        if self._workspace_config is None:
            raise RuntimeError("Workspace configuration is unavailable.")
        max_views, height, width, device, _dtype = self._workspace_config
        if visible_capacity <= 0 or intersection_capacity <= 0:
            raise ValueError("Workspace capacities must be positive.")
        if visible_capacity > _INT32_MAX:
            raise ValueError("Visible storage capacity exceeds the int32 implementation limit.")
        if intersection_capacity > _SORT_CAPACITY_LIMIT:
            raise ValueError("Intersection capacity exceeds the int32 implementation limit "
                             f"(the gfx950 sort accepts at most {_SORT_CAPACITY_LIMIT}).")
        global_tiles = self._global_tiles(max_views)
        if global_tiles > _INT32_MAX:
            raise ValueError("Global tile capacity exceeds the int32 implementation limit.")
        key_bits = compressed_sort_key_bits(global_tiles, self.near_plane, self.far_plane)
        digit_bits = self.sort_digit_bits or default_sort_digit_bits(key_bits)
        scratch_bytes = self._scratch_required(intersection_capacity)
        specs: dict[str, tuple[tuple[int, ...], torch.dtype]] = {
            "visible_records": ((visible_capacity, RECORD_FLOATS), torch.float32),
            "keys_in": ((intersection_capacity,), torch.uint64),
            "keys_out": ((intersection_capacity,), torch.uint64),
            "values_in": ((intersection_capacity,), torch.int32),
            "values_out": ((intersection_capacity,), torch.int32),
            "tile_starts": ((global_tiles,), torch.int32),
            "tile_ends": ((global_tiles,), torch.int32),
            "counters": ((len(COUNTER_NAMES),), torch.int64),
            "scratch": ((scratch_bytes,), torch.uint8),
            "identity_env_xforms": ((max_views, 4, 4), torch.float32),
            "empty_active_ids": ((0,), torch.int64),
            "dummy_depth": ((1,), torch.float32),
            "dummy_alpha": ((1,), torch.float32),
            "dummy_semantic": ((1,), torch.int64),
        }
        if self.max_physical_views is not None:
            specs.update({
                "logical_chunk_counters": ((self._max_logical_chunks, len(COUNTER_NAMES)), torch.int64),
                "physical_capacity_counters": ((len(COUNTER_NAMES),), torch.int64),
                "logical_active_mask": ((self._max_views,), torch.int32),
                "logical_chunk_active_ids": ((self._max_logical_chunks, max_views), torch.int64),
            })
        workspace_bytes = sum(math.prod(shape) * _DTYPE_BYTES[dtype] for shape, dtype in specs.values())
        return {
            "visible_capacity": visible_capacity,
            "visible_storage_capacity": visible_capacity,
            "intersection_capacity": intersection_capacity,
            "global_tiles": global_tiles,
            "workspace_bytes": workspace_bytes,
            "workspace_logical_bytes": workspace_bytes,
            "workspace_unaliased_bytes": workspace_bytes,
            "workspace_alias_savings_bytes": 0,
            "sort_temp_bytes": scratch_bytes,
            "sort_temp_strategy": "device-count-onesweep-radix",
            "sort_key_begin_bit": 0,
            "sort_key_end_bit": _emitted_sort_key_end_bit(max(global_tiles, 1)),
            "sort_key_bits": key_bits,
            "sort_radix_digit_bits_assumption": digit_bits,
            "sort_radix_passes_expected": math.ceil(key_bits / digit_bits),
            "workspace_storage_aliases": {},
            "device": device,
            "specs": specs,
        }

    def _install_workspace(self, *, visible_capacity: int, intersection_capacity: int, reason: str) -> None:
        try:
            plan = self._workspace_plan(visible_capacity=visible_capacity,
                                        intersection_capacity=intersection_capacity)
        except ValueError as error:
            if reason == "initial-allocation":
                raise
            raise RendererCapacityError(str(error), telemetry={
                "schema_version": "renderer-capacity-allocation/v1",
                "reason": reason,
                "failure": "native-int32-limit",
                "requested_capacity": {"visible_records": visible_capacity,
                                       "tile_intersections": intersection_capacity},
                "native_limit": _SORT_CAPACITY_LIMIT,
            }) from error
        requested_bytes = int(plan["workspace_bytes"])
        current_bytes = self.workspace_bytes
        release_before_allocate = reason in {"overflow-retry", "explicit-benchmark-reservation",
                                             "scene-registration"} and bool(self.workspace)
        telemetry: dict[str, Any] = {
            "schema_version": "renderer-capacity-allocation/v1",
            "reason": reason,
            "requested_capacity": {"visible_records": visible_capacity,
                                   "visible_storage_records": int(plan["visible_storage_capacity"]),
                                   "tile_intersections": intersection_capacity},
            "requested_workspace_bytes": requested_bytes,
            "requested_workspace_unaliased_bytes": int(plan["workspace_unaliased_bytes"]),
            "workspace_alias_savings_bytes": int(plan["workspace_alias_savings_bytes"]),
            "workspace_storage_aliases": dict(plan["workspace_storage_aliases"]),
            "sort_temp_bytes": int(plan["sort_temp_bytes"]),
            "sort_temp_strategy": str(plan["sort_temp_strategy"]),
            "sort_key_begin_bit": int(plan["sort_key_begin_bit"]),
            "sort_key_end_bit": int(plan["sort_key_end_bit"]),
            "sort_key_bits": int(plan["sort_key_bits"]),
            "sort_radix_digit_bits_assumption": int(plan["sort_radix_digit_bits_assumption"]),
            "sort_radix_passes_expected": int(plan["sort_radix_passes_expected"]),
            "current_workspace_bytes": current_bytes,
            "allocation_strategy": ("release-old-workspace-first" if release_before_allocate
                                    else "transactional-replacement"),
            "transient_logical_bytes": (max(current_bytes, requested_bytes) if release_before_allocate
                                        else current_bytes + requested_bytes),
            "transactional_transient_logical_bytes": current_bytes + requested_bytes,
            "max_workspace_bytes": self.max_workspace_bytes,
        }
        if self.max_workspace_bytes is not None and requested_bytes > self.max_workspace_bytes:
            telemetry["failure"] = "workspace-memory-limit"
            telemetry["status"] = "failure"
            self._capacity_last_allocation = copy.deepcopy(telemetry)
            raise RendererCapacityError(
                f"Adaptive capacity growth requires {requested_bytes} workspace bytes, exceeding the configured "
                f"limit of {self.max_workspace_bytes} bytes.", telemetry=telemetry)
        device = plan["device"]
        old_plan: dict[str, Any] | None = None
        if release_before_allocate:
            old_plan = self._workspace_plan(visible_capacity=self._allocated_visible_capacity,
                                            intersection_capacity=self._allocated_intersection_capacity)
            self.invalidate_projection_cache()
            self.workspace.clear()
            if device.type == "cuda":
                with torch.cuda.device(device):
                    torch.cuda.empty_cache()
        try:
            replacement = self._allocate_workspace_plan(plan)
        except torch.OutOfMemoryError as error:
            telemetry["failure"] = "allocator-out-of-memory"
            telemetry["status"] = "failure"
            if release_before_allocate:
                assert old_plan is not None
                if device.type == "cuda":
                    with torch.cuda.device(device):
                        torch.cuda.empty_cache()
                try:
                    restored = self._allocate_workspace_plan(old_plan)
                except torch.OutOfMemoryError:
                    self.workspace.clear()
                    self._allocated_visible_capacity = 0
                    self._allocated_intersection_capacity = 0
                    telemetry["restoration"] = {"status": "failed", "workspace_usable": False}
                else:
                    self.workspace.update(restored)
                    telemetry["restoration"] = {"status": "restored-old-capacity", "workspace_usable": True,
                                                "workspace_bytes": self.workspace_bytes}
            self._capacity_last_allocation = copy.deepcopy(telemetry)
            raise RendererCapacityError(
                f"GPU allocator could not satisfy bounded adaptive workspace growth to {requested_bytes} bytes.",
                telemetry=telemetry) from error
        physical_max_views, height, width, _device, _dtype = self._workspace_config
        # Keep the dict identity RendererService holds; replace only its contents.
        self.workspace.clear()
        self.workspace.update(replacement)
        self._allocated_visible_capacity = visible_capacity
        self._allocated_intersection_capacity = intersection_capacity
        self._physical_max_views = physical_max_views
        self._height = height
        self._width = width
        self.invalidate_projection_cache()
        telemetry["status"] = "success"
        telemetry["installed_workspace_bytes"] = self.workspace_bytes
        telemetry["installed_workspace_storage_aliases"] = self.workspace_storage_aliases
        self._capacity_last_allocation = copy.deepcopy(telemetry)

    def _allocate_workspace_plan(self, plan: dict[str, Any]) -> dict[str, torch.Tensor]:
        device = plan["device"]
        replacement: dict[str, torch.Tensor] = {}
        try:
            for name, (shape, dtype) in plan["specs"].items():
                replacement[name] = torch.empty(shape, device=device, dtype=dtype)
            replacement["identity_env_xforms"].copy_(
                torch.eye(4, device=device, dtype=torch.float32).repeat(self._workspace_config[0], 1, 1))
            replacement["counters"].zero_()
            # Diagnostic views into the 48-byte visible records (no extra storage).
            records_i32 = replacement["visible_records"].view(torch.int32)
            replacement["visible_gaussian_ids"] = records_i32[:, 10]
            replacement["visible_camera_ids"] = records_i32[:, 11]
            replacement["visible_depths"] = replacement["visible_records"][:, 6]
        except torch.OutOfMemoryError:
            replacement.clear()
            raise
        return replacement

    def _propose_capacity_growth(self, counters: dict[str, int]) -> tuple[int, int, dict[str, Any] | None]:
        visible_capacity = self._allocated_visible_capacity
        intersection_capacity = self._allocated_intersection_capacity
        visible_required = max(int(counters["visible_gaussians"]),
                               visible_capacity + int(counters["visible_overflow"]))
        proposed_visible = visible_capacity
        if counters["visible_overflow"]:
            proposed_visible = max(visible_capacity + 1, math.ceil(visible_capacity * self.capacity_growth_factor),
                                   math.ceil(visible_required * self.capacity_headroom))
        intersection_required = max(int(counters["tile_intersections"]),
                                    intersection_capacity + int(counters["intersection_overflow"]))
        full_demand = bool(getattr(self._native, "counts_full_intersection_demand", False))
        if proposed_visible > visible_capacity and not full_demand:
            # Counters that cover stored records only: extrapolate the demand of
            # the records that did not fit.
            materialized = max(1, min(int(counters["visible_gaussians"]), visible_capacity))
            intersection_required = max(intersection_required,
                                        math.ceil(int(counters["tile_intersections"]) * visible_required
                                                  / materialized))
        proposed_intersections = intersection_capacity
        if counters["intersection_overflow"] or proposed_visible > visible_capacity:
            proposed_intersections = max(intersection_capacity + 1,
                                         math.ceil(intersection_capacity * self.capacity_growth_factor),
                                         math.ceil(intersection_required * self.capacity_headroom))
        ceiling_adjustment: dict[str, Any] | None = None
        if self.max_workspace_bytes is not None:
            try:
                target_plan = self._workspace_plan(visible_capacity=proposed_visible,
                                                   intersection_capacity=proposed_intersections)
            except ValueError:
                pass  # the native-limit error is raised by _install_workspace
            else:
                target_bytes = int(target_plan["workspace_bytes"])
                if target_bytes > self.max_workspace_bytes:
                    minimum_visible = visible_capacity
                    if counters["visible_overflow"]:
                        minimum_visible = max(visible_capacity + 1, visible_required)
                    minimum_intersections = intersection_capacity
                    if counters["intersection_overflow"] or minimum_visible > visible_capacity:
                        minimum_intersections = max(intersection_capacity + 1, intersection_required)
                    minimum_plan = self._workspace_plan(visible_capacity=minimum_visible,
                                                        intersection_capacity=minimum_intersections)
                    minimum_bytes = int(minimum_plan["workspace_bytes"])
                    ceiling_adjustment = {
                        "policy": "relax-headroom-to-observed-demand",
                        "target_capacity": {"visible_records": proposed_visible,
                                            "tile_intersections": proposed_intersections},
                        "target_workspace_bytes": target_bytes,
                        "minimum_observed_demand_capacity": {"visible_records": minimum_visible,
                                                             "tile_intersections": minimum_intersections},
                        "minimum_observed_demand_workspace_bytes": minimum_bytes,
                        "max_workspace_bytes": self.max_workspace_bytes,
                        "status": ("headroom-relaxed" if minimum_bytes <= self.max_workspace_bytes
                                   else "minimum-observed-demand-exceeds-limit"),
                    }
                    proposed_visible = minimum_visible
                    proposed_intersections = minimum_intersections
        return proposed_visible, proposed_intersections, ceiling_adjustment

    def _capacity_snapshot(self) -> dict[str, int]:
        return {
            "visible_records": self._allocated_visible_capacity,
            "visible_storage_records": self.visible_storage_capacity,
            "tile_intersections": self._allocated_intersection_capacity,
            "workspace_bytes": self.workspace_bytes,
        }

    def _capacity_attempt_record(self, *, retry_count: int, counters: dict[str, int], counter_source: str,
                                 synchronized: bool) -> dict[str, Any]:
        return {
            "attempt": retry_count + 1,
            "requested_capacity": self._capacity_snapshot(),
            "actual_demand": dict(counters),
            "counter_source": counter_source,
            "synchronized": synchronized,
            "overflow": bool(counters["visible_overflow"] or counters["intersection_overflow"]),
        }

    def _capacity_headroom(self, counters: dict[str, int]) -> dict[str, float | None]:
        visible = int(counters["visible_gaussians"])
        intersections = int(counters["tile_intersections"])
        return {
            "visible_records_ratio": self._allocated_visible_capacity / visible if visible else None,
            "tile_intersections_ratio": (self._allocated_intersection_capacity / intersections
                                         if intersections else None),
        }

    def _raise_render_capacity_error(self, request: RenderRequest, *, message: str, attempts: list[dict[str, Any]],
                                     reason: str, cause: BaseException | None = None) -> NoReturn:
        self.invalidate_projection_cache()
        for name, tensor in request.outputs.items():
            if tensor.is_floating_point():
                tensor.fill_(float("nan"))
            elif name == "semantic_id":
                tensor.fill_(-1)
        final_counters = copy.deepcopy(attempts[-1]["actual_demand"]) if attempts else None
        self._capacity_last_render = {
            "status": "failure",
            "validation": "device-counters",
            "synchronized": any(bool(a["synchronized"]) for a in attempts),
            "retry_count": max(0, len(attempts) - 1),
            "attempts": copy.deepcopy(attempts),
            "final_capacity": self._capacity_snapshot(),
            "final_counters": final_counters,
            "headroom": self._capacity_headroom(final_counters) if final_counters is not None else None,
            "failure_reason": reason,
        }
        if any("growth_request" in attempt for attempt in attempts):
            self._capacity_last_growth = copy.deepcopy(self._capacity_last_render)
        telemetry = {"schema_version": "renderer-capacity-error/v1", "reason": reason,
                     "capacity": self.capacity_stats}
        if attempts:
            demand = attempts[-1]["actual_demand"]
            requested = attempts[-1]["requested_capacity"]
            message = (f"{message} Last attempt: visible_gaussians={demand['visible_gaussians']} "
                       f"tile_intersections={demand['tile_intersections']} "
                       f"visible_capacity={requested['visible_records']} "
                       f"intersection_capacity={requested['tile_intersections']} "
                       f"visible_overflow={demand['visible_overflow']} "
                       f"intersection_overflow={demand['intersection_overflow']}.")
        if cause is None:
            raise RendererCapacityError(message, telemetry=telemetry)
        raise RendererCapacityError(message, telemetry=telemetry) from cause

    # ------------------------------------------------------------------ telemetry
    @property
    def workspace_bytes(self) -> int:
        return sum(self.workspace_bytes_by_tensor.values())

    @property
    def workspace_bytes_by_tensor(self) -> dict[str, int]:
        """Bytes of each unique storage, attributed to its first named view."""
        seen: set[tuple[str, int | None, int]] = set()
        result: dict[str, int] = {}
        for name, tensor in self.workspace.items():
            storage = tensor.untyped_storage()
            nbytes = int(storage.nbytes())
            key = (tensor.device.type, tensor.device.index, int(storage.data_ptr()))
            if nbytes == 0 or key in seen:
                result[name] = 0
                continue
            seen.add(key)
            result[name] = nbytes
        return result

    @property
    def workspace_logical_bytes_by_tensor(self) -> dict[str, int]:
        return {name: t.numel() * t.element_size() for name, t in self.workspace.items()}

    @property
    def workspace_storage_aliases(self) -> dict[str, str]:
        owners: dict[tuple[str, int | None, int], str] = {}
        aliases: dict[str, str] = {}
        for name, tensor in self.workspace.items():
            storage = tensor.untyped_storage()
            if storage.nbytes() == 0:
                continue
            key = (tensor.device.type, tensor.device.index, int(storage.data_ptr()))
            owner = owners.setdefault(key, name)
            if owner != name:
                aliases[name] = owner
        return aliases

    @property
    def current_visible_capacity(self) -> int:
        return self._allocated_visible_capacity

    @property
    def current_intersection_capacity(self) -> int:
        return self._allocated_intersection_capacity

    @property
    def visible_storage_capacity(self) -> int:
        records = self.workspace.get("visible_records")
        return int(records.shape[0]) if records is not None else 0

    @property
    def capacity_stats(self) -> dict[str, Any]:
        """JSON-serializable capacity provenance (schema renderer-capacity/v1)."""
        return {
            "schema_version": "renderer-capacity/v1",
            "policy": "adaptive-bounded-grow-retry" if self.adaptive_capacity else "explicit-manual-check",
            "adaptive": self.adaptive_capacity,
            "initial_capacity": {"visible_records": self._initial_visible_capacity,
                                 "tile_intersections": self._initial_intersection_capacity},
            "current_capacity": self._capacity_snapshot(),
            "execution_schedule": {
                "logical_max_views": self._max_views,
                "physical_max_views": self._physical_max_views,
                "maximum_logical_chunks": self._max_logical_chunks,
                "chunking_enabled": self.max_physical_views is not None,
                "counter_reduction": ("device-sum-and-physical-maximum" if self.max_physical_views is not None
                                      else "single-native-submission"),
            },
            "visible_capacity_role": "materialized-projected-candidate-capacity",
            "projected_record_reuse": {"requested": False, "active": False, "fallback_reason": None,
                                       "bytes_per_record": _RECORD_BYTES},
            "parameters": {
                "rasterize_mode": self.rasterize_mode,
                "covariance_epsilon": self.covariance_epsilon,
                "growth_factor": self.capacity_growth_factor,
                "headroom_factor": self.capacity_headroom,
                "max_retries": self.max_capacity_retries,
                "max_workspace_bytes": self.max_workspace_bytes,
                "workspace_ceiling_policy": "relax-headroom-to-observed-demand",
            },
            "totals": {
                "device_validations": self._capacity_validations,
                "validated_input_reuses": self._capacity_validation_reuses,
                "growth_events": self._capacity_growth_events,
                "retries": self._capacity_total_retries,
            },
            "synchronization_policy": (
                "new or explicitly invalidated projection inputs read device counters; identical validated "
                "inputs remain asynchronous" if self.adaptive_capacity
                else "caller must explicitly synchronize and check counters"),
            "intersection_ordering_synchronization": (
                "device-count onesweep radix sort: no device-to-host count copy or stream synchronization"),
            "sort_temp_strategy": "device-count-onesweep-radix",
            "last_render": copy.deepcopy(self._capacity_last_render),
            "last_growth": copy.deepcopy(self._capacity_last_growth),
            "last_allocation": copy.deepcopy(self._capacity_last_allocation),
        }

    @property
    def scene_bytes(self) -> int:
        return sum(t.numel() * t.element_size() for t in self._packed_scene.values())

    @property
    def projection_cache_stats(self) -> dict[str, int | bool]:
        return {"enabled": self.enable_projection_cache, "hits": self._projection_cache_hits,
                "misses": self._projection_cache_misses}

    @property
    def pipeline_name(self) -> str:
        return ("project-count-scan-write-emit-onesweep-radix-linear-ranges-"
                "lds-batched-fused-rgba-depth-semantic")

    @property
    def bin_tile_size(self) -> int:
        return self.tile_size

    @property
    def rasterize_mode(self) -> Literal["classic", "antialiased"]:
        return self._rasterize_mode

    @rasterize_mode.setter
    def rasterize_mode(self, value: Literal["classic", "antialiased"]) -> None:
        if value not in ("classic", "antialiased"):
            raise ValueError("rasterize_mode must be 'classic' or 'antialiased'.")
        if value == self._rasterize_mode:
            return
        self._rasterize_mode = value
        self.invalidate_projection_cache()

    @property
    def resolved_sh_degree(self) -> int:
        return self._resolved_sh_degree

    def invalidate_projection_cache(self) -> None:
        """Drop cached geometry and capacity validation (call after writes the
        tensor version counters cannot see)."""
        self._projection_cache_token = None
        self._projection_cache_tensor_refs = ()
        self._capacity_validation_token = None
        self._capacity_validation_tensor_refs = ()
        self._capacity_validation_counters = None

    # ---------------------------------------------------------------------- scene
    def _rebuild_packed_scene(self) -> None:
        assert self.device is not None
        scenes = list(self._scenes.values())
        packed: dict[str, torch.Tensor] = {}
        for name in ("means", "scales", "rotations", "opacities", "features", "semantic_ids"):
            tensors = [getattr(scene, name) for scene in scenes]
            packed[name] = tensors[0] if len(tensors) == 1 else torch.cat(tensors, dim=0)
        packed["covariances"] = self._precompute_covariances(packed["scales"], packed["rotations"])
        offsets = [0]
        for scene in scenes:
            offsets.append(offsets[-1] + scene.count)
        packed["registered_scene_ids"] = torch.tensor(list(self._scenes), device=self.device, dtype=torch.int64)
        packed["scene_offsets"] = torch.tensor(offsets, device=self.device, dtype=torch.int64)
        self._packed_scene = packed
        self._max_scene_gaussians = max(scene.count for scene in scenes)
        width = int(packed["features"].shape[1])
        self._resolved_sh_degree = self.sh_degree if self.sh_degree is not None else infer_sh_degree(width)
        self._scene_revision += 1
        self.invalidate_projection_cache()

    def _projection_input_token(self, *, batch: int, request: RenderRequest, env_xforms: torch.Tensor,
                                active_ids: torch.Tensor) -> tuple[Any, ...]:
        return (
            self._scene_revision, batch, self._height, self._width, self._max_scene_gaussians, self.near_plane,
            self.far_plane, self.gaussian_support_sigma, self.covariance_epsilon, self.rasterize_mode,
            self.tile_size, self.deterministic, self._resolved_sh_degree,
            self._tensor_cache_token(self._packed_scene["means"]),
            self._tensor_cache_token(self._packed_scene["covariances"]),
            self._tensor_cache_token(self._packed_scene["opacities"]),
            self._tensor_cache_token(self._packed_scene["features"]),
            self._tensor_cache_token(self._packed_scene["registered_scene_ids"]),
            self._tensor_cache_token(self._packed_scene["scene_offsets"]),
            self._tensor_cache_token(request.camera_transforms),
            self._tensor_cache_token(request.intrinsics),
            self._tensor_cache_token(env_xforms),
            self._tensor_cache_token(request.scene_ids),
            self._tensor_cache_token(active_ids),
        )

    @staticmethod
    def _tensor_cache_token(tensor: torch.Tensor) -> tuple[Any, ...]:
        try:
            version: int | object = int(tensor._version)
        except RuntimeError:
            # Inference tensors carry no version counter, so in-place writes are
            # invisible; a fresh object makes the token non-reusable.
            version = object()
        return (tensor.data_ptr(), version, tuple(tensor.shape), tuple(tensor.stride()), tensor.dtype, tensor.device)

    def _precompute_covariances(self, scales: torch.Tensor, rotations: torch.Tensor) -> torch.Tensor:
        assert self.device is not None and self._native is not None
        covariances = torch.empty((scales.shape[0], 6), device=self.device, dtype=torch.float32)
        if self.device.type == "cuda":
            self._native.precompute_covariances(scales, rotations, covariances)
            return covariances
        # Host tensors only exist in CPU contract tests with an injected native fake.
        with torch.no_grad():
            q = rotations * torch.rsqrt(rotations.square().sum(dim=1, keepdim=True).clamp_min(1.0e-20))
            w, x, y, z = q.unbind(dim=1)
            rotation = torch.stack((
                1.0 - 2.0 * (y * y + z * z), 2.0 * (x * y - w * z), 2.0 * (x * z + w * y),
                2.0 * (x * y + w * z), 1.0 - 2.0 * (x * x + z * z), 2.0 * (y * z - w * x),
                2.0 * (x * z - w * y), 2.0 * (y * z + w * x), 1.0 - 2.0 * (x * x + y * y)), dim=1).reshape(-1, 3, 3)
            full = (rotation * scales.square().unsqueeze(1)) @ rotation.transpose(1, 2)
            covariances.copy_(torch.stack((full[:, 0, 0], full[:, 0, 1], full[:, 0, 2], full[:, 1, 1],
                                           full[:, 1, 2], full[:, 2, 2]), dim=1))
        return covariances

    def _require_initialized(self) -> None:
        if self.device is None:
            raise RuntimeError("CustomHipBackend.initialize() must be called first.")


# Drop-in name: code written against the reference imports CustomCudaBackend.
CustomCudaBackend = CustomHipBackend
