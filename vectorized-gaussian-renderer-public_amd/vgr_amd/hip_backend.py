"""gfx950 renderer backend: the drop-in for the reference's ``CustomCudaBackend``.

Public contract kept from ``src/isaacsim_gaussian_renderer/cuda_backend.py``
(Apache-2.0): constructor knobs and defaults (see :mod:`vgr_amd.options`),
the ``GaussianRendererBackend`` lifecycle, the RGB-only / full-sensor output
sets, ``RendererCapacityError(message, telemetry=...)``, the counter block,
``read_counters`` / ``check_capacity`` / ``reserve_workspace_capacity`` /
``invalidate_projection_cache``, bounded adaptive capacity with the
reference's growth arithmetic (:mod:`vgr_amd.capacity`), fail-closed outputs
(NaN floats, semantic -1) and ``capacity_stats`` under the
``renderer-capacity/v1`` schema.

How it works here:
  * one native submission never synchronizes the host (the sort reads the
    intersection count on the device), so ``fixed_capacity_sort`` changes
    nothing and every mode orders intersections by (tile, depth, Gaussian ID);
  * the workspace is one arena allocation (:mod:`vgr_amd.workspace`);
  * adaptive mode validates each new input set once with a counter read; an
    identical validated input set is submitted without reading back;
  * a radix-sort look-back wait that gave up (counter slot 5) invalidates the
    render like an overflow: it is re-rendered without growth and fails
    closed when the retries run out;
  * SH bands 1..3 in ``features`` are evaluated (``sh_degree``; ``sh_degree=0``
    reproduces the reference's base-colour-only rule).
"""

from __future__ import annotations

import copy
import math
from collections.abc import Callable
from dataclasses import dataclass, field, replace
from typing import Any, Literal, NoReturn

import torch

from .backend import RenderRequest, RenderStats
from .capacity import (
    COUNTER_NAMES,
    SORT_FAILURES,
    Capacity,
    CapacityLedger,
    CapacityPolicy,
    headroom_ratios,
    overflowed,
)
from .native import (
    COUNTER_SLOTS,
    RECORD_FLOATS,
    SORT_FAILURE_SLOT,
    RAY_RECORD_FLOATS,
    WORKSPACE_TENSORS,
    dense_cutoff,
    float_bits,
    infer_sh_degree,
    load_native_module,
)
from .options import DEFAULT_GAUSSIAN_SUPPORT_SIGMA, BackendOptions
from .scene import GaussianScene, PackedScene, SceneRegistry, pack_scenes
from .workspace import BufferSpec, WorkspaceArena, WorkspacePlan

__all__ = ["COUNTER_NAMES", "CustomCudaBackend", "CustomHipBackend", "DEFAULT_GAUSSIAN_SUPPORT_SIGMA",
           "RendererCapacityError", "compressed_sort_key_bits", "default_sort_digit_bits"]

FULL_SENSOR_OUTPUTS = ("rgb", "depth", "alpha", "semantic_id")
RGB_ONLY_OUTPUTS = ("rgb",)
_INT32_MAX = 2_147_483_647
_SORT_CAPACITY_LIMIT = (1 << 31) - 1  # int32 positions; look-back flags carry 31-bit prefixes
_RECORD_BYTES = 4 * RECORD_FLOATS
_CAPACITY_ERROR_SCHEMA = "renderer-capacity-error/v1"
_ALLOCATION_SCHEMA = "renderer-capacity-allocation/v1"


def emitted_sort_key_end_bit(global_tiles: int) -> int:
    """32 + bit length of the largest tile index (the reference's table, cuda_backend.py:40-47)."""
    if not 0 < global_tiles <= _INT32_MAX:
        raise ValueError("Global tile capacity must be in [1, INT32_MAX] (int32 implementation limit).")
    return 32 + (global_tiles - 1).bit_length()


def default_sort_digit_bits(key_bits: int) -> int:
    """The library's default radix digit width: 8, or 9 where that saves a pass."""
    return 9 if -(-key_bits // 9) < -(-key_bits // 8) else 8


def compressed_sort_key_bits(global_tiles: int, near_plane: float, far_plane: float) -> int:
    """Significant bits of the order-preserving sort-key image used on the device."""
    tile_bits = (global_tiles - 1).bit_length() if global_tiles > 1 else 0
    return tile_bits + (float_bits(far_plane) - float_bits(near_plane) + 1).bit_length()


class RendererCapacityError(RuntimeError):
    """A render whose buffers or ordering could not be made valid within the bounds."""

    def __init__(self, message: str, *, telemetry: dict[str, Any] | None = None) -> None:
        super().__init__(message)
        self.telemetry = copy.deepcopy(telemetry)


@dataclass
class _Fingerprint:
    """Identity of one render's inputs: storage, version counter and geometry of each tensor.

    The tensors are held so their storage cannot be recycled into a false match.
    Inference tensors have no version counter; they never match (every
    submission with one is treated as new).  ``owned`` tensors are views of the
    backend's own workspace arena (identity transforms, the empty active list):
    every view of the arena shares one version counter, which every render
    bumps, so they are identified by storage and geometry only; they are only
    written when a new arena is installed, which changes their storage.
    """

    scalars: tuple[Any, ...]
    tensors: tuple[torch.Tensor, ...]
    owned: tuple[torch.Tensor, ...] = ()
    keys: tuple[Any, ...] = field(init=False)

    def __post_init__(self) -> None:
        self.keys = tuple(self._key(t, versioned=True) for t in self.tensors) + \
            tuple(self._key(t, versioned=False) for t in self.owned)

    @staticmethod
    def _key(tensor: torch.Tensor, *, versioned: bool) -> Any:
        version: Any = None
        if versioned:
            try:
                version = tensor._version
            except RuntimeError:
                version = object()
        return (tensor.data_ptr(), version, tensor.shape, tensor.stride(), tensor.dtype, tensor.device)

    def matches(self, other: "_Fingerprint | None") -> bool:
        return other is not None and self.scalars == other.scalars and self.keys == other.keys


@dataclass(frozen=True)
class _Layout:
    """Fixed geometry of the installed workspace."""

    physical_views: int
    logical_views: int
    height: int
    width: int
    device: torch.device

    @property
    def chunks(self) -> int:
        return math.ceil(self.logical_views / self.physical_views)


class CustomHipBackend:
    """Bounded adaptive Gaussian renderer on gfx950 with device-only submissions."""

    def __init__(self, *, native_module: Any | None = None,
                 native_loader: Callable[[], Any] = load_native_module,
                 allow_cpu_for_tests: bool = False, **knobs: Any) -> None:
        unknown = sorted(set(knobs) - set(BackendOptions.__dataclass_fields__))
        if unknown:
            raise TypeError(f"Unknown CustomHipBackend option(s): {', '.join(unknown)}.")
        self.options = BackendOptions(**knobs).validate()
        self._native = native_module
        self._native_loader = native_loader
        self.allow_cpu_for_tests = allow_cpu_for_tests
        # Optional instrumentation forwarded to the native render: five recorded
        # torch events (stage boundaries) and an int64[4] statistics tensor.
        self.stage_events: list[torch.cuda.Event] | None = None
        self.statistics: torch.Tensor | None = None
        self.device: torch.device | None = None
        self.stage: Any | None = None
        self.workspace: dict[str, torch.Tensor] = {}
        self._registry = SceneRegistry()
        self._packed: PackedScene | None = None
        self._sh_degree = 0
        self._scene_revision = 0
        self._env: tuple[torch.Tensor, torch.Tensor] | None = None
        self._layout: _Layout | None = None
        self._arena: WorkspaceArena | None = None
        self._capacity = Capacity(0, 0)
        self._initial_capacity = Capacity(0, 0)
        self._ledger = CapacityLedger()
        self._validated: tuple[_Fingerprint, dict[str, int]] | None = None
        self._cached_projection: _Fingerprint | None = None
        self._cache_hits = 0
        self._cache_misses = 0

    # Knob attributes read like the reference backend's (backend.tile_size, ...).
    def __getattr__(self, name: str) -> Any:
        options = self.__dict__.get("options")
        if options is not None and name in BackendOptions.__dataclass_fields__:
            return getattr(options, name)
        raise AttributeError(f"{type(self).__name__!s} has no attribute {name!r}")

    # ------------------------------------------------------------------ lifecycle
    def _require_initialized(self) -> torch.device:
        if self.device is None:
            raise RuntimeError("CustomHipBackend.initialize() must be called first.")
        return self.device

    def initialize(self, stage: Any, device: torch.device) -> None:
        if device.type != "cuda" and not self.allow_cpu_for_tests:
            raise ValueError(f"CustomHipBackend requires a GPU device; got {device}.")
        if self._native is None:
            self._native = self._native_loader()
        self.stage = stage
        self.device = device
        self._ledger.reset()
        self._cache_hits = self._cache_misses = 0
        self.invalidate_projection_cache()

    def register_scene(self, scene_id: int, scene: GaussianScene, offset: int) -> None:
        device = self._require_initialized()
        if offset != self._registry.total_gaussians:
            raise ValueError(f"Scene {scene_id} offset {offset} does not match packed offset "
                             f"{self._registry.total_gaussians}.")
        width = int(scene.features.shape[1])
        if width < 3:
            raise ValueError("CustomHipBackend requires at least three RGB feature channels.")
        if self._packed is not None and width != self._packed.feature_width:
            raise ValueError(f"All packed scenes must use the same feature width; expected "
                             f"{self._packed.feature_width}, got {width}.")
        wanted = self.options.sh_degree
        if wanted is not None and 3 * (wanted + 1) ** 2 > width:
            raise ValueError(f"sh_degree={wanted} needs {3 * (wanted + 1) ** 2} feature channels; "
                             f"scene {scene_id} has {width}.")
        self._registry.register(scene_id, scene)
        self._packed = pack_scenes(self._registry, device=device, covariances=self._covariances)
        self._sh_degree = wanted if wanted is not None else infer_sh_degree(width)
        self._scene_revision += 1
        self.invalidate_projection_cache()
        # Kernel scratch scales with the scene (projection blocks): re-plan if it no longer fits.
        if self._arena is not None and self._plan(self._capacity).nbytes > self._arena.nbytes:
            self._install(self._capacity, reason="scene-registration")

    def allocate_workspace(self, *, max_views: int, height: int, width: int, device: torch.device,
                           dtype: torch.dtype) -> dict[str, torch.Tensor]:
        self._require_initialized()
        if dtype != torch.float32:
            raise ValueError("The gfx950 backend workspace uses float32.")
        if max_views <= 0:
            raise ValueError("max_views must be positive.")
        o = self.options
        physical = min(max_views, o.max_physical_views or max_views)
        self._layout = _Layout(physical, max_views, height, width, device)
        visible = o.max_visible_records or physical * o.visible_capacity_per_view
        intersections = o.max_intersections or min(math.ceil(visible * o.intersections_per_visible),
                                                    _SORT_CAPACITY_LIMIT)
        self._install(Capacity(visible, intersections), reason="initial-allocation")
        self._initial_capacity = self._capacity
        return self.workspace

    def update_scene_transforms(self, scene_ids: torch.Tensor, transforms: torch.Tensor) -> None:
        self._env = (scene_ids, transforms)
        self.invalidate_projection_cache()

    def reserve_workspace_capacity(self, *, visible_records: int, tile_intersections: int) -> None:
        """Install an explicit capacity outside measured work (synchronizes)."""
        self._require_initialized()
        if self._arena is None:
            raise RuntimeError("allocate_workspace() must be called before reservation.")
        self.synchronize()
        self._install(Capacity(visible_records, tile_intersections), reason="explicit-benchmark-reservation")

    def synchronize(self) -> None:
        device = self._require_initialized()
        if device.type == "cuda":
            torch.cuda.synchronize(device)

    def shutdown(self) -> None:
        self.invalidate_projection_cache()
        self.workspace.clear()
        self._arena = None
        self._layout = None
        self._registry = SceneRegistry()
        self._packed = None
        self._env = None
        self._capacity = self._initial_capacity = Capacity(0, 0)
        self.stage = None
        self.device = None

    # ------------------------------------------------------------------ workspace
    def _global_tiles(self) -> int:
        layout = self._layout
        assert layout is not None
        ts = self.bin_tile_size
        return layout.physical_views * (-(-layout.width // ts)) * (-(-layout.height // ts))

    def _scratch_bytes(self, intersections: int) -> int:
        layout = self._layout
        assert layout is not None and self._native is not None
        o = self.options
        query = getattr(self._native, "scratch_bytes", None)
        if query is None:  # a native module with only the reference's surface
            return int(self._native.sort_temp_bytes(intersections, self._global_tiles()))
        gaussians = int(self._packed.means.shape[0]) if self._packed is not None else 0
        buckets, group = dense_cutoff(o.tile_size, o.depth_bucket_count, o.depth_bucket_group_size,
                                      o.compact_projection_cache, o.deterministic)
        return int(query(max_views=layout.physical_views, height=layout.height, width=layout.width,
                         gaussians=gaussians, intersection_capacity=intersections, tile_size=o.tile_size,
                         near_plane=o.near_plane, far_plane=o.far_plane, sort_digit_bits=o.sort_digit_bits,
                         cameras_per_block=o.cameras_per_block, sort_tile_items=o.sort_tile_items,
                         depth_bucket_count=buckets, depth_bucket_group_size=group,
                         ray_evaluation=o.ray_gaussian_evaluation))

    def _check_native_limits(self, capacity: Capacity) -> None:
        if capacity.visible_records <= 0 or capacity.tile_intersections <= 0:
            raise ValueError("Workspace capacities must be positive.")
        if capacity.visible_records > _INT32_MAX:
            raise ValueError("Visible storage capacity exceeds the int32 implementation limit.")
        if capacity.tile_intersections > _SORT_CAPACITY_LIMIT:
            raise ValueError("Intersection capacity exceeds the int32 implementation limit "
                             f"(the gfx950 sort accepts at most {_SORT_CAPACITY_LIMIT}).")
        if self._global_tiles() > _INT32_MAX:
            raise ValueError("Global tile capacity exceeds the int32 implementation limit.")

    def _plan(self, capacity: Capacity) -> WorkspacePlan:
        self._check_native_limits(capacity)
        layout = self._layout
        assert layout is not None
        o = self.options
        tiles = self._global_tiles()
        key_bits = compressed_sort_key_bits(tiles, o.near_plane, o.far_plane)
        digit_bits = o.sort_digit_bits or default_sort_digit_bits(key_bits)
        scratch = self._scratch_bytes(capacity.tile_intersections)
        n, v = capacity.tile_intersections, capacity.visible_records
        buffers = {
            "visible_records": BufferSpec((v, RECORD_FLOATS), torch.float32),
            "keys_in": BufferSpec((n,), torch.uint64),
            "keys_out": BufferSpec((n,), torch.uint64),
            "values_in": BufferSpec((n,), torch.int32),
            "values_out": BufferSpec((n,), torch.int32),
            "tile_starts": BufferSpec((tiles,), torch.int32),
            "tile_ends": BufferSpec((tiles,), torch.int32),
            "counters": BufferSpec((COUNTER_SLOTS,), torch.int64),
            "scratch": BufferSpec((scratch,), torch.uint8),
            "identity_env_xforms": BufferSpec((layout.physical_views, 4, 4), torch.float32),
            "empty_active_ids": BufferSpec((0,), torch.int64),
            "dummy_depth": BufferSpec((1,), torch.float32),
            "dummy_alpha": BufferSpec((1,), torch.float32),
            "dummy_semantic": BufferSpec((1,), torch.int64),
        }
        if o.ray_gaussian_evaluation:
            # Per record: the camera-space precision and its products with the mean.
            buffers["ray_records"] = BufferSpec((v, RAY_RECORD_FLOATS), torch.float32)
        if o.chunked:
            buffers.update({
                "logical_chunk_counters": BufferSpec((layout.chunks, COUNTER_SLOTS), torch.int64),
                "physical_capacity_counters": BufferSpec((COUNTER_SLOTS,), torch.int64),
                "logical_active_mask": BufferSpec((layout.logical_views,), torch.int32),
                "logical_chunk_active_ids": BufferSpec((layout.chunks, layout.physical_views), torch.int64),
            })
        facts = {
            "global_tiles": tiles,
            "sort_temp_bytes": scratch,
            "sort_temp_strategy": "device-count-onesweep-radix",
            "sort_key_begin_bit": 0,
            "sort_key_end_bit": emitted_sort_key_end_bit(max(tiles, 1)),
            "sort_key_bits": key_bits,
            "sort_radix_digit_bits_assumption": digit_bits,
            "sort_radix_passes_expected": math.ceil(key_bits / digit_bits),
        }
        return WorkspacePlan.build(buffers, facts)

    def _price(self, capacity: Capacity) -> int | None:
        try:
            return self._plan(capacity).nbytes
        except ValueError:
            return None

    def _bind(self, arena: WorkspaceArena) -> None:
        """Point the (identity-stable) workspace dict at the arena's views."""
        views = arena.views()
        layout = self._layout
        assert layout is not None
        views["identity_env_xforms"].copy_(torch.eye(4, dtype=torch.float32).expand(layout.physical_views, 4, 4))
        views["counters"].zero_()
        # Diagnostic views into the 48-byte visible records (no extra storage).
        as_int = views["visible_records"].view(torch.int32)
        views["visible_gaussian_ids"] = as_int[:, 10]
        views["visible_camera_ids"] = as_int[:, 11]
        views["visible_depths"] = views["visible_records"][:, 6]
        self.workspace.clear()
        self.workspace.update(views)
        self._arena = arena

    def _install(self, capacity: Capacity, *, reason: str) -> None:
        """Allocate the arena for ``capacity``; on allocator failure restore the old one."""
        layout = self._layout
        assert layout is not None
        try:
            plan = self._plan(capacity)
        except ValueError as error:
            if reason == "initial-allocation":
                raise
            raise RendererCapacityError(str(error), telemetry={
                "schema_version": _ALLOCATION_SCHEMA, "reason": reason, "failure": "native-int32-limit",
                "requested_capacity": capacity.as_dict(), "native_limit": _SORT_CAPACITY_LIMIT}) from error
        current_bytes = self._arena.nbytes if self._arena is not None else 0
        release_first = self._arena is not None and reason != "initial-allocation"
        record: dict[str, Any] = {
            "schema_version": _ALLOCATION_SCHEMA,
            "reason": reason,
            "requested_capacity": {**capacity.as_dict(), "visible_storage_records": capacity.visible_records},
            "requested_workspace_bytes": plan.nbytes,
            "current_workspace_bytes": current_bytes,
            "allocation_strategy": "release-old-workspace-first" if release_first else "transactional-replacement",
            "transient_logical_bytes": max(current_bytes, plan.nbytes) if release_first
            else current_bytes + plan.nbytes,
            "max_workspace_bytes": self.options.max_workspace_bytes,
            "arena_buffers": len(plan.buffers),
            **plan.facts,
        }
        limit = self.options.max_workspace_bytes
        if limit is not None and plan.nbytes > limit:
            record.update(status="failure", failure="workspace-memory-limit")
            self._ledger.last_allocation = copy.deepcopy(record)
            raise RendererCapacityError(
                f"Adaptive capacity growth requires {plan.nbytes} workspace bytes, exceeding the configured limit "
                f"of {limit} bytes.", telemetry=record)
        previous_plan = self._arena.plan if self._arena is not None else None
        if release_first:
            # Drop every reference to the old arena so its memory is reusable by the new one.
            self.workspace.clear()
            self._arena = None
            self.invalidate_projection_cache()
            self._release_cached_memory(layout.device)
        try:
            arena = WorkspaceArena(plan, layout.device)
        except torch.OutOfMemoryError as error:
            record.update(status="failure", failure="allocator-out-of-memory")
            if release_first and previous_plan is not None:
                record["restoration"] = self._restore(previous_plan, layout.device)
            self._ledger.last_allocation = copy.deepcopy(record)
            raise RendererCapacityError(
                f"GPU allocator could not satisfy bounded adaptive workspace growth to {plan.nbytes} bytes.",
                telemetry=record) from error
        self._bind(arena)
        self._capacity = capacity
        self.invalidate_projection_cache()
        record.update(status="success", installed_workspace_bytes=arena.nbytes)
        self._ledger.last_allocation = copy.deepcopy(record)

    def _restore(self, plan: WorkspacePlan, device: torch.device) -> dict[str, Any]:
        self._release_cached_memory(device)
        try:
            arena = WorkspaceArena(plan, device)
        except torch.OutOfMemoryError:
            self._capacity = Capacity(0, 0)
            return {"status": "failed", "workspace_usable": False}
        self._bind(arena)
        return {"status": "restored-old-capacity", "workspace_usable": True, "workspace_bytes": arena.nbytes}

    @staticmethod
    def _release_cached_memory(device: torch.device) -> None:
        if device.type == "cuda":
            with torch.cuda.device(device):
                torch.cuda.empty_cache()

    # --------------------------------------------------------------------- render
    def render(self, request: RenderRequest) -> RenderStats:
        self._require_initialized()
        if self._arena is None:
            raise RuntimeError("allocate_workspace() must be called before render().")
        if self._packed is None:
            raise RuntimeError("At least one Gaussian scene must be registered.")
        names = tuple(request.outputs)
        if names not in (RGB_ONLY_OUTPUTS, FULL_SENSOR_OUTPUTS):
            raise ValueError("CustomHipBackend supports exactly the RGB-only or fused RGB/depth/alpha/semantic_id "
                             f"output contract; got {names}.")
        layout = self._layout
        assert layout is not None
        batch = int(request.camera_transforms.shape[0])
        if batch > layout.logical_views:
            raise ValueError(f"Render batch {batch} exceeds workspace max_views {layout.logical_views}.")
        if tuple(request.outputs["rgb"].shape[1:3]) != (layout.height, layout.width):
            raise ValueError("Output resolution does not match the allocated workspace.")
        full = names == FULL_SENSOR_OUTPUTS
        if batch > layout.physical_views:
            return self._render_chunked(request, full)
        if not self.options.adaptive_capacity:
            stats, _ = self._submit(request, full)
            self._mirror_physical_counters()
            self._ledger.outcome("submitted-unvalidated", "explicit-manual-capacity-mode",
                                 capacity=self._capacity_snapshot(), counters=None, headroom=None, retry_count=0,
                                 attempts=[])
            return stats
        return self._render_validated(request, full)

    def _render_validated(self, request: RenderRequest, full: bool) -> RenderStats:
        """Submit, read the counters once per new input set, grow or re-render until valid."""
        ledger = self._ledger
        ledger.begin()
        retries_allowed = self.options.max_capacity_retries
        retry = 0
        while True:
            stats, fingerprint = self._submit(request, full)
            self._mirror_physical_counters()
            if self._validated is not None and fingerprint.matches(self._validated[0]):
                counters = dict(self._validated[1])
                ledger.validated_input_reuses += 1
                ledger.attempts = []
                ledger.attempt(retry, self._capacity_snapshot(), counters, source="validated-identical-input",
                               synchronized=False)
                ledger.outcome("success", "validated-identical-input", capacity=self._capacity_snapshot(),
                               counters=counters, headroom=headroom_ratios(self._capacity, counters),
                               retry_count=retry)
                return self._stats(stats, counters)
            counters = self.read_counters(synchronize=True)
            attempt = ledger.attempt(retry, self._capacity_snapshot(), counters, source="device-counters",
                                     synchronized=True)
            too_small = overflowed(counters)
            if not too_small and not counters[SORT_FAILURES]:
                self._validated = (fingerprint, dict(counters))
                ledger.device_validations += 1
                ledger.retries += retry
                ledger.outcome("success", "device-counters", capacity=self._capacity_snapshot(), counters=counters,
                               headroom=headroom_ratios(self._capacity, counters), retry_count=retry)
                return self._stats(stats, counters)
            if retry >= retries_allowed:
                ledger.retries += retry
                if too_small:
                    self._fail(request, reason="retry-limit", message=(
                        f"Custom renderer capacity remained insufficient after {retry} bounded retries."))
                self._fail(request, reason="sort-lookback-timeout", message=(
                    f"The intersection sort gave up a look-back wait in each of {retry + 1} attempts; "
                    "the intersection order is invalid."))
            if too_small:
                self._grow(request, counters, attempt, retry)
            retry += 1

    def _grow(self, request: RenderRequest, counters: dict[str, int], attempt: dict[str, Any], retry: int) -> None:
        o = self.options
        policy = CapacityPolicy(o.capacity_growth_factor, o.capacity_headroom, o.max_workspace_bytes)
        full_demand = bool(getattr(self._native, "counts_full_intersection_demand", False))
        plan = policy.propose(self._capacity, counters, full_demand=full_demand, workspace_bytes=self._price)
        attempt["growth_request"] = plan.target.as_dict()
        if plan.ceiling_adjustment is not None:
            attempt["workspace_ceiling_adjustment"] = plan.ceiling_adjustment
        try:
            self._install(plan.target, reason="overflow-retry")
        except RendererCapacityError as error:
            attempt["growth_failure"] = copy.deepcopy(error.telemetry)
            self._ledger.retries += retry
            self._fail(request, reason="growth-allocation-failed", message=str(error), cause=error)
        self._ledger.growth_events += 1

    @staticmethod
    def _stats(submitted: RenderStats, counters: dict[str, int]) -> RenderStats:
        return RenderStats(rendered_views=submitted.rendered_views, visible_gaussians=counters["visible_gaussians"],
                           tile_intersections=counters["tile_intersections"])

    def _fail(self, request: RenderRequest, *, reason: str, message: str,
              cause: BaseException | None = None) -> NoReturn:
        """Poison the outputs (NaN floats, semantic -1) and raise with telemetry."""
        self.invalidate_projection_cache()
        for name, tensor in request.outputs.items():
            tensor.fill_(float("nan") if tensor.is_floating_point() else -1)
        ledger = self._ledger
        last = ledger.attempts[-1] if ledger.attempts else None
        counters = None if last is None else last["actual_demand"]
        ledger.outcome("failure", "device-counters", capacity=self._capacity_snapshot(), counters=counters,
                       headroom=headroom_ratios(self._capacity, counters), retry_count=max(0, len(ledger.attempts) - 1),
                       failure_reason=reason)
        if last is not None:
            asked = last["requested_capacity"]
            message += (f" Last attempt: visible_gaussians={counters['visible_gaussians']} "
                        f"tile_intersections={counters['tile_intersections']} "
                        f"visible_capacity={asked['visible_records']} "
                        f"intersection_capacity={asked['tile_intersections']} "
                        f"visible_overflow={counters['visible_overflow']} "
                        f"intersection_overflow={counters['intersection_overflow']} "
                        f"sort_lookback_failures={counters[SORT_FAILURES]}.")
        error = RendererCapacityError(message, telemetry={"schema_version": _CAPACITY_ERROR_SCHEMA, "reason": reason,
                                                           "capacity": self.capacity_stats})
        if cause is None:
            raise error
        raise error from cause

    # --------------------------------------------------------------- submissions
    def _environment(self, batch: int, start: int = 0) -> torch.Tensor | None:
        if self._env is None:
            return None
        transforms = self._env[1]
        if transforms.shape[0] != batch:
            raise ValueError("update_scene_transforms() must provide one transform per render camera; "
                             f"got {transforms.shape[0]} transforms for batch {batch}.")
        return transforms

    def _output_list(self, outputs: dict[str, torch.Tensor], full: bool, rows: slice | None = None) -> list:
        ws = self.workspace

        def pick(name: str, dummy: str) -> torch.Tensor:
            if not full and name != "rgb":
                return ws[dummy]
            tensor = outputs[name]
            return tensor if rows is None else tensor[rows]

        return [pick("rgb", ""), pick("depth", "dummy_depth"), pick("alpha", "dummy_alpha"),
                pick("semantic_id", "dummy_semantic")]

    @staticmethod
    def _fill_inactive(outputs: dict[str, torch.Tensor], full: bool) -> None:
        """Kernels visit active cameras only: define every inactive camera's outputs first."""
        outputs["rgb"].zero_()
        if full:
            outputs["depth"].fill_(float("inf"))
            outputs["alpha"].zero_()
            outputs["semantic_id"].fill_(-1)

    def _native_render(self, cameras: list[torch.Tensor], outputs: list[torch.Tensor],
                       counters: torch.Tensor | None, *, reuse: bool, full: bool) -> None:
        assert self._native is not None and self._packed is not None and self._layout is not None
        o = self.options
        ws = self.workspace
        workspace = [ws[name] for name in WORKSPACE_TENSORS]
        if counters is not None:
            workspace[WORKSPACE_TENSORS.index("counters")] = counters
        extra: dict[str, Any] = {}
        if self.stage_events is not None:
            extra["stage_events"] = self.stage_events
        if self.statistics is not None:
            extra["statistics"] = self.statistics
        if o.lookback_spin_limit:
            extra["lookback_spin_limit"] = o.lookback_spin_limit
        if o.ray_gaussian_evaluation:
            extra["ray_records"] = ws["ray_records"]
        selector = int(self._native.render(
            self._packed.native_list(), cameras, outputs, workspace, self._layout.height, self._layout.width,
            self._packed.max_scene_gaussians, o.near_plane, o.far_plane, o.gaussian_support_sigma,
            o.covariance_epsilon, o.rasterize_mode == "antialiased", o.semantic_min_alpha,
            o.ray_gaussian_evaluation, o.tile_size,
            o.depth_bucket_count, min(o.depth_bucket_group_size, o.depth_bucket_count), o.compact_projection_cache,
            o.materialize_projected_records, reuse,
            o.output_srgb, o.deterministic, full, o.fixed_capacity_sort, sh_degree=self._sh_degree,
            sort_digit_bits=o.sort_digit_bits, raster_batch=o.raster_batch, cameras_per_block=o.cameras_per_block,
            sort_tile_items=o.sort_tile_items, **extra))
        if selector == 0:
            # The sorted intersections ended in the *_in buffers: rename instead of copying.
            ws["keys_in"], ws["keys_out"] = ws["keys_out"], ws["keys_in"]
            ws["values_in"], ws["values_out"] = ws["values_out"], ws["values_in"]
        elif selector != 1:
            raise RuntimeError(f"Native renderer returned an invalid sorted-buffer selector: {selector}.")

    def _fingerprint(self, request: RenderRequest, env: torch.Tensor, active: torch.Tensor) -> _Fingerprint:
        o = self.options
        packed = self._packed
        assert packed is not None and self._layout is not None
        scalars = (self._scene_revision, int(request.camera_transforms.shape[0]), self._layout.height,
                   self._layout.width, packed.max_scene_gaussians, o.near_plane, o.far_plane,
                   o.gaussian_support_sigma, o.covariance_epsilon, o.rasterize_mode, o.tile_size, o.deterministic,
                   self._sh_degree)
        caller = [*packed.native_list(), request.camera_transforms, request.intrinsics, request.scene_ids]
        owned = []
        for tensor, from_caller in ((env, self._env is not None), (active, request.active_camera_ids is not None)):
            (caller if from_caller else owned).append(tensor)
        return _Fingerprint(scalars, tuple(caller), tuple(owned))

    def _submit(self, request: RenderRequest, full: bool) -> tuple[RenderStats, _Fingerprint]:
        batch = int(request.camera_transforms.shape[0])
        env = self._environment(batch)
        if env is None:
            env = self.workspace["identity_env_xforms"][:batch]
        active = request.active_camera_ids if request.active_camera_ids is not None \
            else self.workspace["empty_active_ids"]
        fingerprint = self._fingerprint(request, env, active)
        reuse = self.options.enable_projection_cache and fingerprint.matches(self._cached_projection)
        if request.active_camera_ids is not None:
            self._fill_inactive(request.outputs, full)
        cameras = [request.camera_transforms, request.intrinsics, env, request.scene_ids, active]
        self._native_render(cameras, self._output_list(request.outputs, full), None, reuse=reuse, full=full)
        if self.options.enable_projection_cache:
            self._cached_projection = fingerprint
            if reuse:
                self._cache_hits += 1
            else:
                self._cache_misses += 1
        views = int(active.numel()) if request.active_camera_ids is not None else batch
        return RenderStats(rendered_views=views), fingerprint

    def _mirror_physical_counters(self) -> None:
        if self.options.chunked:
            self.workspace["physical_capacity_counters"].copy_(self.workspace["counters"])

    def _render_chunked(self, request: RenderRequest, full: bool) -> RenderStats:
        """One logical batch as bounded physical submissions (no host readback)."""
        layout = self._layout
        assert layout is not None and self._native is not None
        o = self.options
        if not o.chunked or not o.fixed_capacity_sort or o.adaptive_capacity or o.enable_projection_cache:
            raise AssertionError("Chunked submission requires preflighted fixed capacities.")
        batch = int(request.camera_transforms.shape[0])
        per = layout.physical_views
        chunks = math.ceil(batch / per)
        if chunks > layout.chunks:
            raise ValueError(f"Logical batch requires {chunks} chunks but only {layout.chunks} were allocated.")
        env = self._environment(batch)
        ws = self.workspace
        active = request.active_camera_ids
        if active is not None:
            self._fill_inactive(request.outputs, full)
            self._native.prepare_chunked_active_ids(active, ws["logical_active_mask"], ws["logical_chunk_active_ids"],
                                                    batch, per)
        for chunk in range(chunks):
            rows = slice(chunk * per, min(batch, (chunk + 1) * per))
            count = rows.stop - rows.start
            cameras = [
                request.camera_transforms[rows], request.intrinsics[rows],
                ws["identity_env_xforms"][:count] if env is None else env[rows],
                request.scene_ids[rows],
                ws["empty_active_ids"] if active is None else ws["logical_chunk_active_ids"][chunk, :count],
            ]
            self._native_render(cameras, self._output_list(request.outputs, full, rows),
                                ws["logical_chunk_counters"][chunk], reuse=False, full=full)
        self._native.aggregate_chunk_counters(ws["logical_chunk_counters"], ws["counters"],
                                              ws["physical_capacity_counters"], chunks)
        return RenderStats(rendered_views=batch if active is None else int(active.numel()))

    # ------------------------------------------------------------------- counters
    def _counter_dict(self, tensor: torch.Tensor, synchronize: bool) -> dict[str, int]:
        if synchronize:
            self.synchronize()
        values = [int(v) for v in tensor.detach().cpu().tolist()]
        counters = dict(zip(COUNTER_NAMES, values))
        counters[SORT_FAILURES] = values[SORT_FAILURE_SLOT] if len(values) > SORT_FAILURE_SLOT else 0
        return counters

    def read_counters(self, *, synchronize: bool = True) -> dict[str, int]:
        """The device counter block (synchronizes; outside steady-state submission)."""
        return self._counter_dict(self.workspace["counters"], synchronize)

    def read_physical_capacity_counters(self, *, synchronize: bool = True) -> dict[str, int]:
        """Largest per-chunk demand of the last chunked render."""
        if "physical_capacity_counters" not in self.workspace:
            return self.read_counters(synchronize=synchronize)
        return self._counter_dict(self.workspace["physical_capacity_counters"], synchronize)

    def check_capacity(self, *, synchronize: bool = True) -> dict[str, int]:
        counters = self.read_counters(synchronize=synchronize)
        if overflowed(counters) or counters[SORT_FAILURES]:
            raise RendererCapacityError(
                "Custom renderer workspace overflow or invalid intersection order: "
                f"visible={counters['visible_overflow']} intersections={counters['intersection_overflow']} "
                f"sort_lookback_failures={counters[SORT_FAILURES]}.",
                telemetry={"schema_version": _CAPACITY_ERROR_SCHEMA, "reason": "explicit-check-overflow",
                           "counters": counters, "capacity": self.capacity_stats})
        return counters

    # ------------------------------------------------------------------ telemetry
    def _capacity_snapshot(self) -> dict[str, int]:
        return {**self._capacity.as_dict(), "visible_storage_records": self.visible_storage_capacity,
                "workspace_bytes": self.workspace_bytes}

    @property
    def workspace_bytes(self) -> int:
        return self._arena.nbytes if self._arena is not None else 0

    @property
    def workspace_bytes_by_tensor(self) -> dict[str, int]:
        """Arena bytes of each named buffer (views of the records add none)."""
        if self._arena is None:
            return {}
        return {name: spec.nbytes for name, spec in self._arena.plan.buffers.items()}

    @property
    def workspace_logical_bytes_by_tensor(self) -> dict[str, int]:
        return {name: t.numel() * t.element_size() for name, t in self.workspace.items()}

    @property
    def workspace_storage_aliases(self) -> dict[str, str]:
        """Views that share another named buffer's bytes (the record diagnostics)."""
        return {"visible_gaussian_ids": "visible_records", "visible_camera_ids": "visible_records",
                "visible_depths": "visible_records"} if self.workspace else {}

    @property
    def current_visible_capacity(self) -> int:
        return self._capacity.visible_records

    @property
    def current_intersection_capacity(self) -> int:
        return self._capacity.tile_intersections

    @property
    def visible_storage_capacity(self) -> int:
        records = self.workspace.get("visible_records")
        return 0 if records is None else int(records.shape[0])

    @property
    def capacity_stats(self) -> dict[str, Any]:
        """JSON-serializable capacity provenance (schema renderer-capacity/v1)."""
        o = self.options
        layout = self._layout
        ledger = self._ledger
        return {
            "schema_version": "renderer-capacity/v1",
            "policy": "adaptive-bounded-grow-retry" if o.adaptive_capacity else "explicit-manual-check",
            "adaptive": o.adaptive_capacity,
            "initial_capacity": self._initial_capacity.as_dict(),
            "current_capacity": self._capacity_snapshot(),
            "execution_schedule": {
                "logical_max_views": layout.logical_views if layout else 0,
                "physical_max_views": layout.physical_views if layout else 0,
                "maximum_logical_chunks": layout.chunks if layout else 1,
                "chunking_enabled": o.chunked,
                "counter_reduction": "device-sum-and-physical-maximum" if o.chunked else "single-native-submission",
            },
            "visible_capacity_role": "materialized-projected-candidate-capacity",
            "projected_record_reuse": {"requested": False, "active": False, "fallback_reason": None,
                                       "bytes_per_record": _RECORD_BYTES},
            "parameters": {
                "rasterize_mode": o.rasterize_mode,
                "covariance_epsilon": o.covariance_epsilon,
                "growth_factor": o.capacity_growth_factor,
                "headroom_factor": o.capacity_headroom,
                "max_retries": o.max_capacity_retries,
                "max_workspace_bytes": o.max_workspace_bytes,
                "workspace_ceiling_policy": "relax-headroom-to-observed-demand",
            },
            "totals": ledger.totals(),
            "synchronization_policy": (
                "new or explicitly invalidated projection inputs read device counters; identical validated "
                "inputs remain asynchronous" if o.adaptive_capacity
                else "caller must explicitly synchronize and check counters"),
            "intersection_ordering_synchronization": (
                "device-count onesweep radix sort: no device-to-host count copy or stream synchronization"),
            "sort_temp_strategy": "device-count-onesweep-radix",
            "last_render": copy.deepcopy(ledger.last_render),
            "last_growth": copy.deepcopy(ledger.last_growth),
            "last_allocation": copy.deepcopy(ledger.last_allocation),
        }

    @property
    def scene_bytes(self) -> int:
        return self._packed.nbytes if self._packed is not None else 0

    @property
    def projection_cache_stats(self) -> dict[str, int | bool]:
        return {"enabled": self.options.enable_projection_cache, "hits": self._cache_hits,
                "misses": self._cache_misses}

    @property
    def pipeline_name(self) -> str:
        return "project-count-scan-write-emit-onesweep-radix-linear-ranges-lds-batched-fused-rgba-depth-semantic"

    @property
    def bin_tile_size(self) -> int:
        """Tile size of the intersection lists: exact-ray evaluation at 16x16 raster tiles
        shares one 32x32 list among four raster tiles (cuda_backend.py:1973-1985)."""
        o = self.options
        return 2 * o.tile_size if o.ray_gaussian_evaluation and o.tile_size == 16 else o.tile_size

    @property
    def rasterize_mode(self) -> Literal["classic", "antialiased"]:
        return self.options.rasterize_mode

    @rasterize_mode.setter
    def rasterize_mode(self, value: Literal["classic", "antialiased"]) -> None:
        updated = replace(self.options, rasterize_mode=value).validate()
        if updated != self.options:
            self.options = updated
            self.invalidate_projection_cache()

    @property
    def resolved_sh_degree(self) -> int:
        return self._sh_degree

    @property
    def max_views(self) -> int:
        return self._layout.logical_views if self._layout else 0

    def invalidate_projection_cache(self) -> None:
        """Forget cached geometry and validated inputs (after writes a version counter cannot see)."""
        self._cached_projection = None
        self._validated = None

    # ---------------------------------------------------------------------- scene
    def _covariances(self, scales: torch.Tensor, rotations: torch.Tensor) -> torch.Tensor:
        out = torch.empty((scales.shape[0], 6), dtype=torch.float32, device=scales.device)
        if scales.device.type == "cuda":
            assert self._native is not None
            self._native.precompute_covariances(scales, rotations, out)
            return out
        # Host tensors exist only in CPU contract tests with a fake native module:
        # Sigma = R diag(s^2) R^T of the normalized quaternion, upper triangle.
        with torch.no_grad():
            q = torch.nn.functional.normalize(rotations, dim=1, eps=1.0e-10)
            w, x, y, z = q.unbind(1)
            rotation = torch.stack([
                torch.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], 1),
                torch.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], 1),
                torch.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], 1),
            ], 1)
            sigma = rotation @ torch.diag_embed(scales.square()) @ rotation.transpose(1, 2)
            rows, cols = torch.triu_indices(3, 3)
            out.copy_(sigma[:, rows, cols])
        return out


# Drop-in name: code written against the reference imports CustomCudaBackend.
CustomCudaBackend = CustomHipBackend
