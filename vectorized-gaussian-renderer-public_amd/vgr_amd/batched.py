"""Batched-persistent rendering: many cameras of uneven resolution over several
scenes in one device-only submission, optionally captured in a hipGraph.

The reference renders one resolution per native submission and answers "many
cameras" with sequential fixed-capacity chunks (``cuda_backend.py:759-967``).
Here one :class:`BatchedRenderer` plan holds every camera's resolution, scene
and output offset on the device, and one submission runs:

  camera table -> frustum cull of (camera group x 256-Gaussian chunk) items
  against conservative per-chunk bounds -> projection + emission over the kept
  items -> onesweep radix ordering -> tile-ID tie order -> tile ranges ->
  persistent compositing over a (camera, tile) work queue.

Nothing in the submission reads a count on the host, so after one eager
calibration render (which sizes the capacities) the whole submission is
captured once with ``torch.cuda.CUDAGraph`` and replayed per frame.

Scenes are stored in Morton order of their means (per scene) so chunk bounds
are tight; records carry registration-order Gaussian IDs, so outputs, tie
order and semantic IDs are those of the plain path.  ``render_plain`` runs the
same batch as one submission per (scene, resolution) group through the plain
path for comparison.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Any

import torch

from .hip_backend import COUNTER_NAMES, RendererCapacityError
from .native import (CAMERA_GRAD_FLOATS, COUNTER_SLOTS, FRONT_MAX_CAMERA_TILES, MAX_INTERSECTIONS,
                     ORDERING_GLOBAL_RADIX, ORDERING_TILE_FRONT, POSE_REPLICAS, RECORD_FLOATS, SORT_FAILURE_SLOT,
                     infer_sh_degree, load_native_module)

ORDERINGS = {"global": ORDERING_GLOBAL_RADIX, "front": ORDERING_TILE_FRONT}


@dataclass(frozen=True)
class CameraSpec:
    """Scene and resolution of one camera of a batched plan."""

    scene_id: int
    width: int
    height: int


def morton_order(means: torch.Tensor) -> torch.Tensor:
    """Permutation sorting points by a 30-bit Morton code over their bounding box."""
    lo = means.min(dim=0).values
    hi = means.max(dim=0).values
    span = torch.clamp(hi - lo, min=1e-12)
    q = ((means - lo) / span * 1023.0).round().clamp(0, 1023).to(torch.int64)

    def spread(v: torch.Tensor) -> torch.Tensor:
        v = (v | (v << 16)) & 0x030000FF
        v = (v | (v << 8)) & 0x0300F00F
        v = (v | (v << 4)) & 0x030C30C3
        v = (v | (v << 2)) & 0x09249249
        return v

    code = spread(q[:, 0]) | (spread(q[:, 1]) << 1) | (spread(q[:, 2]) << 2)
    return torch.argsort(code, stable=True)


class BatchedRenderer:
    """Planned, graph-capturable renderer for a fixed camera layout."""

    def __init__(
        self,
        *,
        device: torch.device | str,
        tile_size: int = 16,
        near_plane: float = 0.01,
        far_plane: float = 100.0,
        gaussian_support_sigma: float = 3.33,
        covariance_epsilon: float = 0.3,
        antialiased: bool = False,
        semantic_min_alpha: float = 0.01,
        output_srgb: bool = False,
        full_sensor_output: bool = True,
        sh_degree: int | None = None,
        spatial_order: bool = True,
        chunk_culling: bool = True,
        persistent_composite: int | None = None,
        sort_digit_bits: int = 0,
        sort_tile_items: int = 0,
        raster_batch: int = 0,
        cameras_per_block: int = 0,
        capacity_headroom: float = 1.15,
        max_visible_per_submission: int = 1 << 28,
        max_intersections_per_submission: int = (1 << 31) - 1,
        ordering: str = "auto",
        front_open_limit: float = 0.05,
        pose_gradients: bool = False,
        native_module: Any | None = None,
    ) -> None:
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise ValueError(f"BatchedRenderer requires a GPU device; got {self.device}.")
        if tile_size not in (8, 16):
            raise ValueError("tile_size must be 8 or 16.")
        self.native = native_module if native_module is not None else load_native_module()
        self.options = dict(
            near_plane=near_plane, far_plane=far_plane, gaussian_support_sigma=gaussian_support_sigma,
            covariance_epsilon=covariance_epsilon, antialiased=antialiased, semantic_min_alpha=semantic_min_alpha,
            output_srgb=output_srgb, tile_size=tile_size)
        self.full_sensor_output = full_sensor_output
        self.requested_sh_degree = sh_degree
        self.spatial_order = spatial_order
        self.chunk_culling = chunk_culling
        if persistent_composite is None:
            persistent_composite = torch.cuda.get_device_properties(self.device).multi_processor_count * 5
        self.persistent_composite = persistent_composite
        self.tuning = dict(sort_digit_bits=sort_digit_bits, sort_tile_items=sort_tile_items,
                           raster_batch=raster_batch, cameras_per_block=cameras_per_block)
        # The plain per-group path keeps the constructed launch parameters: configure()
        # and the autotuner change the batched path only.
        self.plain_tuning = dict(self.tuning)
        self.capacity_headroom = capacity_headroom
        self.max_visible = int(max_visible_per_submission)
        self.max_intersections_limit = int(max_intersections_per_submission)
        if ordering not in ("global", "front", "auto"):
            raise ValueError("ordering must be 'global', 'front' or 'auto'.")
        self.ordering_request = ordering
        self.front_open_limit = float(front_open_limit)
        if pose_gradients and (ordering == "global" or output_srgb or not full_sensor_output):
            raise ValueError("pose_gradients needs the tile-front ordering ('front' or 'auto'), linear RGB and full "
                             "sensor output.")
        self.pose_gradients = bool(pose_gradients)
        self.checkpoints: tuple[torch.Tensor, torch.Tensor] | None = None
        self._last_call = None
        self.ordering = "global"  # resolved by calibrate()
        self.plain_per_submission = 0
        self._scratch_floor = 0
        self.chunk_count = 1
        self._scenes: dict[int, dict[str, torch.Tensor]] = {}
        self._packed: dict[str, torch.Tensor] | None = None
        self.cameras: list[CameraSpec] = []
        self.graph: torch.cuda.CUDAGraph | None = None
        self.workspace: dict[str, torch.Tensor] = {}

    @property
    def max_intersections(self) -> int:
        """Intersection capacity limit of one submission with the current ordering."""
        return min(self.max_intersections_limit, MAX_INTERSECTIONS[ORDERINGS[self.ordering]])

    def front_eligible(self) -> bool:
        """Whether the tile-front ordering can render this plan (16x16 tiles, small cameras)."""
        return self.options["tile_size"] == 16 and bool(self.cameras) and self.max_camera_tiles <= FRONT_MAX_CAMERA_TILES

    def configure(self, *, tile_size: int | None = None, raster_batch: int | None = None,
                  sort_digit_bits: int | None = None, cameras_per_block: int | None = None,
                  sort_tile_items: int | None = None, front_ordering: int | None = None) -> None:
        """Change the batched path's launch parameters; takes effect at the next plan().
        ``front_ordering`` 1 asks for the tile-front ordering where the plan allows it, 0 for the
        global radix sort.  The tile size applies to both paths; the plain per-group path keeps
        the other constructed parameters."""
        if front_ordering is not None:
            self.ordering_request = "front" if int(front_ordering) else "global"
        if tile_size is not None:
            if tile_size not in (8, 16):
                raise ValueError("tile_size must be 8 or 16.")
            self.options["tile_size"] = int(tile_size)
        for name, value in (("raster_batch", raster_batch), ("sort_digit_bits", sort_digit_bits),
                            ("cameras_per_block", cameras_per_block), ("sort_tile_items", sort_tile_items)):
            if value is not None:
                self.tuning[name] = int(value)
        self.graph = None

    # ----------------------------------------------------------------------- scenes
    def register_scene(self, scene_id: int, *, means, scales, rotations, opacities, features, semantic_ids) -> None:
        if scene_id in self._scenes:
            raise ValueError(f"Scene {scene_id} is already registered.")
        tensors = dict(means=means, scales=scales, rotations=rotations, opacities=opacities, features=features,
                       semantic_ids=semantic_ids)
        width = None
        for other in self._scenes.values():
            width = other["features"].shape[1]
        if width is not None and features.shape[1] != width:
            raise ValueError("All packed scenes must use the same feature width.")
        for name, tensor in tensors.items():
            if tensor.device != self.device or not tensor.is_contiguous():
                raise ValueError(f"{name} must be a contiguous tensor on {self.device}.")
        self._scenes[scene_id] = tensors
        self._packed = None
        self.graph = None

    def _pack(self) -> dict[str, torch.Tensor]:
        if self._packed is not None:
            return self._packed
        if not self._scenes:
            raise RuntimeError("register_scene() before planning.")
        order: list[torch.Tensor] = []
        offset = 0
        offsets = [0]
        for tensors in self._scenes.values():
            count = tensors["means"].shape[0]
            local = morton_order(tensors["means"]) if self.spatial_order and count > 1 else torch.arange(
                count, device=self.device)
            order.append(local + offset)
            offset += count
            offsets.append(offset)
        permutation = torch.cat(order)
        plain = {name: torch.cat([t[name] for t in self._scenes.values()]).contiguous()
                 for name in ("means", "scales", "rotations", "opacities", "features", "semantic_ids")}
        packed = {name: plain[name][permutation].contiguous()
                  for name in ("means", "scales", "rotations", "opacities", "features")}
        packed["semantic_ids"] = plain["semantic_ids"]  # registration order: indexed by Gaussian ID
        packed["gaussian_ids"] = permutation.to(torch.int32).contiguous()
        g = packed["means"].shape[0]
        packed["covariances"] = torch.empty((g, 6), device=self.device, dtype=torch.float32)
        self.native.precompute_covariances(packed["scales"], packed["rotations"], packed["covariances"])
        packed["chunk_bounds"] = torch.empty((math.ceil(g / 256), 8), device=self.device, dtype=torch.float32)
        self.native.chunk_bounds(packed["means"], packed["covariances"], packed["chunk_bounds"])
        packed["registered_scene_ids"] = torch.tensor(list(self._scenes), device=self.device, dtype=torch.int64)
        packed["scene_offsets"] = torch.tensor(offsets, device=self.device, dtype=torch.int64)
        plain["covariances"] = torch.empty((g, 6), device=self.device, dtype=torch.float32)
        self.native.precompute_covariances(plain["scales"], plain["rotations"], plain["covariances"])
        packed["plain"] = plain
        width = packed["features"].shape[1]
        self.sh_degree = self.requested_sh_degree if self.requested_sh_degree is not None else infer_sh_degree(width)
        self._packed = packed
        return packed

    # ------------------------------------------------------------------------- plan
    def plan(self, cameras: list[CameraSpec], *, visible_capacity: int | None = None,
             intersection_capacity: int | None = None) -> None:
        """Fix the camera layout and allocate static inputs, outputs and workspace."""
        packed = self._pack()
        if not cameras:
            raise ValueError("plan() needs at least one camera.")
        for spec in cameras:
            if spec.scene_id not in self._scenes:
                raise ValueError(f"Camera refers to unregistered scene {spec.scene_id}.")
            if spec.width <= 0 or spec.height <= 0:
                raise ValueError("camera resolutions must be positive.")
        self.cameras = list(cameras)
        b = len(cameras)
        ts = self.options["tile_size"]
        pixels = [c.width * c.height for c in cameras]
        self.pixel_offsets_host = [0]
        for count in pixels[:-1]:
            self.pixel_offsets_host.append(self.pixel_offsets_host[-1] + count)
        self.total_pixels = sum(pixels)
        self.total_tiles = sum(math.ceil(c.width / ts) * math.ceil(c.height / ts) for c in cameras)
        self.max_camera_tiles = max(math.ceil(c.width / ts) * math.ceil(c.height / ts) for c in cameras)
        self.ordering = "global"
        self.plain_per_submission = 0
        dev = self.device
        self.widths = torch.tensor([c.width for c in cameras], dtype=torch.int32, device=dev)
        self.heights = torch.tensor([c.height for c in cameras], dtype=torch.int32, device=dev)
        self.pixel_offsets = torch.tensor(self.pixel_offsets_host, dtype=torch.int64, device=dev)
        self.scene_ids = torch.tensor([c.scene_id for c in cameras], dtype=torch.int64, device=dev)
        self.viewmats = torch.eye(4, device=dev).repeat(b, 1, 1).contiguous()
        self.intrinsics = torch.eye(3, device=dev).repeat(b, 1, 1).contiguous()
        self.env_xforms = torch.eye(4, device=dev).repeat(b, 1, 1).contiguous()
        self.empty_active = torch.empty(0, dtype=torch.int64, device=dev)
        self.outputs = {
            "rgb": torch.zeros((self.total_pixels, 3), device=dev, dtype=torch.float32),
            "depth": torch.zeros((self.total_pixels, 1), device=dev, dtype=torch.float32),
            "alpha": torch.zeros((self.total_pixels, 1), device=dev, dtype=torch.float32),
            "semantic_id": torch.full((self.total_pixels, 1), -1, device=dev, dtype=torch.int64),
        }
        if self.pose_gradients:
            # Per-pixel checkpoints of the forward and the pose-gradient buffers.
            self.checkpoints = (torch.ones(self.total_pixels, device=dev, dtype=torch.float32),
                                torch.zeros(self.total_pixels, device=dev, dtype=torch.int32))
            self.camera_grads = torch.zeros((POSE_REPLICAS * b, CAMERA_GRAD_FLOATS), device=dev,
                                            dtype=torch.float32)
            self.grad_viewmats = torch.zeros((b, 4, 4), device=dev, dtype=torch.float32)
        g = packed["means"].shape[0]
        if visible_capacity is None:
            visible_capacity = min(b * g, max(1, b * 200_000))
        visible_capacity = min(int(visible_capacity), self.max_visible)
        if intersection_capacity is None:
            intersection_capacity = 3 * visible_capacity
        intersection_capacity = min(int(intersection_capacity), self.max_intersections)
        self._layout_chunks(self.chunk_count)
        self._allocate(visible_capacity, intersection_capacity)
        # Plain-path groups: one submission per (scene, resolution).
        groups: dict[tuple[int, int, int], list[int]] = {}
        for index, spec in enumerate(cameras):
            groups.setdefault((spec.scene_id, spec.width, spec.height), []).append(index)
        self.plain_groups = []
        for (scene_id, width, height), indices in groups.items():
            idx = torch.tensor(indices, dtype=torch.int64, device=dev)
            outputs = {
                "rgb": torch.zeros((len(indices), height, width, 3), device=dev),
                "depth": torch.zeros((len(indices), height, width, 1), device=dev),
                "alpha": torch.zeros((len(indices), height, width, 1), device=dev),
                "semantic_id": torch.full((len(indices), height, width, 1), -1, device=dev, dtype=torch.int64),
            }
            self.plain_groups.append(dict(scene_id=scene_id, width=width, height=height, indices=idx,
                                          index_list=indices, outputs=outputs))
        self.graph = None

    def _layout_chunks(self, count: int) -> None:
        """Split the cameras into ``count`` consecutive submissions of near-equal size."""
        b = len(self.cameras)
        count = max(1, min(int(count), b))
        ts = self.options["tile_size"]
        bounds = [round(i * b / count) for i in range(count + 1)]
        self.chunk_count = count
        self.chunks = []
        for start, stop in zip(bounds[:-1], bounds[1:]):
            tiles = sum(math.ceil(c.width / ts) * math.ceil(c.height / ts) for c in self.cameras[start:stop])
            self.chunks.append((start, stop, tiles))
        self.chunk_counters = torch.zeros((count, COUNTER_SLOTS), dtype=torch.int64, device=self.device)
        self.graph = None

    def _scratch_bytes(self, intersection_capacity: int) -> int:
        packed = self._pack()
        g = packed["means"].shape[0]
        batched = max(self.native.scratch_bytes(
            max_views=stop - start, height=1, width=1, gaussians=g,
            intersection_capacity=min(intersection_capacity, self.max_intersections),
            tile_size=self.options["tile_size"], near_plane=self.options["near_plane"],
            far_plane=self.options["far_plane"], sort_digit_bits=self.tuning["sort_digit_bits"],
            cameras_per_block=self.tuning["cameras_per_block"], sort_tile_items=self.tuning["sort_tile_items"],
            chunk_culling=self.chunk_culling, resolutions=[(c.width, c.height) for c in self.cameras[start:stop]],
            ordering=ORDERINGS[self.ordering])
            for start, stop, _ in self.chunks)
        plain_views = self.plain_per_submission or max(stop - start for start, stop, _ in self.chunks)
        plain = self.native.scratch_bytes(
            max_views=plain_views, height=max(c.height for c in self.cameras),
            width=max(c.width for c in self.cameras), gaussians=g, intersection_capacity=intersection_capacity,
            tile_size=self.options["tile_size"], near_plane=self.options["near_plane"],
            far_plane=self.options["far_plane"], sort_digit_bits=self.plain_tuning["sort_digit_bits"],
            cameras_per_block=self.plain_tuning["cameras_per_block"],
            sort_tile_items=self.plain_tuning["sort_tile_items"])
        # Buffers are shared by both orderings and the plain path: never shrink below
        # what an earlier calibration needed.
        self._scratch_floor = max(self._scratch_floor, batched, plain)
        return self._scratch_floor

    def _allocate(self, visible_capacity: int, intersection_capacity: int) -> None:
        dev = self.device
        self.workspace = {}
        torch.cuda.empty_cache()
        ts = self.options["tile_size"]
        plain_views = self.plain_per_submission or max(stop - start for start, stop, _ in self.chunks)
        plain_tiles = plain_views * \
            math.ceil(max(c.width for c in self.cameras) / ts) * math.ceil(max(c.height for c in self.cameras) / ts)
        tiles = max(max(tiles for _, _, tiles in self.chunks), plain_tiles)
        self.workspace = {
            "visible_records": torch.empty((visible_capacity, RECORD_FLOATS), device=dev, dtype=torch.float32),
            "keys_in": torch.empty(intersection_capacity, device=dev, dtype=torch.uint64),
            "keys_out": torch.empty(intersection_capacity, device=dev, dtype=torch.uint64),
            "values_in": torch.empty(intersection_capacity, device=dev, dtype=torch.int32),
            "values_out": torch.empty(intersection_capacity, device=dev, dtype=torch.int32),
            "tile_starts": torch.empty(tiles, device=dev, dtype=torch.int32),
            "tile_ends": torch.empty(tiles, device=dev, dtype=torch.int32),
            "counters": torch.zeros(COUNTER_SLOTS, device=dev, dtype=torch.int64),
            "scratch": torch.empty(self._scratch_bytes(intersection_capacity), device=dev, dtype=torch.uint8),
        }
        self.visible_capacity = visible_capacity
        self.intersection_capacity = intersection_capacity
        self.graph = None

    # ---------------------------------------------------------------------- submit
    def _workspace_list(self, ordering: str | None = None) -> list[torch.Tensor]:
        """Workspace tensors for one submission; the intersection buffers are viewed at the
        ordering's capacity limit (the buffers are shared and may be larger)."""
        ws = self.workspace
        limit = MAX_INTERSECTIONS[ORDERINGS[ordering or self.ordering]]
        return [ws["visible_records"], ws["keys_in"][:limit], ws["keys_out"][:limit], ws["values_in"][:limit],
                ws["values_out"][:limit], ws["tile_starts"], ws["tile_ends"], ws["counters"], ws["scratch"]]

    def _submit(self, *, stage_events=None, statistics=None) -> None:
        packed = self._pack()
        o = self.outputs
        scene = [packed["means"], packed["covariances"], packed["opacities"], packed["features"],
                 packed["semantic_ids"], packed["registered_scene_ids"], packed["scene_offsets"]]
        cameras = [self.viewmats, self.intrinsics, self.env_xforms, self.scene_ids, self.empty_active]
        outputs = [o["rgb"], o["depth"], o["alpha"], o["semantic_id"]]
        extra: dict[str, Any] = {}
        if stage_events is not None:
            extra["stage_events"] = stage_events
        if self.pose_gradients and self.ordering == "front":
            extra["checkpoints"] = self.checkpoints
        chunk_statistics = None
        if statistics is not None:
            statistics.zero_()
            chunk_statistics = torch.zeros_like(statistics)
            extra["statistics"] = chunk_statistics
        for index, (start, stop, tiles) in enumerate(self.chunks):
            if self.chunk_count > 1:
                cameras = [t[start:stop] for t in (self.viewmats, self.intrinsics, self.env_xforms, self.scene_ids)]
                cameras.append(self.empty_active)
            resolutions = (self.widths[start:stop], self.heights[start:stop], self.pixel_offsets[start:stop], tiles)
            self._selector = self.native.render(
                scene, cameras, outputs, self._workspace_list(), 1, 1, 0, self.options["near_plane"],
                self.options["far_plane"], self.options["gaussian_support_sigma"], self.options["covariance_epsilon"],
                self.options["antialiased"], self.options["semantic_min_alpha"], False, self.options["tile_size"],
                4096, 64, False, False, False, self.options["output_srgb"], True, self.full_sensor_output, False,
                sh_degree=self.sh_degree, persistent_composite=self.persistent_composite,
                gaussian_ids=packed["gaussian_ids"] if self.spatial_order else None,
                chunk_bounds=packed["chunk_bounds"] if self.chunk_culling else None,
                resolutions=resolutions, ordering=ORDERINGS[self.ordering], max_camera_tiles=self.max_camera_tiles,
                **self.tuning, **extra)
            self.chunk_counters[index].copy_(self.workspace["counters"])
            if chunk_statistics is not None:
                statistics.add_(chunk_statistics)
        self._last_call = self.native.last_call

    def read_counters(self) -> dict[str, int]:
        """Counters summed over the submissions of the plan (``max_*`` keys: the largest one)."""
        torch.cuda.synchronize(self.device)
        rows = self.chunk_counters.cpu().tolist()
        totals = dict(zip(COUNTER_NAMES, (int(sum(column)) for column in zip(*rows))))
        totals["sort_lookback_failures"] = int(sum(r[SORT_FAILURE_SLOT] for r in rows))
        totals["max_visible_gaussians"] = max(int(r[0]) for r in rows)
        totals["max_tile_intersections"] = max(int(r[1]) for r in rows)
        totals["submissions"] = len(rows)
        return totals

    def check_capacity(self) -> dict[str, int]:
        counters = self.read_counters()
        if counters["sort_lookback_failures"]:
            raise RendererCapacityError(
                f"Batched renderer sort gave up {counters['sort_lookback_failures']} look-back wait(s); "
                "the intersection order and the images of this render are invalid.",
                telemetry={"schema_version": "renderer-capacity-error/v1", "reason": "sort-lookback-timeout",
                           "counters": counters})
        if counters["visible_overflow"] or counters["intersection_overflow"]:
            raise RendererCapacityError(
                "Batched renderer workspace overflow: "
                f"visible={counters['visible_overflow']} intersections={counters['intersection_overflow']}.",
                telemetry={"schema_version": "renderer-capacity-error/v1", "reason": "batched-overflow",
                           "counters": counters,
                           "capacity": {"visible_records": self.visible_capacity,
                                        "tile_intersections": self.intersection_capacity}})
        return counters

    def set_cameras(self, viewmats: torch.Tensor, intrinsics: torch.Tensor,
                    env_xforms: torch.Tensor | None = None) -> None:
        """Copy a frame's camera tensors into the plan's static inputs."""
        b = len(self.cameras)
        if viewmats.shape != (b, 4, 4) or intrinsics.shape != (b, 3, 3):
            raise ValueError(f"expected viewmats [{b},4,4] and intrinsics [{b},3,3].")
        self.viewmats.copy_(viewmats)
        self.intrinsics.copy_(intrinsics)
        if env_xforms is not None:
            self.env_xforms.copy_(env_xforms)
        else:
            self.env_xforms.copy_(torch.eye(4, device=self.device).expand(b, 4, 4))

    def calibrate(self, viewmats: torch.Tensor, intrinsics: torch.Tensor, env_xforms: torch.Tensor | None = None,
                  max_retries: int = 4) -> dict[str, int]:
        """Eager renders that size the capacities to the demand (x headroom) and resolve the ordering.

        The global radix ordering is calibrated first: its submission size is also the plain
        path's.  With ordering 'front' (or 'auto' on an eligible plan) the tile-front ordering is
        then calibrated on the same buffers; 'auto' keeps it unless more than ``front_open_limit``
        of the non-empty tiles had a pixel open at the end of their sorted front (those tiles are
        sorted in full a second time, which costs more than it saves)."""
        self.set_cameras(viewmats, intrinsics, env_xforms)
        self.ordering = "global"
        counters = self._calibrate_capacities(max_retries)
        self.plain_per_submission = max(stop - start for start, stop, _ in self.chunks)
        if self.pose_gradients and not self.front_eligible():
            raise ValueError("pose gradients need the tile-front ordering, which this plan cannot use "
                             f"(16x16 tiles and at most {FRONT_MAX_CAMERA_TILES} tiles per camera).")
        if self.ordering_request == "global" or not self.front_eligible():
            return counters
        global_chunks = self.chunk_count
        self.ordering = "front"
        self._layout_chunks(1)
        # One submission now spans more tiles (and needs other scratch): size the buffers for it.
        self._allocate(self.visible_capacity, self.intersection_capacity)
        front_counters = self._calibrate_capacities(max_retries)
        if self.ordering_request == "auto" and not self.pose_gradients:
            statistics = torch.zeros(4, dtype=torch.int64, device=self.device)
            self._submit(statistics=statistics)
            opened = int(statistics[3])
            if opened > self.front_open_limit * max(1, front_counters.get("active_tiles", 0)):
                self.ordering = "global"
                self._layout_chunks(global_chunks)
                self._allocate(self.visible_capacity, min(self.intersection_capacity, self.max_intersections))
                return self._calibrate_capacities(max_retries)
        return front_counters

    def _calibrate_capacities(self, max_retries: int) -> dict[str, int]:
        counters: dict[str, int] = {}
        for _ in range(max_retries + 1):
            self._submit()
            counters = self.read_counters()
            # The tile-front ordering stores no visible records (compositing recomputes them).
            records = self.ordering == "global"
            over_limit = ((records and counters["max_visible_gaussians"] > self.max_visible)
                          or counters["max_tile_intersections"] > self.max_intersections)
            overflow = over_limit or counters["visible_overflow"] or counters["intersection_overflow"]
            if not overflow and not counters["sort_lookback_failures"]:
                return counters
            if not overflow:
                continue  # a sort look-back wait gave up: render again with the same buffers
            visible = (math.ceil(counters["max_visible_gaussians"] * self.capacity_headroom) if records
                       else self.visible_capacity)
            intersections = math.ceil(counters["max_tile_intersections"] * self.capacity_headroom)
            if over_limit or visible > self.max_visible or intersections > self.max_intersections:
                # One submission cannot hold a chunk: split the cameras further (the
                # counters count the full demand, so the split is sized in one step).
                if self.chunk_count >= len(self.cameras):
                    raise RendererCapacityError("a single camera exceeds the per-submission capacity limits",
                                                telemetry={"counters": counters})
                factor = max(visible / self.max_visible, intersections / self.max_intersections)
                self._layout_chunks(min(len(self.cameras), math.ceil(self.chunk_count * factor)))
                visible = min(visible, self.max_visible)
                intersections = min(intersections, self.max_intersections)
            self._allocate(min(max(self.visible_capacity, visible), self.max_visible),
                           min(max(self.intersection_capacity, intersections), self.max_intersections))
        raise RendererCapacityError("capacity calibration did not converge", telemetry={"counters": counters})

    def capture(self) -> torch.cuda.CUDAGraph:
        """Capture one whole submission into a graph (after calibrate())."""
        stream = torch.cuda.Stream(device=self.device)
        stream.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(stream):
            self._submit()  # warm-up on the capture stream
        torch.cuda.current_stream(self.device).wait_stream(stream)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=stream):
            self._submit()
        self.graph = graph
        return graph

    def render(self, viewmats: torch.Tensor | None = None, intrinsics: torch.Tensor | None = None,
               env_xforms: torch.Tensor | None = None) -> dict[str, torch.Tensor]:
        """Render the planned batch (graph replay when captured); returns flat [P, C] outputs."""
        if viewmats is not None or intrinsics is not None:
            self.set_cameras(viewmats if viewmats is not None else self.viewmats,
                             intrinsics if intrinsics is not None else self.intrinsics, env_xforms)
        if self.graph is not None:
            self.graph.replay()
        else:
            self._submit()
        return self.outputs

    def backward_pose(self, grad_rgb: torch.Tensor, grad_depth: torch.Tensor | None = None,
                      grad_alpha: torch.Tensor | None = None) -> torch.Tensor:
        """dL/d viewmats [B,4,4] of the last render from the flat upstream gradients
        (``grad_rgb`` [P,3], ``grad_depth`` / ``grad_alpha`` [P,1] or None).

        Needs ``pose_gradients=True`` and the tile-front ordering in one submission: the pass
        walks each tile's list back to front from the checkpointed transmittance, recomputing
        the records, and feeds every (Gaussian, tile) pair's terms through the pair's projection
        backward into per-camera pose partials (no scene gradients).  The returned tensor is the
        renderer's own buffer, overwritten by the next call."""
        if not self.pose_gradients:
            raise RuntimeError("backward_pose needs a renderer built with pose_gradients=True.")
        if self.ordering != "front" or self.chunk_count != 1 or self._last_call is None:
            raise RuntimeError("backward_pose follows a calibrated tile-front render in one submission.")
        for name, tensor, width in (("grad_rgb", grad_rgb, 3), ("grad_depth", grad_depth, 1),
                                    ("grad_alpha", grad_alpha, 1)):
            if tensor is not None and tensor.numel() != self.total_pixels * width:
                raise ValueError(f"{name} must hold {width} value(s) per pixel of the plan ([P, {width}]).")
        self.grad_viewmats.zero_()
        self.native.render_backward(self._last_call, grad_rgb.contiguous(),
                                    None if grad_depth is None else grad_depth.contiguous(),
                                    None if grad_alpha is None else grad_alpha.contiguous(), None,
                                    scene_grads=None, grad_viewmats=self.grad_viewmats,
                                    camera_grads=self.camera_grads)
        return self.grad_viewmats

    def render_pose(self, viewmats: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
        """Differentiable render w.r.t. ``viewmats`` [B,4,4]: returns copies of the flat rgb, depth
        and alpha outputs; their backward calls :meth:`backward_pose`."""
        return _BatchedPoseRender.apply(viewmats, self)

    def camera_outputs(self, index: int, outputs: dict[str, torch.Tensor] | None = None) -> dict[str, torch.Tensor]:
        """[H, W, C] views of camera ``index`` in the flat outputs."""
        outputs = outputs if outputs is not None else self.outputs
        spec = self.cameras[index]
        start = self.pixel_offsets_host[index]
        stop = start + spec.width * spec.height
        return {name: tensor[start:stop].view(spec.height, spec.width, tensor.shape[1])
                for name, tensor in outputs.items()}

    # ------------------------------------------------------------------- plain path
    def render_plain(self) -> list[dict[str, Any]]:
        """Same batch as one plain submission per (scene, resolution) group."""
        packed = self._pack()
        plain = packed["plain"]
        scene = [plain["means"], plain["covariances"], plain["opacities"], plain["features"],
                 plain["semantic_ids"], packed["registered_scene_ids"], packed["scene_offsets"]]
        per_submission = self.plain_per_submission or max(stop - start for start, stop, _ in self.chunks)
        for group in self.plain_groups:
            idx = group["indices"]
            o = group["outputs"]
            for start in range(0, idx.numel(), per_submission):
                part = idx[start:start + per_submission]
                cameras = [self.viewmats.index_select(0, part), self.intrinsics.index_select(0, part),
                           self.env_xforms.index_select(0, part), self.scene_ids.index_select(0, part),
                           self.empty_active]
                stop = start + part.numel()
                outputs = [o["rgb"][start:stop], o["depth"][start:stop], o["alpha"][start:stop],
                           o["semantic_id"][start:stop]]
                self.native.render(
                    scene, cameras, outputs, self._workspace_list("global"),
                    group["height"], group["width"], 0, self.options["near_plane"], self.options["far_plane"],
                    self.options["gaussian_support_sigma"], self.options["covariance_epsilon"],
                    self.options["antialiased"], self.options["semantic_min_alpha"], False, self.options["tile_size"],
                    4096, 64, False, False, False, self.options["output_srgb"], True, self.full_sensor_output, False,
                    sh_degree=self.sh_degree, **self.plain_tuning)
        return self.plain_groups


class _BatchedPoseRender(torch.autograd.Function):
    """Batched forward (graph replay when captured) with camera-pose gradients."""

    @staticmethod
    def forward(ctx, viewmats: torch.Tensor, renderer: BatchedRenderer):
        outputs = renderer.render(viewmats.detach())
        ctx.renderer = renderer
        return outputs["rgb"].clone(), outputs["depth"].clone(), outputs["alpha"].clone()

    @staticmethod
    def backward(ctx, grad_rgb, grad_depth, grad_alpha):
        renderer: BatchedRenderer = ctx.renderer
        if grad_rgb is None:
            grad_rgb = torch.zeros((renderer.total_pixels, 3), device=renderer.device, dtype=torch.float32)
        grad = renderer.backward_pose(grad_rgb, grad_depth, grad_alpha)
        return grad.clone(), None
