"""Differentiable rendering: ``torch.autograd.Function`` over the gfx950 kernels.

The reference renderer is forward-only (``KERNEL_ARCHITECTURE.md:454-460``);
this module adds the training path on top of the same forward:

* forward  = ``vgr_precompute_covariances`` + ``vgr_render`` with per-pixel
  checkpoints (final transmittance, last blended list position);
* backward = ``vgr_render_backward`` (compositing back to front from the
  checkpoint, then projection / EWA / SH / pose per visible pair) and
  ``vgr_covariance_backward`` (scales and quaternions).

Parameters follow the 3DGS convention: means, log-scales, unnormalized WXYZ
quaternions, opacity logits and SH features; ``exp`` and ``sigmoid`` are
applied here in torch so autograd chains them.  Camera poses are the
world->camera ``viewmats`` [B,4,4]; :func:`compose_pose` maps 6-DoF twists
onto base poses for pose optimization (``viewmats = exp(xi^) @ base`` in one
kernel each way, ``vgr_pose_compose``); :func:`se3_exp` is the same map in
torch operations.

Every forward reads the device counters once to verify that the visible and
intersection buffers held the whole batch (and re-renders with grown
capacity when they did not), so gradients never come from a truncated list.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import torch

from . import native as native_module
from .native import (CAMERA_GRAD_FLOATS, COUNTER_SLOTS, POSE_REPLICAS, RECORD_FLOATS, RECORD_GRAD_FLOATS,
                     SORT_FAILURE_SLOT, infer_sh_degree)

_GROWTH = 1.25


def se3_exp(xi: torch.Tensor) -> torch.Tensor:
    """[B,6] twists (rho, phi) -> [B,4,4] rigid transforms exp(xi^) (Rodrigues, differentiable)."""
    if xi.ndim != 2 or xi.shape[1] != 6:
        raise ValueError("se3_exp expects [B,6] twists")
    rho, phi = xi[:, :3], xi[:, 3:]
    theta2 = (phi * phi).sum(-1, keepdim=True)
    small = theta2 < 1e-8
    theta2_safe = torch.where(small, torch.ones_like(theta2), theta2)
    theta = torch.sqrt(theta2_safe)
    a = torch.where(small, 1.0 - theta2 / 6.0, torch.sin(theta) / theta)
    b = torch.where(small, 0.5 - theta2 / 24.0, (1.0 - torch.cos(theta)) / theta2_safe)
    c = torch.where(small, 1.0 / 6.0 - theta2 / 120.0, (theta - torch.sin(theta)) / (theta2_safe * theta))
    zero = torch.zeros_like(phi[:, 0])
    k = torch.stack((zero, -phi[:, 2], phi[:, 1], phi[:, 2], zero, -phi[:, 0], -phi[:, 1], phi[:, 0], zero),
                    dim=-1).reshape(-1, 3, 3)
    eye = torch.eye(3, dtype=xi.dtype, device=xi.device).expand_as(k)
    kk = k @ k
    rotation = eye + a[..., None] * k + b[..., None] * kk
    jacobian = eye + b[..., None] * k + c[..., None] * kk
    translation = (jacobian @ rho.unsqueeze(-1)).squeeze(-1)
    top = torch.cat((rotation, translation.unsqueeze(-1)), dim=-1)
    bottom = torch.zeros((xi.shape[0], 1, 4), dtype=xi.dtype, device=xi.device)
    bottom[:, 0, 3] = 1.0
    return torch.cat((top, bottom), dim=1)


class _PoseCompose(torch.autograd.Function):
    """views = exp(xi^) @ base on the device (forward and analytic backward per camera)."""

    @staticmethod
    def forward(ctx, twists: torch.Tensor, base_views: torch.Tensor, native) -> torch.Tensor:
        twists_c = twists.detach().float().contiguous()
        base_c = base_views.detach().float().contiguous()
        views = torch.empty_like(base_c)
        native.pose_compose(twists_c, base_c, views)
        ctx.save_for_backward(twists_c, base_c)
        ctx.native = native
        return views

    @staticmethod
    def backward(ctx, grad_views: torch.Tensor):
        twists, base = ctx.saved_tensors
        grad_twists = torch.empty_like(twists)
        grad_base = torch.empty_like(base) if ctx.needs_input_grad[1] else None
        ctx.native.pose_compose_backward(twists, base, grad_views.float().contiguous(), grad_twists, grad_base)
        return grad_twists, grad_base, None


def compose_pose(xi: torch.Tensor, base_views: torch.Tensor,
                 native: native_module.NativeRenderer | None = None) -> torch.Tensor:
    """[B,6] twists (rho, phi) and [B,4,4] base poses -> exp(xi^) @ base [B,4,4], differentiable
    in both; the device form of ``se3_exp(xi) @ base_views``."""
    if xi.ndim != 2 or xi.shape[1] != 6:
        raise ValueError("compose_pose expects [B,6] twists")
    if base_views.shape != (xi.shape[0], 4, 4):
        raise ValueError("compose_pose expects [B,4,4] base views matching the twists")
    if xi.dtype != torch.float32 or base_views.dtype != torch.float32:
        raise ValueError("compose_pose expects float32 twists and base views")
    if not xi.is_cuda or base_views.device != xi.device:
        raise ValueError("compose_pose runs on the GPU: twists and base views on one device")
    return _PoseCompose.apply(xi, base_views, native if native is not None else native_module.load_native_module())


@dataclass
class _Workspace:
    records: torch.Tensor
    keys_in: torch.Tensor
    keys_out: torch.Tensor
    values_in: torch.Tensor
    values_out: torch.Tensor
    tile_starts: torch.Tensor
    tile_ends: torch.Tensor
    counters: torch.Tensor
    scratch: torch.Tensor
    slot_records: torch.Tensor  # intersection slot -> visible record (the sort carries slots)

    def as_list(self) -> list[torch.Tensor]:
        return [self.records, self.keys_in, self.keys_out, self.values_in, self.values_out, self.tile_starts,
                self.tile_ends, self.counters, self.scratch]


class DifferentiableRenderer:
    """Training-path renderer: same forward as the service backend, plus gradients.

    ``render(...)`` returns rgb [B,H,W,3], depth [B,H,W,1], alpha [B,H,W,1]
    (differentiable) and semantic_id [B,H,W,1] (int64, not differentiable).
    Linear RGB only (no sRGB transfer on this path).
    """

    def __init__(self, *, height: int, width: int, device: torch.device | str | None = None,
                 near_plane: float = 0.01, far_plane: float = 100.0, gaussian_support_sigma: float = 3.33,
                 covariance_epsilon: float = 0.3, rasterize_mode: str = "classic", semantic_min_alpha: float = 0.01,
                 tile_size: int = 16, raster_batch: int = 0, cameras_per_block: int = 0, sort_digit_bits: int = 0,
                 visible_capacity: int | None = None, intersections_per_visible: float = 4.0,
                 capacity_check: str = "sync", native: native_module.NativeRenderer | None = None) -> None:
        if rasterize_mode not in ("classic", "antialiased"):
            raise ValueError("rasterize_mode must be 'classic' or 'antialiased'.")
        if tile_size not in (8, 16):
            raise ValueError("tile_size must be 8 or 16.")
        if height <= 0 or width <= 0:
            raise ValueError("height and width must be positive.")
        self.height = int(height)
        self.width = int(width)
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.near_plane = float(near_plane)
        self.far_plane = float(far_plane)
        self.gaussian_support_sigma = float(gaussian_support_sigma)
        self.covariance_epsilon = float(covariance_epsilon)
        self.antialiased = rasterize_mode == "antialiased"
        self.semantic_min_alpha = float(semantic_min_alpha)
        self.tile_size = int(tile_size)
        self.raster_batch = int(raster_batch)
        self.cameras_per_block = int(cameras_per_block)
        self.sort_digit_bits = int(sort_digit_bits)
        self.visible_capacity = visible_capacity
        self.intersections_per_visible = float(intersections_per_visible)
        self.intersection_capacity: int | None = None
        self.native = native if native is not None else native_module.load_native_module()
        if capacity_check not in ("sync", "deferred"):
            raise ValueError("capacity_check must be 'sync' or 'deferred'.")
        self.last_counters: dict[str, int] | None = None
        # "sync": every forward reads its counters (one host synchronization) and
        # re-renders with grown buffers on overflow.  "deferred": after the first
        # (synchronous) call, counters are copied to pinned memory without waiting
        # and checked at the next call, which raises RendererCapacityError (after
        # growing the buffers) if an earlier render overflowed; the backward of an
        # overflowed render yields NaN gradients on the device, never silent ones.
        self.capacity_check = capacity_check
        self._calibrated = False
        self._pending: list[tuple[torch.cuda.Event, torch.Tensor, int, int]] = []
        self.max_visible = 1 << 27
        self.max_intersections = (1 << 31) - 1
        self.camera_chunks = 1

    # -- workspace ------------------------------------------------------------------------
    def _tiles(self, batch: int) -> int:
        ts = self.tile_size
        return batch * ((self.width + ts - 1) // ts) * ((self.height + ts - 1) // ts)

    def _workspace(self, batch: int, gaussians: int) -> _Workspace:
        if self.visible_capacity is None:
            # First guess; the counter check after the render grows it to the demand.
            self.visible_capacity = max(1, min(batch * gaussians, 1 << 25))
        if self.intersection_capacity is None:
            self.intersection_capacity = max(1, int(math.ceil(self.visible_capacity * self.intersections_per_visible)))
        visible = min(int(self.visible_capacity), 0x7FFFFFFF)
        intersections = min(int(self.intersection_capacity), (1 << 31) - 1)
        device = self.device
        scratch_bytes = self.native.scratch_bytes(
            max_views=batch, height=self.height, width=self.width, gaussians=gaussians,
            intersection_capacity=intersections, tile_size=self.tile_size, near_plane=self.near_plane,
            far_plane=self.far_plane, sort_digit_bits=self.sort_digit_bits,
            cameras_per_block=self.cameras_per_block or min(64, max(1, batch)))
        tiles = self._tiles(batch)
        return _Workspace(
            records=torch.empty((visible, RECORD_FLOATS), dtype=torch.float32, device=device),
            keys_in=torch.empty(intersections, dtype=torch.uint64, device=device),
            keys_out=torch.empty(intersections, dtype=torch.uint64, device=device),
            values_in=torch.empty(intersections, dtype=torch.int32, device=device),
            values_out=torch.empty(intersections, dtype=torch.int32, device=device),
            tile_starts=torch.empty(tiles, dtype=torch.int32, device=device),
            tile_ends=torch.empty(tiles, dtype=torch.int32, device=device),
            counters=torch.zeros(COUNTER_SLOTS, dtype=torch.int64, device=device),
            scratch=torch.empty(scratch_bytes, dtype=torch.uint8, device=device),
            slot_records=torch.empty(intersections, dtype=torch.int32, device=device))

    # -- forward ---------------------------------------------------------------------------
    def _forward(self, means, scales, quats, opacities, features, viewmats, intrinsics, env_xforms, semantic_ids,
                 sh_degree):
        batch = int(viewmats.shape[0])
        gaussians = int(means.shape[0])
        device = self.device
        # One camera group per projection workgroup (up to 64 cameras) lets the backward
        # own each Gaussian's gradient in one thread (no atomics).
        cameras_per_block = self.cameras_per_block or min(64, max(1, batch))
        covariances = torch.empty((gaussians, 6), dtype=torch.float32, device=device)
        self.native.precompute_covariances(scales, quats, covariances)
        pixels = batch * self.height * self.width
        outputs = {
            "rgb": torch.empty((batch, self.height, self.width, 3), dtype=torch.float32, device=device),
            "depth": torch.empty((batch, self.height, self.width, 1), dtype=torch.float32, device=device),
            "alpha": torch.empty((batch, self.height, self.width, 1), dtype=torch.float32, device=device),
            "semantic_id": torch.empty((batch, self.height, self.width, 1), dtype=torch.int64, device=device),
        }
        transmittance = torch.empty(pixels, dtype=torch.float32, device=device)
        last_blended = torch.empty(pixels, dtype=torch.int32, device=device)
        registered = torch.zeros(1, dtype=torch.int64, device=device)
        offsets = torch.tensor([0, gaussians], dtype=torch.int64, device=device)
        scene_ids = torch.zeros(batch, dtype=torch.int64, device=device)
        active = torch.empty(0, dtype=torch.int64, device=device)
        scene = [means, covariances, opacities, features, semantic_ids, registered, offsets]
        cameras = [viewmats, intrinsics, env_xforms, scene_ids, active]
        for _ in range(4):
            workspace = self._workspace(batch, gaussians)
            self.native.render(
                scene, cameras, [outputs["rgb"], outputs["depth"], outputs["alpha"], outputs["semantic_id"]],
                workspace.as_list(), height=self.height, width=self.width, max_scene_gaussians=gaussians,
                near_plane=self.near_plane, far_plane=self.far_plane,
                gaussian_support_sigma=self.gaussian_support_sigma, covariance_epsilon=self.covariance_epsilon,
                antialiased=self.antialiased, semantic_min_alpha=self.semantic_min_alpha,
                ray_gaussian_evaluation=False, tile_size=self.tile_size, depth_bucket_count=0,
                depth_bucket_group_size=0, compact_projection_cache=False, materialize_projected_records=False,
                reuse_projection=False, output_srgb=False, deterministic=True, full_sensor_output=True,
                fixed_capacity_sort=False, sh_degree=sh_degree, sort_digit_bits=self.sort_digit_bits,
                raster_batch=self.raster_batch, cameras_per_block=cameras_per_block,
                checkpoints=(transmittance, last_blended), slot_records=workspace.slot_records)
            call = self.native.last_call
            if self.capacity_check == "deferred" and self._calibrated:
                pinned = torch.empty(COUNTER_SLOTS, dtype=torch.int64, pin_memory=True)
                pinned.copy_(workspace.counters, non_blocking=True)
                event = torch.cuda.Event()
                event.record()
                self._pending.append((event, pinned, int(workspace.records.shape[0]), int(workspace.keys_in.numel())))
                break
            block = workspace.counters.tolist()
            visible, intersections, sort_failures = int(block[0]), int(block[1]), int(block[SORT_FAILURE_SLOT])
            self.last_counters = {"visible_gaussians": visible, "tile_intersections": intersections,
                                  "sort_lookback_failures": sort_failures}
            if visible > self.max_visible or intersections > self.max_intersections:
                factor = max(visible / self.max_visible, intersections / self.max_intersections)
                raise _SplitCameras(math.ceil(factor * _GROWTH))
            grow = False
            if visible > workspace.records.shape[0]:
                self.visible_capacity = min(int(math.ceil(visible * _GROWTH)), 0x7FFFFFFF)
                self.intersection_capacity = max(int(self.intersection_capacity or 0),
                                                 min(int(math.ceil(intersections * _GROWTH)), (1 << 31) - 1))
                grow = True
            if intersections > workspace.keys_in.numel():
                self.intersection_capacity = min(int(math.ceil(intersections * _GROWTH)), (1 << 31) - 1)
                grow = True
            if sort_failures and not grow:
                continue  # a sort look-back wait gave up: the order is wrong, render again
            if not grow:
                self._calibrated = True
                break
        else:
            from .hip_backend import RendererCapacityError

            raise RendererCapacityError("capacity growth or the intersection sort did not converge",
                                        telemetry={"counters": dict(self.last_counters)})
        state = {"call": call, "workspace": workspace, "covariances": covariances, "transmittance": transmittance,
                 "last_blended": last_blended, "scene": scene, "cameras": cameras, "outputs": outputs}
        return outputs, state

    def check_pending(self, wait: bool = False) -> None:
        """Check counters of earlier deferred renders; raise if one overflowed."""
        keep = []
        overflow = None
        for event, pinned, visible_capacity, intersection_capacity in self._pending:
            if not (wait or event.query()):
                keep.append((event, pinned, visible_capacity, intersection_capacity))
                continue
            event.synchronize()
            visible, intersections = int(pinned[0]), int(pinned[1])
            sort_failures = int(pinned[SORT_FAILURE_SLOT])
            self.last_counters = {"visible_gaussians": visible, "tile_intersections": intersections,
                                  "sort_lookback_failures": sort_failures}
            if visible > visible_capacity or intersections > intersection_capacity or sort_failures:
                overflow = (visible, intersections, visible_capacity, intersection_capacity)
        self._pending = keep
        if overflow is None:
            return
        visible, intersections, visible_capacity, intersection_capacity = overflow
        self._pending = []
        self._calibrated = False  # the next call renders synchronously with grown buffers
        self.visible_capacity = min(max(int(self.visible_capacity or 0), math.ceil(visible * _GROWTH)), 0x7FFFFFFF)
        self.intersection_capacity = min(max(int(self.intersection_capacity or 0),
                                             math.ceil(intersections * _GROWTH)), (1 << 31) - 1)
        from .hip_backend import RendererCapacityError

        raise RendererCapacityError(
            f"an earlier deferred render needed {visible} visible records / {intersections} intersections "
            f"(capacity {visible_capacity} / {intersection_capacity}); its outputs and gradients are invalid "
            "(gradients are NaN). Buffers have been grown: repeat the step.",
            telemetry={"schema_version": "renderer-capacity-error/v1", "reason": "deferred-overflow",
                       "counters": {"visible_gaussians": visible, "tile_intersections": intersections}})

    def render(self, means: torch.Tensor, log_scales: torch.Tensor, quats: torch.Tensor,
               opacity_logits: torch.Tensor, features: torch.Tensor, viewmats: torch.Tensor,
               intrinsics: torch.Tensor, *, semantic_ids: torch.Tensor | None = None, sh_degree: int | None = None,
               env_xforms: torch.Tensor | None = None) -> dict[str, torch.Tensor]:
        for name, tensor in (("means", means), ("log_scales", log_scales), ("quats", quats),
                             ("opacity_logits", opacity_logits), ("features", features), ("viewmats", viewmats),
                             ("intrinsics", intrinsics)):
            if tensor.device != self.device or tensor.dtype != torch.float32:
                raise ValueError(f"{name} must be float32 on {self.device}")
        g = int(means.shape[0])
        if means.shape != (g, 3) or log_scales.shape != (g, 3) or quats.shape != (g, 4) or opacity_logits.shape != (g,):
            raise ValueError("means [G,3], log_scales [G,3], quats [G,4], opacity_logits [G] expected")
        if features.ndim != 2 or features.shape[0] != g:
            raise ValueError("features must be [G,F]")
        batch = int(viewmats.shape[0])
        if viewmats.shape != (batch, 4, 4) or intrinsics.shape != (batch, 3, 3):
            raise ValueError("viewmats [B,4,4] and intrinsics [B,3,3] expected")
        degree = infer_sh_degree(int(features.shape[1])) if sh_degree is None else int(sh_degree)
        if semantic_ids is None:
            semantic_ids = torch.zeros(g, dtype=torch.int64, device=self.device)
        if env_xforms is None:
            env_xforms = torch.eye(4, dtype=torch.float32, device=self.device).expand(batch, 4, 4).contiguous()
        if self._pending:
            self.check_pending(wait=len(self._pending) > 8)
        scales = torch.exp(log_scales)
        opacities = torch.sigmoid(opacity_logits)
        while True:
            chunks = max(1, min(self.camera_chunks, batch))
            bounds = [round(i * batch / chunks) for i in range(chunks + 1)]
            try:
                parts = [_RasterizeFunction.apply(means, scales, quats, opacities, features, viewmats[a:b],
                                                  intrinsics[a:b], env_xforms[a:b], semantic_ids, self, degree)
                         for a, b in zip(bounds[:-1], bounds[1:])]
                break
            except _SplitCameras as split:
                if chunks >= batch:
                    raise RuntimeError("a single camera exceeds the workspace limits") from None
                # Each call holds its workspace until backward; smaller calls bound both.
                self.camera_chunks = min(batch, chunks * max(2, split.factor))
        names = ("rgb", "depth", "alpha", "semantic_id")
        if len(parts) == 1:
            return dict(zip(names, parts[0]))
        return {name: torch.cat([p[i] for p in parts], dim=0) for i, name in enumerate(names)}


class _SplitCameras(Exception):
    """One call's demand exceeds the int32 workspace limits: render fewer cameras per call."""

    def __init__(self, factor: int) -> None:
        super().__init__(f"split cameras by {factor}")
        self.factor = int(factor)


class _RasterizeFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, means, scales, quats, opacities, features, viewmats, intrinsics, env_xforms, semantic_ids,
                renderer: DifferentiableRenderer, sh_degree: int):
        contiguous = [t.contiguous() for t in (means, scales, quats, opacities, features, viewmats, intrinsics,
                                               env_xforms)]
        means_c, scales_c, quats_c, opacities_c, features_c, viewmats_c, intrinsics_c, env_c = contiguous
        outputs, state = renderer._forward(means_c, scales_c, quats_c, opacities_c, features_c, viewmats_c,
                                           intrinsics_c, env_c, semantic_ids.contiguous(), sh_degree)
        ctx.renderer = renderer
        ctx.state = state
        ctx.inputs = (scales_c, quats_c)
        ctx.pose = bool(ctx.needs_input_grad[5])
        ctx.scene_grads = any(ctx.needs_input_grad[:5])
        ctx.mark_non_differentiable(outputs["semantic_id"])
        return outputs["rgb"], outputs["depth"], outputs["alpha"], outputs["semantic_id"]

    @staticmethod
    def backward(ctx, grad_rgb, grad_depth, grad_alpha, grad_semantic):
        renderer: DifferentiableRenderer = ctx.renderer
        state = ctx.state
        device = renderer.device
        outputs = state["outputs"]
        if grad_rgb is None:
            grad_rgb = torch.zeros_like(outputs["rgb"])
        grad_rgb = grad_rgb.contiguous().float()
        # depth is +inf where nothing was blended; the compositing backward reads dL/d depth
        # only where the accumulated alpha is positive, so whatever it holds there is ignored.
        grad_depth = grad_depth.contiguous().float() if grad_depth is not None else None
        grad_alpha = grad_alpha.contiguous().float() if grad_alpha is not None else None
        means, covariances, opacities, features = (state["scene"][0], state["covariances"], state["scene"][2],
                                                   state["scene"][3])
        workspace: _Workspace = state["workspace"]
        # One row of gradient terms per intersection slot (every slot of every tile list
        # is written by the compositing backward, so no zero fill is needed).
        record_grads = torch.empty((workspace.keys_in.numel(), RECORD_GRAD_FLOATS), dtype=torch.float32,
                                   device=device)
        scene_grads = None
        if ctx.scene_grads:
            scene_grads = (torch.zeros_like(means), torch.zeros_like(covariances), torch.zeros_like(opacities),
                           torch.zeros_like(features))
        grad_viewmats = camera_grads = None
        batch = int(state["cameras"][0].shape[0])
        if ctx.pose:
            grad_viewmats = torch.zeros((batch, 4, 4), dtype=torch.float32, device=device)
            camera_grads = torch.empty((POSE_REPLICAS * batch, CAMERA_GRAD_FLOATS), dtype=torch.float32,
                                       device=device)
        renderer.native.render_backward(state["call"], grad_rgb, grad_depth, grad_alpha, record_grads,
                                        scene_grads=scene_grads, grad_viewmats=grad_viewmats,
                                        camera_grads=camera_grads)
        grad_means = grad_scales = grad_quats = grad_opacities = grad_features = None
        if scene_grads is not None:
            grad_means, grad_covariances, grad_opacities, grad_features = scene_grads
            scales, quats = ctx.inputs
            grad_scales = torch.empty_like(scales)
            grad_quats = torch.empty_like(quats)
            renderer.native.covariance_backward(scales, quats, grad_covariances, grad_scales, grad_quats)
        ctx.state = None
        return (grad_means, grad_scales, grad_quats, grad_opacities, grad_features, grad_viewmats, None, None, None,
                None, None)
