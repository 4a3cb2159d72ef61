"""Process-local renderer service: the public call a simulator makes.

Public surface of the reference service (``src/isaacsim_gaussian_renderer/
renderer_service.py:15-237``, Apache-2.0): constructor arguments,
``initialize / load_scene / update_scene_transforms / prepare_outputs /
render / synchronize / shutdown``, strict tensor validation, service-owned
output sets initialised to rgb 0 / depth 0 / alpha 0 / semantic -1,
caller-provided outputs, active-camera subsets, and a render cadence that
returns the previous outputs on skipped frames.
"""

from __future__ import annotations

from collections.abc import Iterable
from typing import Any

import torch

from .backend import GaussianRendererBackend, RenderRequest
from .contracts import (
    OUTPUT_SPECS,
    TensorSpec,
    validate_active_camera_ids,
    validate_camera_batch,
    validate_outputs,
)
from .scene import GaussianScene, SceneRegistry

# Value every owned output starts with (semantic -1 = "no label").
_OUTPUT_FILL = {"rgb": 0.0, "depth": 0.0, "alpha": 0.0, "semantic_id": -1}


class _OutputPool:
    """Service-owned output sets, one per batch size, allocated once and reused."""

    def __init__(self, names: tuple[str, ...], height: int, width: int) -> None:
        self.names = names
        self.height = height
        self.width = width
        self._sets: dict[int, dict[str, torch.Tensor]] = {}

    def get(self, batch: int, device: torch.device) -> dict[str, torch.Tensor]:
        existing = self._sets.get(batch)
        if existing is not None and tuple(existing) == self.names:
            return existing
        fresh = {}
        for name in self.names:
            channels, dtype = OUTPUT_SPECS[name]
            fresh[name] = torch.full((batch, self.height, self.width, channels), _OUTPUT_FILL[name], dtype=dtype,
                                     device=device)
        self._sets[batch] = fresh
        return fresh

    def clear(self) -> None:
        self._sets = {}


class _Cadence:
    """Admits the 1st, (n+1)th, (2n+1)th ... render request."""

    def __init__(self, every_n: int) -> None:
        self.every_n = every_n
        self.requests = 0

    def admit(self) -> bool:
        self.requests += 1
        return self.every_n == 1 or self.requests % self.every_n == 1


class RendererService:
    """Stable API between a simulator's camera loop and the renderer kernels."""

    def __init__(
        self,
        backend: GaussianRendererBackend,
        *,
        height: int,
        width: int,
        outputs: Iterable[str] = ("rgb", "depth", "alpha", "semantic_id"),
        max_views: int | None = None,
        render_every_n: int = 1,
        allow_cpu_for_tests: bool = False,
    ) -> None:
        if min(height, width) <= 0:
            raise ValueError(f"height and width must be positive; got {height}x{width}.")
        if render_every_n < 1:
            raise ValueError(f"render_every_n must be positive; got {render_every_n}.")
        names = tuple(outputs)
        unsupported = [name for name in names if name not in OUTPUT_SPECS]
        if unsupported:
            raise ValueError(f"Unsupported output name(s): {', '.join(sorted(unsupported))}.")
        self.backend = backend
        self.height = height
        self.width = width
        self.output_names = names
        self.max_views = max_views
        self.render_every_n = render_every_n
        self.allow_cpu_for_tests = allow_cpu_for_tests
        self.device: torch.device | None = None
        self.stage: Any | None = None
        self.registry = SceneRegistry()
        self.workspace: dict[str, torch.Tensor] = {}
        self._pool = _OutputPool(names, height, width)
        self._cadence = _Cadence(render_every_n)
        self._previous: dict[str, torch.Tensor] | None = None

    # ------------------------------------------------------------------ lifecycle
    @property
    def _require_cuda(self) -> bool:
        return not self.allow_cpu_for_tests

    def _device(self) -> torch.device:
        if self.device is None:
            raise RuntimeError("RendererService.initialize(stage, device) must be called first.")
        return self.device

    def initialize(self, stage: Any, device: torch.device | str) -> None:
        resolved = torch.device(device)
        if resolved.type == "cuda" and resolved.index is None:
            resolved = torch.device("cuda", torch.cuda.current_device())
        if resolved.type != "cuda" and self._require_cuda:
            raise ValueError(f"RendererService requires a CUDA device; got {resolved}.")
        self.backend.initialize(stage, resolved)
        self.stage = stage
        self.device = resolved
        if self.max_views is not None:
            self.workspace = self.backend.allocate_workspace(max_views=self.max_views, height=self.height,
                                                             width=self.width, device=resolved,
                                                             dtype=torch.float32)

    def load_scene(self, scene_id: int, *, means: torch.Tensor, scales: torch.Tensor, rotations: torch.Tensor,
                   opacities: torch.Tensor, features: torch.Tensor, semantic_ids: torch.Tensor) -> None:
        device = self._device()
        scene = GaussianScene(means, scales, rotations, opacities, features, semantic_ids)
        scene.validate(device=device, require_cuda=self._require_cuda)
        offset = self.registry.register(scene_id, scene)
        self.backend.register_scene(scene_id, scene, offset)
        # Build the device offsets now so render() only reads a resident tensor.
        self.registry.offsets_tensor(device=device, require_cuda=self._require_cuda)

    def update_scene_transforms(self, scene_ids: torch.Tensor, transforms: torch.Tensor) -> None:
        device = self._device()
        TensorSpec((None, 4, 4), torch.float32).check(transforms, "transforms", device=device,
                                                       require_cuda=self._require_cuda)
        count = int(transforms.shape[0])
        if tuple(scene_ids.shape) != (count,):
            raise ValueError(f"scene_ids must have one entry per transform ({count}); "
                             f"got shape {tuple(scene_ids.shape)}.")
        TensorSpec((count,), torch.int64).check(scene_ids, "scene_ids", device=device,
                                               require_cuda=self._require_cuda)
        if not self._require_cuda:
            self.registry.require_scene_ids(scene_ids)
        self.backend.update_scene_transforms(scene_ids, transforms)

    # --------------------------------------------------------------------- render
    def _check_batch_size(self, batch: int) -> None:
        if self.max_views is not None and batch > self.max_views:
            raise ValueError(f"Batch has {batch} cameras but service max_views is {self.max_views}.")

    def render(
        self,
        camera_transforms: torch.Tensor,
        intrinsics: torch.Tensor,
        scene_ids: torch.Tensor,
        outputs: dict[str, torch.Tensor] | None = None,
        active_camera_ids: torch.Tensor | None = None,
    ) -> dict[str, torch.Tensor]:
        device = self._device()
        strict = self._require_cuda
        batch = validate_camera_batch(camera_transforms, intrinsics, scene_ids, device=device, require_cuda=strict)
        self._check_batch_size(batch)
        active = validate_active_camera_ids(active_camera_ids, batch=batch, device=device, require_cuda=strict)
        if not strict:
            self.registry.require_scene_ids(scene_ids)
        target = self._pool.get(batch, device) if outputs is None else outputs
        validate_outputs(target, batch=batch, height=self.height, width=self.width, device=device,
                         require_cuda=strict)
        if not self._cadence.admit():
            if self._previous is None:
                raise RuntimeError("Render cadence skipped before any output exists.")
            return self._previous
        self.backend.render(RenderRequest(
            camera_transforms=camera_transforms, intrinsics=intrinsics, scene_ids=scene_ids,
            scene_offsets=self.registry.offsets_tensor(device=device, require_cuda=strict),
            outputs=target, active_camera_ids=active, workspace=self.workspace))
        self._previous = target
        return target

    def prepare_outputs(self, batch: int) -> dict[str, torch.Tensor]:
        """Allocate (once) the service-owned output set of ``batch`` cameras outside the render loop."""
        device = self._device()
        if batch < 1:
            raise ValueError(f"batch must be positive; got {batch}.")
        self._check_batch_size(batch)
        return self._pool.get(batch, device)

    def synchronize(self) -> None:
        self._device()
        self.backend.synchronize()

    def shutdown(self) -> None:
        if self.device is not None:
            self.backend.shutdown()
        self.device = None
        self.workspace = {}
        self._pool.clear()
        self._previous = None
