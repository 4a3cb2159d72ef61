"""Backend protocol and request/stats records of the renderer service.

Same fields and method set as the reference's injection seam
(``src/isaacsim_gaussian_renderer/backend.py:12-67``).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Protocol

import torch


@dataclass(frozen=True)
class RenderStats:
    """Counters a backend reports for one render request."""

    rendered_views: int
    visible_gaussians: int | None = None
    tile_intersections: int | None = None


@dataclass(frozen=True)
class RenderRequest:
    """Validated tensors handed from the service to a backend."""

    camera_transforms: torch.Tensor
    intrinsics: torch.Tensor
    scene_ids: torch.Tensor
    scene_offsets: torch.Tensor
    outputs: dict[str, torch.Tensor]
    active_camera_ids: torch.Tensor | None = None
    workspace: dict[str, torch.Tensor] = field(default_factory=dict)


class GaussianRendererBackend(Protocol):
    """Kernel backend seam used by ``RendererService``.

    A backend renders into caller-owned output tensors, owns its reusable
    workspace, and never advances simulation state.
    """

    def initialize(self, stage: Any, device: torch.device) -> None:
        """Bind resources to a stage handle and a device."""

    def register_scene(self, scene_id: int, scene: Any, offset: int) -> None:
        """Receive an immutable scene and its packed offset."""

    def allocate_workspace(self, *, max_views: int, height: int, width: int, device: torch.device,
                           dtype: torch.dtype) -> dict[str, torch.Tensor]:
        """Allocate reusable non-output work tensors."""

    def update_scene_transforms(self, scene_ids: torch.Tensor, transforms: torch.Tensor) -> None:
        """Set per-camera scene-to-world transforms."""

    def render(self, request: RenderRequest) -> RenderStats:
        """Render into ``request.outputs``."""

    def synchronize(self) -> None:
        """Wait for submitted work."""

    def shutdown(self) -> None:
        """Release backend resources."""
