"""Scene registration and packed-offset bookkeeping.

Mirrors ``src/isaacsim_gaussian_renderer/scene.py:13-127``: canonical
activated tensors per scene, registration order defines the packed layout and
``scene_offsets[S+1]``.
"""

from __future__ import annotations

from dataclasses import dataclass

import torch

from .contracts import ensure_tensor


@dataclass(frozen=True)
class GaussianScene:
    """Canonical scene tensors owned by the caller.

    ``scales`` are positive activated scales, ``opacities`` activated
    probabilities, ``rotations`` WXYZ quaternions; ``features[:, :3]`` is the
    linear base RGB and any further channels hold SH bands 1..3 as
    ``features[g, 3k + c]``.
    """

    means: torch.Tensor
    scales: torch.Tensor
    rotations: torch.Tensor
    opacities: torch.Tensor
    features: torch.Tensor
    semantic_ids: torch.Tensor

    @property
    def count(self) -> int:
        return int(self.means.shape[0])

    def validate(self, *, device: torch.device, require_cuda: bool) -> "GaussianScene":
        n = self.count
        ensure_tensor(self.means, name="means", shape=(None, 3), dtype=torch.float32, device=device,
                      require_cuda=require_cuda)
        ensure_tensor(self.scales, name="scales", shape=(n, 3), dtype=torch.float32, device=device,
                      require_cuda=require_cuda)
        ensure_tensor(self.rotations, name="rotations", shape=(n, 4), dtype=torch.float32, device=device,
                      require_cuda=require_cuda)
        ensure_tensor(self.opacities, name="opacities", shape=(n,), dtype=torch.float32, device=device,
                      require_cuda=require_cuda)
        ensure_tensor(self.features, name="features", shape=(n, None), dtype=torch.float32, device=device,
                      require_cuda=require_cuda)
        ensure_tensor(self.semantic_ids, name="semantic_ids", shape=(n,), dtype=torch.int64, device=device,
                      require_cuda=require_cuda)
        return self


class SceneRegistry:
    """Scene ids, packed offsets and total Gaussian count."""

    def __init__(self) -> None:
        self._scenes: dict[int, GaussianScene] = {}
        self._offsets: list[int] = [0]
        self._offset_tensors: dict[torch.device, torch.Tensor] = {}

    @property
    def scene_ids(self) -> tuple[int, ...]:
        return tuple(self._scenes)

    @property
    def total_gaussians(self) -> int:
        return self._offsets[-1]

    def register(self, scene_id: int, scene: GaussianScene) -> int:
        if scene_id in self._scenes:
            raise ValueError(f"Scene {scene_id} is already registered.")
        offset = self._offsets[-1]
        self._scenes[scene_id] = scene
        self._offsets.append(offset + scene.count)
        self._offset_tensors.clear()
        return offset

    def require_scene_ids(self, scene_ids: torch.Tensor) -> None:
        if scene_ids.device.type != "cpu":
            raise ValueError("Host scene-ID validation accepts CPU tensors only.")
        missing = sorted({int(v) for v in scene_ids.tolist()} - set(self._scenes))
        if missing:
            raise ValueError(f"Unregistered scene ID(s): {missing}.")

    def offsets_tensor(self, *, device: torch.device, require_cuda: bool) -> torch.Tensor:
        tensor = self._offset_tensors.get(device)
        if tensor is None:
            tensor = torch.tensor(self._offsets, dtype=torch.int64, device=device)
            self._offset_tensors[device] = tensor
        if require_cuda and tensor.device.type != "cuda":
            raise ValueError("scene_offsets must be CUDA-resident.")
        return tensor
