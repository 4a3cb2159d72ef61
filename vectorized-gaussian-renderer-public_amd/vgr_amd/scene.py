"""Scenes and their packed, multi-scene device layout.

A :class:`GaussianScene` is the canonical activated tensor set of one scene
(the reference's ``GaussianScene`` fields, ``src/isaacsim_gaussian_renderer/
scene.py:13-84``).  Scenes are packed in registration order into one SoA arrays
set; ``scene_offsets[S+1]`` and ``registered_scene_ids[S]`` (int64) locate each
scene inside it, as the native renderer expects (``renderer_cuda.cu:48-56``).

:class:`SceneRegistry` keeps the id -> scene table and the host offsets;
:func:`pack_scenes` builds the device arrays once per registration so a render
only reads resident tensors.
"""

from __future__ import annotations

from collections.abc import Callable
from dataclasses import dataclass, fields

import torch

from .contracts import ANY, TensorSpec

_SCENE_FIELDS = ("means", "scales", "rotations", "opacities", "features", "semantic_ids")


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

    def specs(self) -> dict[str, TensorSpec]:
        """Per-field contract; every field has one row per Gaussian of ``means``."""
        n = self.count
        return {
            "means": TensorSpec((ANY, 3), torch.float32),
            "scales": TensorSpec((n, 3), torch.float32),
            "rotations": TensorSpec((n, 4), torch.float32),
            "opacities": TensorSpec((n,), torch.float32),
            "features": TensorSpec((n, ANY), torch.float32),
            "semantic_ids": TensorSpec((n,), torch.int64),
        }

    def validate(self, *, device: torch.device, require_cuda: bool) -> "GaussianScene":
        for name, spec in self.specs().items():
            spec.check(getattr(self, name), name, device=device, require_cuda=require_cuda)
        return self

    def tensors(self) -> dict[str, torch.Tensor]:
        return {f.name: getattr(self, f.name) for f in fields(self)}


class SceneRegistry:
    """Registered scenes in packing order, with their host-side packed offsets."""

    def __init__(self) -> None:
        self._scenes: dict[int, GaussianScene] = {}
        self._device_offsets: dict[torch.device, torch.Tensor] = {}

    @property
    def scene_ids(self) -> tuple[int, ...]:
        return tuple(self._scenes)

    @property
    def scenes(self) -> tuple[GaussianScene, ...]:
        return tuple(self._scenes.values())

    @property
    def offsets(self) -> list[int]:
        """Packed start of every scene, then the total (``scene_offsets``)."""
        running = [0]
        for scene in self._scenes.values():
            running.append(running[-1] + scene.count)
        return running

    @property
    def total_gaussians(self) -> int:
        return self.offsets[-1]

    def __len__(self) -> int:
        return len(self._scenes)

    def __contains__(self, scene_id: object) -> bool:
        return scene_id in self._scenes

    def register(self, scene_id: int, scene: GaussianScene) -> int:
        """Append a scene; returns its packed offset."""
        if scene_id in self._scenes:
            raise ValueError(f"Scene {scene_id} is already registered.")
        start = self.total_gaussians
        self._scenes[scene_id] = scene
        self._device_offsets = {}
        return start

    def require_scene_ids(self, scene_ids: torch.Tensor) -> None:
        """Host-side check that every camera names a registered scene (CPU tensors only)."""
        if scene_ids.device.type != "cpu":
            raise ValueError("Host scene-ID validation accepts CPU tensors only.")
        unknown = sorted(set(int(v) for v in scene_ids.unique().tolist()).difference(self._scenes))
        if unknown:
            raise ValueError(f"Unregistered scene ID(s): {unknown}.")

    def offsets_tensor(self, *, device: torch.device, require_cuda: bool) -> torch.Tensor:
        """``scene_offsets`` resident on ``device`` (built once per registration)."""
        if require_cuda and device.type != "cuda":
            raise ValueError("scene_offsets must be CUDA-resident.")
        if device not in self._device_offsets:
            self._device_offsets[device] = torch.tensor(self.offsets, dtype=torch.int64, device=device)
        return self._device_offsets[device]


@dataclass
class PackedScene:
    """Device arrays of every registered scene, concatenated in registration order."""

    means: torch.Tensor
    covariances: torch.Tensor
    opacities: torch.Tensor
    features: torch.Tensor
    semantic_ids: torch.Tensor
    registered_scene_ids: torch.Tensor
    scene_offsets: torch.Tensor
    scales: torch.Tensor
    rotations: torch.Tensor
    max_scene_gaussians: int

    def native_list(self) -> list[torch.Tensor]:
        """The seven scene tensors of the native render call (``renderer_cuda.cu:48-56``)."""
        return [self.means, self.covariances, self.opacities, self.features, self.semantic_ids,
                self.registered_scene_ids, self.scene_offsets]

    @property
    def feature_width(self) -> int:
        return int(self.features.shape[1])

    @property
    def nbytes(self) -> int:
        return sum(t.numel() * t.element_size() for t in (self.means, self.covariances, self.opacities,
                                                          self.features, self.semantic_ids,
                                                          self.registered_scene_ids, self.scene_offsets))


def pack_scenes(registry: SceneRegistry, *, device: torch.device,
                covariances: Callable[[torch.Tensor, torch.Tensor], torch.Tensor]) -> PackedScene:
    """Concatenate the registered scenes; ``covariances(scales, rotations)`` builds [G,6]."""
    scenes = registry.scenes
    if not scenes:
        raise RuntimeError("At least one Gaussian scene must be registered.")

    def joined(name: str) -> torch.Tensor:
        parts = [getattr(scene, name) for scene in scenes]
        return parts[0] if len(parts) == 1 else torch.cat(parts)

    arrays = {name: joined(name) for name in _SCENE_FIELDS}
    return PackedScene(
        means=arrays["means"], covariances=covariances(arrays["scales"], arrays["rotations"]),
        opacities=arrays["opacities"], features=arrays["features"], semantic_ids=arrays["semantic_ids"],
        registered_scene_ids=torch.tensor(registry.scene_ids, dtype=torch.int64, device=device),
        scene_offsets=torch.tensor(registry.offsets, dtype=torch.int64, device=device),
        scales=arrays["scales"], rotations=arrays["rotations"],
        max_scene_gaussians=max(scene.count for scene in scenes))
