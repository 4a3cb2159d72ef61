"""Tensor contracts of the renderer-service boundary.

The boundary keeps the reference's output names, channel counts and dtypes
(``src/isaacsim_gaussian_renderer/contracts.py:7-12``, Apache-2.0) and its
error types: a non-tensor is a ``TypeError``, every other violation a
``ValueError`` naming the argument.  The checks are expressed here as
declarative :class:`TensorSpec` objects.  Device-resident active-camera ids are
never read back on the host: reading their range would synchronize the render
path, so the kernels mask out-of-range ids instead.
"""

from __future__ import annotations

from dataclasses import dataclass

import torch

# name -> (channels, dtype); outputs are [B, H, W, C] contiguous device tensors.
OUTPUT_SPECS: dict[str, tuple[int, torch.dtype]] = {
    "rgb": (3, torch.float32),
    "depth": (1, torch.float32),
    "alpha": (1, torch.float32),
    "semantic_id": (1, torch.int64),
}

ANY = None  # wildcard extent in a TensorSpec shape


@dataclass(frozen=True)
class TensorSpec:
    """Rank, extents (``ANY`` = free), dtype and contiguity one argument must have."""

    shape: tuple[int | None, ...]
    dtype: torch.dtype | None

    def check(self, tensor: object, name: str, *, device: torch.device, require_cuda: bool) -> torch.Tensor:
        if not isinstance(tensor, torch.Tensor):
            raise TypeError(f"{name} must be a torch.Tensor.")
        problem = self._placement_problem(tensor, device, require_cuda) or self._layout_problem(tensor)
        if problem is not None:
            raise ValueError(f"{name} {problem}.")
        return tensor

    def _placement_problem(self, tensor: torch.Tensor, device: torch.device, require_cuda: bool) -> str | None:
        if require_cuda and tensor.device.type != "cuda":
            return f"must be CUDA-resident; got device {tensor.device}"
        if tensor.device != device:
            return f"must be on {device}; got {tensor.device}"
        if self.dtype is not None and tensor.dtype != self.dtype:
            return f"must have dtype {self.dtype}; got {tensor.dtype}"
        return None

    def _layout_problem(self, tensor: torch.Tensor) -> str | None:
        if tensor.dim() != len(self.shape):
            return f"must have rank {len(self.shape)}; got shape {tuple(tensor.shape)}"
        mismatched = [(axis, want, got) for axis, (want, got) in enumerate(zip(self.shape, tensor.shape))
                      if want is not ANY and want != got]
        if mismatched:
            axis, want, got = mismatched[0]
            return f"dimension {axis} must be {want}; got {got}"
        if not tensor.is_contiguous():
            return "must be contiguous"
        return None


def ensure_tensor(
    tensor: torch.Tensor,
    *,
    name: str,
    shape: tuple[int | None, ...],
    dtype: torch.dtype | None,
    device: torch.device,
    require_cuda: bool,
) -> torch.Tensor:
    """Validate one tensor against ``shape`` (``None`` = any extent) and ``dtype``."""
    return TensorSpec(tuple(shape), dtype).check(tensor, name, device=device, require_cuda=require_cuda)


def _camera_specs(batch: int) -> tuple[tuple[str, TensorSpec], ...]:
    return (
        ("camera_transforms", TensorSpec((ANY, 4, 4), torch.float32)),
        ("intrinsics", TensorSpec((batch, 3, 3), torch.float32)),
        ("scene_ids", TensorSpec((batch,), torch.int64)),
    )


def validate_camera_batch(
    camera_transforms: torch.Tensor,
    intrinsics: torch.Tensor,
    scene_ids: torch.Tensor,
    *,
    device: torch.device,
    require_cuda: bool,
) -> int:
    """Check the [B,4,4] / [B,3,3] / [B] camera tensors and return B."""
    batch = int(camera_transforms.shape[0]) if isinstance(camera_transforms, torch.Tensor) \
        and camera_transforms.dim() > 0 else 0
    for (name, spec), tensor in zip(_camera_specs(batch), (camera_transforms, intrinsics, scene_ids)):
        spec.check(tensor, name, device=device, require_cuda=require_cuda)
    return batch


def validate_active_camera_ids(
    active_camera_ids: torch.Tensor | None,
    *,
    batch: int,
    device: torch.device,
    require_cuda: bool,
) -> torch.Tensor | None:
    """Check an optional [A] int64 subset of camera indices.

    Only host tensors have their values range-checked (see the module note).
    """
    if active_camera_ids is None:
        return None
    TensorSpec((ANY,), torch.int64).check(active_camera_ids, "active_camera_ids", device=device,
                                          require_cuda=require_cuda)
    if not require_cuda and active_camera_ids.numel() > 0:
        low, high = (int(v) for v in torch.aminmax(active_camera_ids))
        if low < 0 or high >= batch:
            raise ValueError(f"active_camera_ids contains an index outside the camera batch [0, {batch}).")
    return active_camera_ids


def validate_outputs(
    outputs: dict[str, torch.Tensor],
    *,
    batch: int,
    height: int,
    width: int,
    device: torch.device,
    require_cuda: bool,
) -> dict[str, torch.Tensor]:
    """Check a caller-provided output set against ``OUTPUT_SPECS``."""
    if not outputs:
        raise ValueError("At least one output tensor is required.")
    unknown = sorted(name for name in outputs if name not in OUTPUT_SPECS)
    if unknown:
        raise ValueError(f"Unsupported output tensor(s): {', '.join(unknown)}.")
    for name, tensor in outputs.items():
        channels, dtype = OUTPUT_SPECS[name]
        TensorSpec((batch, height, width, channels), dtype).check(
            tensor, f"outputs[{name!r}]", device=device, require_cuda=require_cuda)
    return outputs
