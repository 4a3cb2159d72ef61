"""MI355X (gfx950) batched Gaussian-splat camera renderer.

Drop-in host layer for the reference renderer service and its native CUDA
backend: ``RendererService`` + ``CustomHipBackend`` (alias
``CustomCudaBackend``) over hand-written HIP kernels in ``libvgr_hip.so``.
"""

from .backend import GaussianRendererBackend, RenderRequest, RenderStats
from .contracts import OUTPUT_SPECS
from .hip_backend import (
    COUNTER_NAMES,
    FULL_SENSOR_OUTPUTS,
    RGB_ONLY_OUTPUTS,
    CustomCudaBackend,
    CustomHipBackend,
    RendererCapacityError,
)
from .ply import canonicalize_3dgs_scene, load_ply_to_canonical_gaussians, load_ply_to_gaussians
from .renderer_service import RendererService
from .scene import GaussianScene, SceneRegistry

__all__ = [
    "COUNTER_NAMES",
    "FULL_SENSOR_OUTPUTS",
    "OUTPUT_SPECS",
    "RGB_ONLY_OUTPUTS",
    "CustomCudaBackend",
    "CustomHipBackend",
    "GaussianRendererBackend",
    "GaussianScene",
    "RenderRequest",
    "RenderStats",
    "RendererCapacityError",
    "RendererService",
    "SceneRegistry",
    "canonicalize_3dgs_scene",
    "load_ply_to_canonical_gaussians",
    "load_ply_to_gaussians",
]
