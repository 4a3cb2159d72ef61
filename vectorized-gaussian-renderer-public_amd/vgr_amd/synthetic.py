"""Seeded synthetic scenes and camera rigs for the benchmark workloads.

Tensors are generated on the host with ``torch.Generator("cpu")`` so the CPU
oracle and the GPU see byte-identical inputs (the reference's manifest
pattern, ``benchmark_manifest.py:492-522``).  Distributions:
means ~ N(0, 3^2) per axis, log-scales ~ N(-4, 0.5^2), unit WXYZ quaternions
from normalized N(0, 1) with w >= 0, opacity = sigmoid(N(0, 1.5^2)), base RGB
~ U[0, 1], SH bands 1..3 ~ N(0, 0.1^2), semantic ids uniform over the class
count.  Cameras are OpenCV world->camera look-at rigs.
"""

from __future__ import annotations

import hashlib
import math
from dataclasses import dataclass

import torch


@dataclass
class SyntheticScene:
    means: torch.Tensor
    scales: torch.Tensor
    rotations: torch.Tensor
    opacities: torch.Tensor
    features: torch.Tensor
    semantic_ids: torch.Tensor

    @property
    def count(self) -> int:
        return int(self.means.shape[0])

    def to(self, device: torch.device | str) -> "SyntheticScene":
        return SyntheticScene(**{k: getattr(self, k).to(device).contiguous() for k in self.__dataclass_fields__})

    def as_kwargs(self) -> dict[str, torch.Tensor]:
        return {k: getattr(self, k) for k in self.__dataclass_fields__}


@dataclass
class CameraRig:
    viewmats: torch.Tensor      # [B,4,4]
    intrinsics: torch.Tensor    # [B,3,3]
    widths: list[int]
    heights: list[int]

    @property
    def batch(self) -> int:
        return int(self.viewmats.shape[0])


def gaussian_scene(count: int, *, seed: int, sh_degree: int = 0, semantic_classes: int = 64,
                   layout: str = "normal") -> SyntheticScene:
    """Scene of ``count`` Gaussians; ``layout`` 'normal' (N(0,3^2) means) or 'room'."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    if layout == "normal":
        means = 3.0 * torch.randn((count, 3), generator=g)
    elif layout == "room":
        means = room_means(count, g)
    else:
        raise ValueError(f"unknown layout {layout!r}")
    scales = torch.exp(-4.0 + 0.5 * torch.randn((count, 3), generator=g))
    quats = torch.nn.functional.normalize(torch.randn((count, 4), generator=g), dim=-1)
    quats = torch.where(quats[:, :1] < 0, -quats, quats)
    opacities = torch.sigmoid(1.5 * torch.randn((count,), generator=g))
    bands = (sh_degree + 1) ** 2
    features = torch.empty((count, 3 * bands), dtype=torch.float32)
    features[:, :3] = torch.rand((count, 3), generator=g)
    if bands > 1:
        features[:, 3:] = 0.1 * torch.randn((count, 3 * (bands - 1)), generator=g)
    semantic = torch.randint(0, semantic_classes, (count,), generator=g, dtype=torch.int64)
    return SyntheticScene(means.float().contiguous(), scales.float().contiguous(), quats.float().contiguous(),
                          opacities.float().contiguous(), features.contiguous(), semantic.contiguous())


def room_means(count: int, generator: torch.Generator) -> torch.Tensor:
    """70 % uniform in [-10,10]^2 x [0,4]; 30 % on the four walls with N(0, 0.02^2) normal jitter."""
    interior = int(round(0.7 * count))
    walls = count - interior
    means = torch.empty((count, 3), dtype=torch.float32)
    u = torch.rand((interior, 3), generator=generator)
    means[:interior, 0] = -10.0 + 20.0 * u[:, 0]
    means[:interior, 1] = -10.0 + 20.0 * u[:, 1]
    means[:interior, 2] = 4.0 * u[:, 2]
    wall_index = torch.randint(0, 4, (walls,), generator=generator)
    along = -10.0 + 20.0 * torch.rand((walls,), generator=generator)
    height = 4.0 * torch.rand((walls,), generator=generator)
    jitter = 0.02 * torch.randn((walls,), generator=generator)
    sign = torch.where(wall_index % 2 == 0, 1.0, -1.0)
    on_x = wall_index < 2
    wx = torch.where(on_x, sign * 10.0 + jitter, along)
    wy = torch.where(on_x, along, sign * 10.0 + jitter)
    means[interior:, 0] = wx
    means[interior:, 1] = wy
    means[interior:, 2] = height
    return means


def look_at(centers: torch.Tensor, targets: torch.Tensor) -> torch.Tensor:
    """OpenCV world->camera matrices (x right, y down, z forward); up hint +Z, +Y near the poles."""
    centers = centers.double()
    forward = torch.nn.functional.normalize(targets.double() - centers, dim=-1)
    up = torch.zeros_like(forward)
    up[:, 2] = 1.0
    polar = forward[:, 2].abs() > 0.99
    up[polar] = torch.tensor([0.0, 1.0, 0.0], dtype=torch.float64)
    right = torch.nn.functional.normalize(torch.linalg.cross(forward, up, dim=-1), dim=-1)
    down = torch.linalg.cross(forward, right, dim=-1)
    rotation = torch.stack((right, down, forward), dim=1)
    view = torch.zeros((centers.shape[0], 4, 4), dtype=torch.float64)
    view[:, :3, :3] = rotation
    view[:, :3, 3] = -(rotation @ centers.unsqueeze(-1)).squeeze(-1)
    view[:, 3, 3] = 1.0
    return view.float().contiguous()


def pinhole(batch: int, width: int, height: int, focal: float) -> torch.Tensor:
    k = torch.zeros((batch, 3, 3), dtype=torch.float32)
    k[:, 0, 0] = focal
    k[:, 1, 1] = focal
    k[:, 0, 2] = width / 2.0
    k[:, 1, 2] = height / 2.0
    k[:, 2, 2] = 1.0
    return k


def fibonacci_directions(n: int) -> torch.Tensor:
    i = torch.arange(n, dtype=torch.float64) + 0.5
    phi = torch.acos(1.0 - 2.0 * i / n)
    theta = math.pi * (1.0 + 5.0 ** 0.5) * i
    return torch.stack((torch.cos(theta) * torch.sin(phi), torch.sin(theta) * torch.sin(phi), torch.cos(phi)), dim=1)


def orbit_rig(batch: int, width: int, height: int, focal: float, *, seed: int, radius: tuple[float, float]) -> CameraRig:
    """Cameras on Fibonacci-sphere directions at radius U[r0, r1], looking at the origin."""
    g = torch.Generator(device="cpu").manual_seed(seed + 7919)
    directions = fibonacci_directions(batch)
    radii = radius[0] + (radius[1] - radius[0]) * torch.rand((batch, 1), generator=g, dtype=torch.float64)
    centers = directions * radii
    viewmats = look_at(centers, torch.zeros_like(centers))
    return CameraRig(viewmats, pinhole(batch, width, height, focal), [width] * batch, [height] * batch)


def single_camera(width: int, height: int, focal: float, position=(12.0, 0.0, 0.0)) -> CameraRig:
    center = torch.tensor([position], dtype=torch.float64)
    return CameraRig(look_at(center, torch.zeros_like(center)), pinhole(1, width, height, focal), [width], [height])


def catmull_rom(points: torch.Tensor, samples: int) -> tuple[torch.Tensor, torch.Tensor]:
    """Positions and unit tangents of a uniform Catmull-Rom spline through ``points`` [K,3]."""
    k = points.shape[0]
    padded = torch.cat((2 * points[:1] - points[1:2], points, 2 * points[-1:] - points[-2:-1]), dim=0)
    t_global = torch.linspace(0.0, k - 1 - 1e-6, samples, dtype=torch.float64)
    seg = t_global.floor().long()
    t = (t_global - seg).unsqueeze(1)
    p0, p1, p2, p3 = padded[seg], padded[seg + 1], padded[seg + 2], padded[seg + 3]
    pos = 0.5 * ((2 * p1) + (-p0 + p2) * t + (2 * p0 - 5 * p1 + 4 * p2 - p3) * t ** 2
                 + (-p0 + 3 * p1 - 3 * p2 + p3) * t ** 3)
    tan = 0.5 * ((-p0 + p2) + 2 * (2 * p0 - 5 * p1 + 4 * p2 - p3) * t + 3 * (-p0 + 3 * p1 - 3 * p2 + p3) * t ** 2)
    return pos, torch.nn.functional.normalize(tan, dim=-1)


def trajectory_rig(trajectories: int, samples: int, width: int, height: int, focal: float, *, seed: int) -> CameraRig:
    """Robotics rig: Catmull-Rom trajectories through 6 waypoints in [-9,9]^2 x [0.3,1.8], heading = tangent."""
    g = torch.Generator(device="cpu").manual_seed(seed + 104729)
    centers, targets = [], []
    for _ in range(trajectories):
        u = torch.rand((6, 3), generator=g, dtype=torch.float64)
        way = torch.stack((-9.0 + 18.0 * u[:, 0], -9.0 + 18.0 * u[:, 1], 0.3 + 1.5 * u[:, 2]), dim=1)
        pos, tan = catmull_rom(way, samples)
        centers.append(pos)
        targets.append(pos + tan)
    centers_t = torch.cat(centers)
    viewmats = look_at(centers_t, torch.cat(targets))
    batch = centers_t.shape[0]
    return CameraRig(viewmats, pinhole(batch, width, height, focal), [width] * batch, [height] * batch)


def tensor_sha256(*tensors: torch.Tensor) -> str:
    digest = hashlib.sha256()
    for tensor in tensors:
        digest.update(tensor.detach().cpu().contiguous().numpy().tobytes())
    return digest.hexdigest()


# ---------------------------------------------------------------- benchmark workloads
# The five workloads of the benchmark, built from the generators above with fixed
# seeds so the benchmark, the parity tests and the CPU baseline share their bytes.

ORBIT_SEED, UNEVEN_SEED, TRAINING_SEED, ROBOTICS_SEED = 20261202, 20261203, 20261204, 20261205
UNEVEN_SIZES = (200_000, 500_000, 1_000_000, 2_000_000)
UNEVEN_RESOLUTIONS = ((128, 128), (320, 240), (640, 480), (1280, 720))


@dataclass
class UnevenBatch:
    """Cameras of several resolutions over several scenes (scene id = camera index mod scenes)."""

    scenes: list[tuple[int, SyntheticScene]]
    scene_ids: list[int]
    resolutions: list[tuple[int, int]]
    viewmats: torch.Tensor
    intrinsics: torch.Tensor


def single_camera_workload() -> tuple[SyntheticScene, CameraRig]:
    """1 camera 640x480, fx = fy = 554, 100k Gaussians SH0 (the CPU-runnable workload)."""
    return gaussian_scene(100_000, seed=20261201, sh_degree=0), single_camera(640, 480, 554.0)


def orbit_workload(cameras: int = 64, gaussians: int = 1_000_000) -> tuple[SyntheticScene, CameraRig]:
    """64 cameras 640x480 (fx = fy = 554) at radius 10-15 around 1M Gaussians SH3."""
    return (gaussian_scene(gaussians, seed=ORBIT_SEED, sh_degree=3),
            orbit_rig(cameras, 640, 480, 554.0, seed=ORBIT_SEED, radius=(10.0, 15.0)))


def uneven_workload(cameras: int = 256, sizes: tuple[int, ...] = UNEVEN_SIZES) -> UnevenBatch:
    """256 cameras, resolutions U{128^2, 320x240, 640x480, 1280x720} with a 60 degree horizontal
    FOV, at radius 10-15 around 4 scenes of 200k/500k/1M/2M Gaussians SH3, round-robin."""
    g = torch.Generator(device="cpu").manual_seed(UNEVEN_SEED)
    choice = torch.randint(0, len(UNEVEN_RESOLUTIONS), (cameras,), generator=g).tolist()
    scenes = [(k, gaussian_scene(count, seed=UNEVEN_SEED + k, sh_degree=3)) for k, count in enumerate(sizes)]
    rig = orbit_rig(cameras, 1, 1, 1.0, seed=UNEVEN_SEED, radius=(10.0, 15.0))
    resolutions = [UNEVEN_RESOLUTIONS[c] for c in choice]
    intrinsics = torch.stack([pinhole(1, w, h, (w / 2) / math.tan(math.radians(30)))[0] for w, h in resolutions])
    return UnevenBatch(scenes, [i % len(sizes) for i in range(cameras)], resolutions, rig.viewmats,
                       intrinsics.contiguous())


def robotics_workload(trajectories: int = 32, samples: int = 32,
                      gaussians: int = 5_000_000) -> tuple[SyntheticScene, CameraRig]:
    """1024 cameras 160x120 (fx = fy = 139) along 32 Catmull-Rom trajectories in a 5M-Gaussian
    SH1 room with 256 semantic classes."""
    scene = gaussian_scene(gaussians, seed=ROBOTICS_SEED, sh_degree=1, semantic_classes=256, layout="room")
    return scene, trajectory_rig(trajectories, samples, 160, 120, 139.0, seed=ROBOTICS_SEED)
