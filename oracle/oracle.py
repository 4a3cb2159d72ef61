"""ctypes wrapper of the CPU oracle (``liboracle.so``).

TEST INFRASTRUCTURE ONLY: imported by the test suite, the smoke entry point and
the benchmark's CPU-baseline leg as the checker.  The renderer product never
imports this module.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
LIBRARY_PATH = ORACLE_DIR / "liboracle.so"


class _Scene(ctypes.Structure):
    _fields_ = [("means", ctypes.c_void_p), ("covariances", ctypes.c_void_p), ("opacities", ctypes.c_void_p),
                ("features", ctypes.c_void_p), ("semantic_ids", ctypes.c_void_p),
                ("registered_scene_ids", ctypes.c_void_p), ("scene_offsets", ctypes.c_void_p),
                ("gaussian_count", ctypes.c_int64), ("feature_width", ctypes.c_int32),
                ("scene_count", ctypes.c_int32), ("sh_degree", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class _Cameras(ctypes.Structure):
    _fields_ = [("viewmats", ctypes.c_void_p), ("intrinsics", ctypes.c_void_p), ("env_xforms", ctypes.c_void_p),
                ("scene_ids", ctypes.c_void_p), ("active_ids", ctypes.c_void_p), ("batch", ctypes.c_int32),
                ("active_count", ctypes.c_int32), ("widths", ctypes.c_void_p), ("heights", ctypes.c_void_p),
                ("pixel_offsets", ctypes.c_void_p)]


class _Options(ctypes.Structure):
    _fields_ = [("height", ctypes.c_int32), ("width", ctypes.c_int32), ("near_plane", ctypes.c_float),
                ("far_plane", ctypes.c_float), ("gaussian_support_sigma", ctypes.c_float),
                ("covariance_epsilon", ctypes.c_float), ("semantic_min_alpha", ctypes.c_float),
                ("antialiased", ctypes.c_int32), ("tile_size", ctypes.c_int32), ("output_srgb", ctypes.c_int32),
                ("full_sensor_output", ctypes.c_int32), ("threads", ctypes.c_int32),
                ("depth_bucket_count", ctypes.c_int32), ("depth_bucket_group_size", ctypes.c_int32),
                ("ray_evaluation", ctypes.c_int32)]


class _Outputs(ctypes.Structure):
    _fields_ = [("rgb", ctypes.c_void_p), ("depth", ctypes.c_void_p), ("alpha", ctypes.c_void_p),
                ("semantic", ctypes.c_void_p), ("counters", ctypes.c_int64 * 5), ("tile_counts", ctypes.c_void_p),
                ("sorted_keys", ctypes.c_void_p), ("sorted_gaussians", ctypes.c_void_p),
                ("sorted_capacity", ctypes.c_int64), ("global_tiles", ctypes.c_int64)]


_LIB: ctypes.CDLL | None = None


def build() -> None:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True)


def library() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        if not LIBRARY_PATH.is_file():
            build()
        lib = ctypes.CDLL(str(LIBRARY_PATH))
        lib.oracle_render.restype = ctypes.c_int32
        lib.oracle_render.argtypes = [ctypes.POINTER(_Scene), ctypes.POINTER(_Cameras), ctypes.POINTER(_Options),
                                      ctypes.POINTER(_Outputs)]
        lib.oracle_precompute_covariances.restype = ctypes.c_int32
        lib.oracle_precompute_covariances.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                                      ctypes.c_void_p]
        for name in ("oracle_expf", "oracle_expf_nonpositive", "oracle_logf", "oracle_srgb"):
            getattr(lib, name).restype = ctypes.c_float
            getattr(lib, name).argtypes = [ctypes.c_float]
        lib.oracle_sh_colour.restype = None
        lib.oracle_sh_colour.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_float, ctypes.c_float,
                                         ctypes.c_float, ctypes.c_void_p]
        lib.oracle_last_error.restype = ctypes.c_char_p
        _LIB = lib
    return _LIB


def _array(value, dtype) -> np.ndarray:
    if hasattr(value, "detach"):
        value = value.detach().cpu().numpy()
    return np.ascontiguousarray(np.asarray(value, dtype=dtype))


def _p(array: np.ndarray | None) -> int | None:
    return None if array is None or array.size == 0 else array.ctypes.data


def covariances(scales, rotations) -> np.ndarray:
    s = _array(scales, np.float32)
    q = _array(rotations, np.float32)
    out = np.empty((s.shape[0], 6), dtype=np.float32)
    library().oracle_precompute_covariances(_p(s), _p(q), s.shape[0], _p(out))
    return out


def expf(x: float) -> float:
    return float(library().oracle_expf(x))


def expf_nonpositive(x: float) -> float:
    return float(library().oracle_expf_nonpositive(x))


def logf(x: float) -> float:
    return float(library().oracle_logf(x))


def srgb(x: float) -> float:
    return float(library().oracle_srgb(x))


def sh_colour(features, degree: int, direction) -> np.ndarray:
    f = _array(features, np.float32)
    out = np.empty(3, dtype=np.float32)
    library().oracle_sh_colour(_p(f), degree, float(direction[0]), float(direction[1]), float(direction[2]), _p(out))
    return out


def render(
    scenes: list[dict],
    scene_ids_registered: list[int],
    viewmats,
    intrinsics,
    camera_scene_ids,
    *,
    height: int,
    width: int,
    env_xforms=None,
    active_ids=None,
    resolutions: list[tuple[int, int]] | None = None,
    near_plane: float = 0.01,
    far_plane: float = 100.0,
    gaussian_support_sigma: float = 3.33,
    covariance_epsilon: float = 0.3,
    semantic_min_alpha: float = 0.01,
    antialiased: bool = False,
    tile_size: int = 16,
    output_srgb: bool = False,
    full_sensor_output: bool = True,
    sh_degree: int = 0,
    threads: int = 0,
    want_sorted: bool = False,
    background: bool = True,
    depth_bucket_count: int = 0,
    depth_bucket_group_size: int = 0,
    ray_evaluation: bool = False,
) -> dict[str, np.ndarray]:
    """Render a batch on the CPU.

    ``scenes`` are dicts with means/scales/rotations/opacities/features/semantic_ids
    in registration order.  Returns rgb/depth/alpha/semantic ([B,H,W,C], or flat
    [P,C] with ``resolutions``), counters, tile_counts and optionally the sorted
    (key, Gaussian ID) sequence.
    """
    means = np.concatenate([_array(s["means"], np.float32) for s in scenes]).reshape(-1, 3)
    scales = np.concatenate([_array(s["scales"], np.float32) for s in scenes]).reshape(-1, 3)
    rotations = np.concatenate([_array(s["rotations"], np.float32) for s in scenes]).reshape(-1, 4)
    opacities = np.concatenate([_array(s["opacities"], np.float32) for s in scenes]).reshape(-1)
    features = np.concatenate([_array(s["features"], np.float32) for s in scenes])
    semantic_ids = np.concatenate([_array(s["semantic_ids"], np.int64) for s in scenes]).reshape(-1)
    cov = covariances(scales, rotations)
    offsets = np.zeros(len(scenes) + 1, dtype=np.int64)
    for i, s in enumerate(scenes):
        offsets[i + 1] = offsets[i] + _array(s["means"], np.float32).reshape(-1, 3).shape[0]
    registered = np.asarray(scene_ids_registered, dtype=np.int64)
    view = _array(viewmats, np.float32).reshape(-1, 4, 4)
    k = _array(intrinsics, np.float32).reshape(-1, 3, 3)
    batch = view.shape[0]
    env = (np.tile(np.eye(4, dtype=np.float32), (batch, 1, 1)) if env_xforms is None
           else _array(env_xforms, np.float32).reshape(-1, 4, 4))
    cam_scenes = _array(camera_scene_ids, np.int64).reshape(-1)
    active = None if active_ids is None or len(active_ids) == 0 else _array(active_ids, np.int64).reshape(-1)

    c_scene = _Scene(_p(means), _p(cov), _p(opacities), _p(features), _p(semantic_ids), _p(registered), _p(offsets),
                     means.shape[0], features.shape[1], len(scenes), sh_degree, 0)
    widths = heights = pixel_offsets = None
    bin_size = 32 if ray_evaluation and tile_size == 16 else tile_size
    if resolutions is None:
        pixels = batch * height * width
        tiles = batch * ((width + bin_size - 1) // bin_size) * ((height + bin_size - 1) // bin_size)
        shape = (batch, height, width)
    else:
        widths = np.asarray([r[0] for r in resolutions], dtype=np.int32)
        heights = np.asarray([r[1] for r in resolutions], dtype=np.int32)
        sizes = widths.astype(np.int64) * heights
        pixel_offsets = np.concatenate(([0], np.cumsum(sizes)[:-1])).astype(np.int64)
        pixels = int(sizes.sum())
        tiles = int(sum(((w + tile_size - 1) // tile_size) * ((h + tile_size - 1) // tile_size)
                        for w, h in resolutions))
        shape = (pixels,)
    c_cams = _Cameras(_p(view), _p(k), _p(env), _p(cam_scenes), _p(active), batch,
                      0 if active is None else active.shape[0], _p(widths), _p(heights), _p(pixel_offsets))
    c_opt = _Options(height, width, near_plane, far_plane, gaussian_support_sigma, covariance_epsilon,
                     semantic_min_alpha, int(antialiased), tile_size, int(output_srgb), int(full_sensor_output),
                     threads, int(depth_bucket_count), int(depth_bucket_group_size), int(bool(ray_evaluation)))
    rgb = np.zeros(shape + (3,), dtype=np.float32)
    depth = np.full(shape + (1,), np.inf if background else 0.0, dtype=np.float32)
    alpha = np.zeros(shape + (1,), dtype=np.float32)
    semantic = np.full(shape + (1,), -1, dtype=np.int64)
    tile_counts = np.zeros(tiles, dtype=np.int32)
    out = _Outputs()
    out.rgb, out.depth, out.alpha, out.semantic = _p(rgb), _p(depth), _p(alpha), _p(semantic)
    out.tile_counts = _p(tile_counts)
    out.global_tiles = tiles
    sorted_keys = sorted_gaussians = None
    if want_sorted:
        # First pass for counts, then the sorted sequence.
        status = library().oracle_render(ctypes.byref(c_scene), ctypes.byref(c_cams), ctypes.byref(c_opt),
                                         ctypes.byref(out))
        if status != 0:
            raise RuntimeError(library().oracle_last_error().decode())
        total = int(out.counters[1])
        sorted_keys = np.zeros(max(total, 1), dtype=np.uint64)
        sorted_gaussians = np.zeros(max(total, 1), dtype=np.int32)
        out.sorted_keys = _p(sorted_keys)
        out.sorted_gaussians = _p(sorted_gaussians)
        out.sorted_capacity = total
    status = library().oracle_render(ctypes.byref(c_scene), ctypes.byref(c_cams), ctypes.byref(c_opt),
                                     ctypes.byref(out))
    if status != 0:
        raise RuntimeError(library().oracle_last_error().decode())
    result = {
        "rgb": rgb,
        "depth": depth,
        "alpha": alpha,
        "semantic_id": semantic,
        "counters": np.asarray(list(out.counters), dtype=np.int64),
        "tile_counts": tile_counts,
        "covariances": cov,
    }
    if want_sorted:
        total = int(out.counters[1])
        result["sorted_keys"] = sorted_keys[:total]
        result["sorted_gaussians"] = sorted_gaussians[:total]
    return result


def cpu_count() -> int:
    return os.cpu_count() or 1
