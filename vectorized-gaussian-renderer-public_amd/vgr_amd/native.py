"""ctypes binding of ``libvgr_hip.so`` (the gfx950 kernels behind a C ABI).

The module object returned by :func:`load_native_module` exposes the same
method names and positional argument order as the reference's pybind module
(``src/isaacsim_gaussian_renderer/native/renderer.cpp:312-337``):
``render``, ``sort_temp_bytes``, ``radix_sort_double_buffer_temp_bytes``,
``precompute_covariances``, ``prepare_chunked_active_ids`` and
``aggregate_chunk_counters``, so the backend's ``native_module`` seam is
unchanged.  Tensors are handed over as raw device pointers; work is enqueued on
the current torch stream of the tensors' device and nothing synchronizes the
host.

There is no fallback: if the library is missing, or a call fails, this module
raises.
"""

from __future__ import annotations

import ctypes
import math
import os
import struct
from pathlib import Path
from typing import Sequence

import torch

LIBRARY_NAME = "libvgr_hip.so"
LIBRARY_PATH = Path(__file__).resolve().with_name(LIBRARY_NAME)

# Workspace tensors passed positionally to render() (this build's layout; the
# reference used 24 tensors, renderer_cuda.cu:73-98).
WORKSPACE_TENSORS = (
    "visible_records",
    "keys_in",
    "keys_out",
    "values_in",
    "values_out",
    "tile_starts",
    "tile_ends",
    "counters",
    "scratch",
)
RECORD_FLOATS = 12  # 48-byte visible record, see vgr_record in include/vgr.h
RECORD_GRAD_FLOATS = 10  # dL/d (mean_x, mean_y, conic_0..2, opacity, depth, r, g, b) per intersection slot
CAMERA_GRAD_FLOATS = 24  # per-camera pose partials of the backward pass
POSE_REPLICAS = 32  # copies of the pose partials (VGR_POSE_REPLICAS)
RAY_RECORD_FLOATS = 12  # exact-ray precision data per visible record (workspace.ray_records)
ABI_VERSION = 7
# Device counter block: the reference's five counters (cuda_backend.py:22-28) plus
# one slot counting radix-sort look-back waits that gave up (the order is then wrong).
COUNTER_SLOTS = 6
SORT_FAILURE_SLOT = 5
# Intersection orderings (vgr_ordering in include/vgr.h).
ORDERING_GLOBAL_RADIX = 0
ORDERING_TILE_FRONT = 1
# Largest intersection capacity of one submission per ordering: the radix sort's
# look-back flags carry 31-bit prefixes; tile-front buckets use int32 offsets.
MAX_INTERSECTIONS = {ORDERING_GLOBAL_RADIX: (1 << 31) - 1, ORDERING_TILE_FRONT: (1 << 31) - 1}
# Tile-front ordering limits: tiles per camera (16x16 tiles).
FRONT_MAX_CAMERA_TILES = 2048


class _Scene(ctypes.Structure):
    _fields_ = [
        ("means", ctypes.c_void_p),
        ("covariances", ctypes.c_void_p),
        ("opacities", ctypes.c_void_p),
        ("features", ctypes.c_void_p),
        ("semantic_ids", ctypes.c_void_p),
        ("registered_scene_ids", ctypes.c_void_p),
        ("scene_offsets", ctypes.c_void_p),
        ("gaussian_count", ctypes.c_int64),
        ("max_scene_gaussians", ctypes.c_int64),
        ("feature_width", ctypes.c_int32),
        ("scene_count", ctypes.c_int32),
        ("sh_degree", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("gaussian_ids", ctypes.c_void_p),
        ("chunk_bounds", ctypes.c_void_p),
    ]


class _Cameras(ctypes.Structure):
    _fields_ = [
        ("viewmats", ctypes.c_void_p),
        ("intrinsics", ctypes.c_void_p),
        ("env_xforms", ctypes.c_void_p),
        ("scene_ids", ctypes.c_void_p),
        ("active_ids", ctypes.c_void_p),
        ("batch", ctypes.c_int32),
        ("active_count", ctypes.c_int32),
        ("widths", ctypes.c_void_p),
        ("heights", ctypes.c_void_p),
        ("pixel_offsets", ctypes.c_void_p),
        ("max_global_tiles", ctypes.c_int64),
        ("max_camera_tiles", ctypes.c_int64),
    ]


class _Outputs(ctypes.Structure):
    _fields_ = [
        ("rgb", ctypes.c_void_p),
        ("depth", ctypes.c_void_p),
        ("alpha", ctypes.c_void_p),
        ("semantic", ctypes.c_void_p),
        ("final_transmittance", ctypes.c_void_p),
        ("last_blended", ctypes.c_void_p),
    ]


class _Gradients(ctypes.Structure):
    _fields_ = [
        ("grad_rgb", ctypes.c_void_p),
        ("grad_depth", ctypes.c_void_p),
        ("grad_alpha", ctypes.c_void_p),
        ("record_grads", ctypes.c_void_p),
        ("grad_means", ctypes.c_void_p),
        ("grad_covariances", ctypes.c_void_p),
        ("grad_opacities", ctypes.c_void_p),
        ("grad_features", ctypes.c_void_p),
        ("grad_viewmats", ctypes.c_void_p),
        ("camera_grads", ctypes.c_void_p),
        ("sorted_selector", ctypes.c_int64),
    ]


class _Workspace(ctypes.Structure):
    _fields_ = [
        ("records", ctypes.c_void_p),
        ("radii", ctypes.c_void_p),
        ("keys_in", ctypes.c_void_p),
        ("keys_out", ctypes.c_void_p),
        ("values_in", ctypes.c_void_p),
        ("values_out", ctypes.c_void_p),
        ("tile_starts", ctypes.c_void_p),
        ("tile_ends", ctypes.c_void_p),
        ("counters", ctypes.c_void_p),
        ("scratch", ctypes.c_void_p),
        ("visible_capacity", ctypes.c_int64),
        ("intersection_capacity", ctypes.c_int64),
        ("tile_capacity", ctypes.c_int64),
        ("scratch_bytes", ctypes.c_int64),
        ("slot_records", ctypes.c_void_p),
        ("ray_records", ctypes.c_void_p),
    ]


class _Options(ctypes.Structure):
    _fields_ = [
        ("height", ctypes.c_int32),
        ("width", ctypes.c_int32),
        ("near_plane", ctypes.c_float),
        ("far_plane", ctypes.c_float),
        ("gaussian_support_sigma", ctypes.c_float),
        ("covariance_epsilon", ctypes.c_float),
        ("semantic_min_alpha", ctypes.c_float),
        ("antialiased", ctypes.c_int32),
        ("tile_size", ctypes.c_int32),
        ("output_srgb", ctypes.c_int32),
        ("deterministic", ctypes.c_int32),
        ("full_sensor_output", ctypes.c_int32),
        ("fixed_capacity_sort", ctypes.c_int32),
        ("sort_digit_bits", ctypes.c_int32),
        ("raster_batch", ctypes.c_int32),
        ("cameras_per_block", ctypes.c_int32),
        ("reuse_projection", ctypes.c_int32),
        ("stage_events", ctypes.POINTER(ctypes.c_void_p)),
        ("stage_event_count", ctypes.c_int32),
        ("sort_tile_items", ctypes.c_int32),
        ("statistics", ctypes.c_void_p),
        ("persistent_composite", ctypes.c_int32),
        ("chunk_culling", ctypes.c_int32),
        ("lookback_spin_limit", ctypes.c_int32),
        ("ordering", ctypes.c_int32),
        ("depth_bucket_count", ctypes.c_int32),
        ("depth_bucket_group_size", ctypes.c_int32),
        ("ray_evaluation", ctypes.c_int32),
    ]


def dense_cutoff(tile_size: int, depth_bucket_count: int, depth_bucket_group_size: int,
                 compact_projection_cache: bool, deterministic: bool) -> tuple[int, int]:
    """(buckets, buckets per pass) of the depth cutoff for the reference's knobs, (0, 0) for none.

    The reference runs the cutoff at tile size 1 in nondeterministic mode, over
    ``depth_bucket_group_size`` buckets per pass, or over every bucket in one pass with
    ``compact_projection_cache`` (renderer_cuda.cu:2888-3144).  Images do not depend on it.
    """
    if tile_size != 1 or depth_bucket_count <= 0 or (deterministic and not compact_projection_cache):
        return 0, 0
    group = depth_bucket_count if compact_projection_cache else min(depth_bucket_group_size, depth_bucket_count)
    return int(depth_bucket_count), int(group)


class _Sizes(ctypes.Structure):
    _fields_ = [
        ("scratch_bytes", ctypes.c_int64),
        ("tile_capacity", ctypes.c_int64),
        ("sort_key_bits", ctypes.c_int32),
        ("sort_passes", ctypes.c_int32),
    ]


_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_P = ctypes.c_void_p

EXPORTED_SYMBOLS: dict[str, tuple[type, tuple]] = {
    "vgr_abi_version": (_I32, ()),
    "vgr_last_error": (ctypes.c_char_p, ()),
    "vgr_emitted_sort_key_end_bit": (_I32, (_I64,)),
    "vgr_workspace_sizes_query": (
        _I32, (ctypes.POINTER(_Cameras), ctypes.POINTER(_Options), _I64, _I64, ctypes.POINTER(_Sizes))),
    "vgr_sort_temp_bytes": (_I64, (_I64, _I64)),
    "vgr_precompute_covariances": (_I32, (_P, _P, _I64, _P, _P)),
    "vgr_chunk_bounds": (_I32, (_P, _P, _I64, _P, _P)),
    "vgr_render": (
        _I32,
        (ctypes.POINTER(_Scene), ctypes.POINTER(_Cameras), ctypes.POINTER(_Outputs), ctypes.POINTER(_Workspace),
         ctypes.POINTER(_Options), _P, ctypes.POINTER(_I64)),
    ),
    "vgr_static_scratch_bytes": (_I64, ()),
    "vgr_render_static_scratch": (
        _I32,
        (ctypes.POINTER(_Scene), ctypes.POINTER(_Cameras), ctypes.POINTER(_Outputs), ctypes.POINTER(_Workspace),
         ctypes.POINTER(_Options), _P, ctypes.POINTER(_I64)),
    ),
    "vgr_render_backward": (
        _I32,
        (ctypes.POINTER(_Scene), ctypes.POINTER(_Cameras), ctypes.POINTER(_Outputs), ctypes.POINTER(_Workspace),
         ctypes.POINTER(_Options), ctypes.POINTER(_Gradients), _P),
    ),
    "vgr_covariance_backward": (_I32, (_P, _P, _P, _I64, _P, _P, _P)),
    "vgr_prepare_chunked_active_ids": (_I32, (_P, _I64, _P, _P, _I64, _I64, _P)),
    "vgr_aggregate_chunk_counters": (_I32, (_P, _I64, _I64, _P, _P, _P)),
    "vgr_fill_i64": (_I32, (_P, _I64, _I64, _P)),
    "vgr_fill_f32": (_I32, (_P, _I64, ctypes.c_float, _P)),
    "vgr_pose_compose": (_I32, (_P, _P, _I64, _P, _P)),
    "vgr_pose_compose_backward": (_I32, (_P, _P, _P, _I64, _P, _P, _P)),
}


def load_library(path: Path | str | None = None) -> ctypes.CDLL:
    """Load the C-ABI library and declare every exported signature.

    ``import torch`` has already run, so the HIP runtime torch ships is the
    one this library binds to (both carry the soname libamdhip64.so.7).
    """
    target = Path(path) if path is not None else LIBRARY_PATH
    if not target.is_file():
        raise RuntimeError(
            f"{target} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
            "(or make -C vectorized-gaussian-renderer-public_amd/csrc) first.")
    library = ctypes.CDLL(str(target), mode=os.RTLD_NOW | os.RTLD_LOCAL)
    for name, (restype, argtypes) in EXPORTED_SYMBOLS.items():
        function = getattr(library, name)
        function.restype = restype
        function.argtypes = list(argtypes)
    return library


def _ptr(tensor: torch.Tensor | None) -> int | None:
    if tensor is None or tensor.numel() == 0:
        return None
    return tensor.data_ptr()


def _stream(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def infer_sh_degree(feature_width: int) -> int:
    """Largest SH degree whose bands fit in ``feature_width`` as 3*(d+1)^2 channels."""
    bands = feature_width // 3
    if bands * 3 != feature_width:
        return 0
    degree = int(math.isqrt(bands)) - 1
    if (degree + 1) ** 2 != bands or degree > 3:
        return 0
    return degree


class NativeRenderer:
    """Python face of libvgr_hip with the reference native module's method names."""

    # Intersection counter includes intersections of records that did not fit
    # the visible capacity (the reference counts only stored records).
    counts_full_intersection_demand = True

    def __init__(self, library: ctypes.CDLL | None = None) -> None:
        self.lib = library if library is not None else load_library()
        version = int(self.lib.vgr_abi_version())
        if version != ABI_VERSION:
            raise RuntimeError(f"libvgr_hip ABI version {version} is not supported by this binding "
                               f"(expected {ABI_VERSION}).")

    # -- helpers --------------------------------------------------------------------------
    def _check(self, status: int, what: str) -> None:
        if status != 0:
            message = self.lib.vgr_last_error().decode("utf-8", "replace")
            raise RuntimeError(f"{what} failed (status {status}): {message}")

    @staticmethod
    def _options(**values) -> _Options:
        options = _Options()
        for name, value in values.items():
            setattr(options, name, value)
        return options

    # -- reference-compatible surface -----------------------------------------------------
    def emitted_sort_key_end_bit(self, global_tile_capacity: int) -> int:
        bit = int(self.lib.vgr_emitted_sort_key_end_bit(int(global_tile_capacity)))
        if bit < 0:
            raise ValueError("Global tile capacity must be in [1, INT32_MAX].")
        return bit

    def sort_temp_bytes(self, n_items: int, n_segments: int) -> int:
        value = int(self.lib.vgr_sort_temp_bytes(int(n_items), int(n_segments)))
        if value < 0:
            raise ValueError("sort_temp_bytes: invalid arguments")
        return value

    radix_sort_double_buffer_temp_bytes = sort_temp_bytes

    def scratch_bytes(self, *, max_views: int, height: int, width: int, gaussians: int, intersection_capacity: int,
                      tile_size: int, near_plane: float, far_plane: float, sort_digit_bits: int = 0,
                      cameras_per_block: int = 0, sort_tile_items: int = 0, chunk_culling: bool = False,
                      resolutions: Sequence[tuple[int, int]] | None = None,
                      ordering: int = ORDERING_GLOBAL_RADIX, depth_bucket_count: int = 0,
                      depth_bucket_group_size: int = 0, ray_evaluation: bool = False) -> int:
        """Scratch bytes one render of up to ``max_views`` cameras needs (``depth_bucket_count`` /
        ``depth_bucket_group_size``: the effective cutoff of :func:`dense_cutoff`)."""
        cameras = _Cameras()
        cameras.batch = max_views
        cameras.active_count = max_views
        if resolutions is not None:
            cameras.widths = 1  # marks per-camera resolutions; only the tile total is read
            cameras.heights = 1
            per_camera = [((w + tile_size - 1) // tile_size) * ((h + tile_size - 1) // tile_size)
                          for w, h in resolutions]
            cameras.max_global_tiles = sum(per_camera)
            cameras.max_camera_tiles = max(per_camera)
        options = self._options(height=height, width=width, near_plane=near_plane, far_plane=far_plane,
                                gaussian_support_sigma=3.0, covariance_epsilon=0.0, tile_size=tile_size,
                                sort_digit_bits=sort_digit_bits, cameras_per_block=cameras_per_block,
                                sort_tile_items=sort_tile_items, chunk_culling=int(bool(chunk_culling)),
                                ordering=int(ordering), depth_bucket_count=int(depth_bucket_count),
                                depth_bucket_group_size=int(depth_bucket_group_size),
                                ray_evaluation=int(bool(ray_evaluation)))
        sizes = _Sizes()
        self._check(self.lib.vgr_workspace_sizes_query(ctypes.byref(cameras), ctypes.byref(options), int(gaussians),
                                                        int(intersection_capacity), ctypes.byref(sizes)),
                    "vgr_workspace_sizes_query")
        return int(sizes.scratch_bytes)

    def static_scratch_bytes(self) -> int:
        """Bytes of the per-device scratch array behind ``render(..., static_scratch=True)``."""
        return int(self.lib.vgr_static_scratch_bytes())

    def precompute_covariances(self, scales: torch.Tensor, rotations: torch.Tensor, out: torch.Tensor) -> None:
        count = int(scales.shape[0])
        if count == 0:
            return
        with torch.cuda.device(scales.device):
            self._check(self.lib.vgr_precompute_covariances(_ptr(scales), _ptr(rotations), count, _ptr(out),
                                                            _stream(scales.device)),
                        "vgr_precompute_covariances")

    def chunk_bounds(self, means: torch.Tensor, covariances: torch.Tensor, out: torch.Tensor) -> None:
        """Per-chunk (256 Gaussians) bounds: min xyz, max xyz, sqrt(max trace(cov)), 0."""
        count = int(means.shape[0])
        if count == 0:
            return
        if out.dtype != torch.float32 or out.numel() < 8 * ((count + 255) // 256):
            raise ValueError("chunk bounds output must be float32 [ceil(G/256), 8]")
        with torch.cuda.device(means.device):
            self._check(self.lib.vgr_chunk_bounds(_ptr(means), _ptr(covariances), count, _ptr(out),
                                                  _stream(means.device)), "vgr_chunk_bounds")

    def prepare_chunked_active_ids(self, active_ids: torch.Tensor, active_mask: torch.Tensor,
                                   chunk_active_ids: torch.Tensor, logical_batch: int, physical_batch: int) -> None:
        with torch.cuda.device(active_mask.device):
            self._check(self.lib.vgr_prepare_chunked_active_ids(
                _ptr(active_ids), int(active_ids.numel()), _ptr(active_mask), _ptr(chunk_active_ids),
                int(logical_batch), int(physical_batch), _stream(active_mask.device)),
                "vgr_prepare_chunked_active_ids")

    def aggregate_chunk_counters(self, chunk_counters: torch.Tensor, counters: torch.Tensor,
                                 physical_capacity_counters: torch.Tensor, chunk_count: int) -> None:
        with torch.cuda.device(counters.device):
            self._check(self.lib.vgr_aggregate_chunk_counters(
                _ptr(chunk_counters), int(chunk_counters.stride(0)), int(chunk_count), _ptr(counters),
                _ptr(physical_capacity_counters), _stream(counters.device)), "vgr_aggregate_chunk_counters")

    def render(
        self,
        scene: Sequence[torch.Tensor],
        cameras: Sequence[torch.Tensor],
        outputs: Sequence[torch.Tensor],
        workspace: Sequence[torch.Tensor],
        height: int,
        width: int,
        max_scene_gaussians: int,
        near_plane: float,
        far_plane: float,
        gaussian_support_sigma: float,
        covariance_epsilon: float,
        antialiased: bool,
        semantic_min_alpha: float,
        ray_gaussian_evaluation: bool,
        tile_size: int,
        depth_bucket_count: int,
        depth_bucket_group_size: int,
        compact_projection_cache: bool,
        materialize_projected_records: bool,
        reuse_projection: bool,
        output_srgb: bool,
        deterministic: bool,
        full_sensor_output: bool,
        fixed_capacity_sort: bool,
        *,
        sh_degree: int = 0,
        sort_digit_bits: int = 0,
        raster_batch: int = 0,
        cameras_per_block: int = 0,
        sort_tile_items: int = 0,
        persistent_composite: int = 0,
        gaussian_ids: torch.Tensor | None = None,
        chunk_bounds: torch.Tensor | None = None,
        resolutions: tuple[torch.Tensor, torch.Tensor, torch.Tensor, int] | None = None,
        stage_events: Sequence[torch.cuda.Event] | None = None,
        statistics: torch.Tensor | None = None,
        checkpoints: tuple[torch.Tensor, torch.Tensor] | None = None,
        lookback_spin_limit: int = 0,
        static_scratch: bool = False,
        ordering: int = ORDERING_GLOBAL_RADIX,
        max_camera_tiles: int = 0,
        slot_records: torch.Tensor | None = None,
        ray_records: torch.Tensor | None = None,
    ) -> int:
        """One render submission; returns the sorted-buffer selector (1: *_out holds the result).

        ``stage_events`` (5 recorded torch events) are re-recorded at the stage
        boundaries on the render stream; ``statistics`` (int64[4] on the device)
        receives the compositor's evaluation count.  ``checkpoints`` (float32
        final transmittance and int32 last-blended position, one per pixel) makes
        the forward store what :meth:`render_backward` needs.  ``lookback_spin_limit``
        (0 = the library default) bounds the sort's look-back waits; a wait that
        gives up is counted in ``counters[SORT_FAILURE_SLOT]``.  ``static_scratch`` calls
        ``vgr_render_static_scratch``: the library's own per-device scratch array replaces the
        ``scratch`` workspace tensor (which may then be empty), and no device memory is allocated.
        ``ordering`` selects the intersection ordering (``ORDERING_TILE_FRONT``: per-tile buckets with
        an LDS sort of each tile's nearest entries, see include/vgr.h); ``max_camera_tiles`` is the
        largest tile count of one camera when ``resolutions`` are given (0: unknown).  Training
        forwards (``checkpoints``) also take ``slot_records`` (int32, one per intersection slot):
        the sort then carries slots, which :meth:`render_backward` needs.
        """
        if ray_gaussian_evaluation and ray_records is None:
            raise ValueError("ray_gaussian_evaluation needs the ray_records workspace (RAY_RECORD_FLOATS per "
                             "visible record)")
        if (compact_projection_cache or materialize_projected_records) and int(tile_size) != 1:
            raise ValueError("compact_projection_cache requires tile_size=1.")
        # Records are always materialized here (48-byte records written by the projection
        # pass), so materialize_projected_records selects nothing further.
        cutoff_buckets, cutoff_group = dense_cutoff(int(tile_size), int(depth_bucket_count),
                                                    int(depth_bucket_group_size), bool(compact_projection_cache),
                                                    bool(deterministic))
        if len(scene) != 7 or len(cameras) != 5 or len(outputs) != 4 or len(workspace) != len(WORKSPACE_TENSORS):
            raise ValueError("render expects 7 scene, 5 camera, 4 output and "
                             f"{len(WORKSPACE_TENSORS)} workspace tensors.")
        means, covariances, opacities, features, semantic_ids, registered, offsets = scene
        viewmats, intrinsics, env_xforms, camera_scene_ids, active_ids = cameras
        rgb, depth, alpha, semantic = outputs
        records, keys_in, keys_out, values_in, values_out, tile_starts, tile_ends, counters, scratch = workspace
        _require_dtypes(
            (means, torch.float32, "means"), (covariances, torch.float32, "covariances"),
            (opacities, torch.float32, "opacities"), (features, torch.float32, "features"),
            (semantic_ids, torch.int64, "semantic_ids"), (registered, torch.int64, "registered_scene_ids"),
            (offsets, torch.int64, "scene_offsets"), (viewmats, torch.float32, "viewmats"),
            (intrinsics, torch.float32, "intrinsics"), (env_xforms, torch.float32, "env_xforms"),
            (camera_scene_ids, torch.int64, "camera_scene_ids"), (active_ids, torch.int64, "active_camera_ids"),
            (rgb, torch.float32, "rgb"), (depth, torch.float32, "depth"), (alpha, torch.float32, "alpha"),
            (semantic, torch.int64, "semantic"), (records, torch.float32, "visible_records"),
            (keys_in, torch.uint64, "keys_in"), (keys_out, torch.uint64, "keys_out"),
            (values_in, torch.int32, "values_in"), (values_out, torch.int32, "values_out"),
            (tile_starts, torch.int32, "tile_starts"), (tile_ends, torch.int32, "tile_ends"),
            (counters, torch.int64, "counters"), (scratch, torch.uint8, "scratch"))
        tensors = [t for t in (*scene, *cameras, *outputs, *workspace) if t.numel() > 0]
        device = means.device
        for tensor in tensors:
            if tensor.device != device:
                raise ValueError(f"all render tensors must be on {device}; found {tensor.device}.")
            if not tensor.is_contiguous():
                raise ValueError("all render tensors must be contiguous.")
        if device.type != "cuda":
            raise ValueError(f"render tensors must be on a GPU device; got {device}.")
        if int(counters.numel()) < COUNTER_SLOTS:
            raise ValueError(f"counters workspace must contain at least {COUNTER_SLOTS} int64 values")
        if env_xforms.shape[0] < viewmats.shape[0]:
            raise ValueError("env_xforms must provide one transform per camera")

        c_scene = _Scene(
            means=_ptr(means), covariances=_ptr(covariances), opacities=_ptr(opacities), features=_ptr(features),
            semantic_ids=_ptr(semantic_ids), registered_scene_ids=_ptr(registered), scene_offsets=_ptr(offsets),
            gaussian_count=int(means.shape[0]), max_scene_gaussians=int(max_scene_gaussians),
            feature_width=int(features.shape[1]) if features.ndim == 2 else 0,
            scene_count=int(registered.numel()), sh_degree=int(sh_degree), reserved=0,
            gaussian_ids=_ptr(gaussian_ids), chunk_bounds=_ptr(chunk_bounds))
        if gaussian_ids is not None and (gaussian_ids.dtype != torch.int32 or gaussian_ids.numel() != means.shape[0]):
            raise ValueError("gaussian_ids must be int32 with one entry per packed Gaussian")
        if chunk_bounds is not None and (chunk_bounds.dtype != torch.float32
                                         or chunk_bounds.numel() < 8 * ((int(means.shape[0]) + 255) // 256)):
            raise ValueError("chunk_bounds must be float32 [ceil(G/256), 8]")
        batch = int(viewmats.shape[0])
        c_cameras = _Cameras(
            viewmats=_ptr(viewmats), intrinsics=_ptr(intrinsics), env_xforms=_ptr(env_xforms),
            scene_ids=_ptr(camera_scene_ids), active_ids=_ptr(active_ids), batch=batch,
            active_count=int(active_ids.numel()) if active_ids.numel() else batch,
            widths=None, heights=None, pixel_offsets=None, max_global_tiles=0)
        if resolutions is not None:
            widths, heights, pixel_offsets, total_tiles = resolutions
            c_cameras.widths = _ptr(widths)
            c_cameras.heights = _ptr(heights)
            c_cameras.pixel_offsets = _ptr(pixel_offsets)
            c_cameras.max_global_tiles = int(total_tiles)
            c_cameras.max_camera_tiles = int(max_camera_tiles)
        c_outputs = _Outputs(rgb=_ptr(rgb), depth=_ptr(depth) if full_sensor_output else None,
                             alpha=_ptr(alpha) if full_sensor_output else None,
                             semantic=_ptr(semantic) if full_sensor_output else None)
        if checkpoints is not None:
            transmittance, last_blended = checkpoints
            _require_dtypes((transmittance, torch.float32, "final_transmittance"),
                            (last_blended, torch.int32, "last_blended"))
            if transmittance.numel() * 3 != rgb.numel() or last_blended.numel() * 3 != rgb.numel():
                raise ValueError("checkpoints need one value per output pixel")
            if transmittance.device != device or last_blended.device != device:
                raise ValueError("checkpoints must be on the render device")
            c_outputs.final_transmittance = _ptr(transmittance)
            c_outputs.last_blended = _ptr(last_blended)
        c_workspace = _Workspace(
            records=_ptr(records), radii=None, keys_in=_ptr(keys_in), keys_out=_ptr(keys_out),
            values_in=_ptr(values_in), values_out=_ptr(values_out), tile_starts=_ptr(tile_starts),
            tile_ends=_ptr(tile_ends), counters=_ptr(counters), scratch=_ptr(scratch),
            visible_capacity=int(records.shape[0]), intersection_capacity=int(keys_in.numel()),
            tile_capacity=int(tile_starts.numel()), scratch_bytes=int(scratch.numel()))
        if ray_gaussian_evaluation:
            if (ray_records.dtype != torch.float32 or ray_records.device != device or not ray_records.is_contiguous()
                    or ray_records.numel() < RAY_RECORD_FLOATS * int(records.shape[0])):
                raise ValueError("ray_records must be a contiguous float32 tensor of RAY_RECORD_FLOATS per visible "
                                 "record on the render device")
            c_workspace.ray_records = ray_records.data_ptr()
        if cutoff_buckets > 0:
            # The cutoff reads the records' radii between the records pass and the
            # emission pass; keys_out is idle until the sort, so it holds them.
            if keys_out.numel() < int(records.shape[0]):
                raise ValueError("the depth cutoff keeps the records' radii in keys_out: the intersection "
                                 "capacity must be at least the visible capacity")
            c_workspace.radii = keys_out.data_ptr()
        if slot_records is not None:
            if (slot_records.dtype != torch.int32 or slot_records.device != device
                    or slot_records.numel() < keys_in.numel() or not slot_records.is_contiguous()):
                raise ValueError("slot_records must be a contiguous int32 tensor with one entry per intersection "
                                 "slot on the render device")
            c_workspace.slot_records = slot_records.data_ptr()
        c_options = self._options(
            height=int(height), width=int(width), near_plane=float(near_plane), far_plane=float(far_plane),
            gaussian_support_sigma=float(gaussian_support_sigma), covariance_epsilon=float(covariance_epsilon),
            semantic_min_alpha=float(semantic_min_alpha), antialiased=int(bool(antialiased)),
            tile_size=int(tile_size), output_srgb=int(bool(output_srgb)), deterministic=int(bool(deterministic)),
            full_sensor_output=int(bool(full_sensor_output)), fixed_capacity_sort=int(bool(fixed_capacity_sort)),
            sort_digit_bits=int(sort_digit_bits), raster_batch=int(raster_batch),
            cameras_per_block=int(cameras_per_block), reuse_projection=int(bool(reuse_projection)),
            sort_tile_items=int(sort_tile_items), persistent_composite=int(persistent_composite),
            chunk_culling=int(chunk_bounds is not None), lookback_spin_limit=int(lookback_spin_limit),
            ordering=int(ordering), depth_bucket_count=cutoff_buckets, depth_bucket_group_size=cutoff_group,
            ray_evaluation=int(bool(ray_gaussian_evaluation)))
        event_array = None
        if stage_events is not None:
            if len(stage_events) != 5:
                raise ValueError("stage_events must hold 5 events")
            event_array = (ctypes.c_void_p * 5)(*[int(e.cuda_event) for e in stage_events])
            c_options.stage_events = ctypes.cast(event_array, ctypes.POINTER(ctypes.c_void_p))
            c_options.stage_event_count = 5
        if statistics is not None:
            if statistics.dtype != torch.int64 or statistics.numel() < 4 or statistics.device != device:
                raise ValueError("statistics must be an int64[4] tensor on the render device")
            c_options.statistics = statistics.data_ptr()
        selector = _I64(1)
        with torch.cuda.device(device):
            entry = self.lib.vgr_render_static_scratch if static_scratch else self.lib.vgr_render
            self._check(entry(ctypes.byref(c_scene), ctypes.byref(c_cameras), ctypes.byref(c_outputs),
                              ctypes.byref(c_workspace), ctypes.byref(c_options), _stream(device),
                              ctypes.byref(selector)),
                        "vgr_render_static_scratch" if static_scratch else "vgr_render")
        self.last_call = RenderCall(c_scene, c_cameras, c_outputs, c_workspace, c_options, int(selector.value),
                                    device, event_array)
        return int(selector.value)

    def render_backward(self, call: "RenderCall", grad_rgb: torch.Tensor, grad_depth: torch.Tensor | None,
                        grad_alpha: torch.Tensor | None, record_grads: torch.Tensor | None, *,
                        scene_grads: tuple[torch.Tensor, torch.Tensor, torch.Tensor, torch.Tensor] | None,
                        grad_viewmats: torch.Tensor | None = None,
                        camera_grads: torch.Tensor | None = None) -> None:
        """Backward of the checkpointed forward ``call`` (its tensors untouched since).

        ``scene_grads`` = (means [G,3], covariances [G,6], opacities [G],
        features [G,F]) float32, zero-filled by the caller and added to.
        ``grad_viewmats`` [B,4,4] float32 (zeroed, added to) with ``camera_grads``
        workspace [POSE_REPLICAS * active, 24] requests camera-pose gradients.  After a
        tile-front forward (``ordering=ORDERING_TILE_FRONT`` with checkpoints) the pass computes
        pose gradients only: ``record_grads`` and ``scene_grads`` are None there.
        """
        if call.outputs.final_transmittance is None:
            raise ValueError("render_backward needs a forward rendered with checkpoints")
        device = call.device
        for name, tensor in (("grad_rgb", grad_rgb), ("grad_depth", grad_depth), ("grad_alpha", grad_alpha),
                             ("record_grads", record_grads), ("grad_viewmats", grad_viewmats),
                             ("camera_grads", camera_grads)):
            if tensor is None:
                continue
            if tensor.dtype != torch.float32 or tensor.device != device or not tensor.is_contiguous():
                raise ValueError(f"{name} must be a contiguous float32 tensor on {device}")
        if int(call.options.ordering) == ORDERING_TILE_FRONT:
            if scene_grads is not None or grad_viewmats is None:
                raise ValueError("the tile-front backward computes camera-pose gradients only: pass grad_viewmats "
                                 "and no scene_grads")
        else:
            if call.workspace.slot_records is None:
                raise ValueError("render_backward needs a forward rendered with slot_records")
            if record_grads is None or record_grads.numel() < call.workspace.intersection_capacity * RECORD_GRAD_FLOATS:
                raise ValueError("record_grads must hold 10 floats per intersection slot")
        gradients = _Gradients(grad_rgb=_ptr(grad_rgb), grad_depth=_ptr(grad_depth), grad_alpha=_ptr(grad_alpha),
                               record_grads=_ptr(record_grads), sorted_selector=call.selector)
        if scene_grads is not None:
            means, covariances, opacities, features = scene_grads
            for tensor in scene_grads:
                if tensor.dtype != torch.float32 or tensor.device != device or not tensor.is_contiguous():
                    raise ValueError("scene gradients must be contiguous float32 tensors on the render device")
            if (means.numel() != 3 * call.scene.gaussian_count or covariances.numel() != 6 * call.scene.gaussian_count
                    or opacities.numel() != call.scene.gaussian_count
                    or features.numel() != call.scene.feature_width * call.scene.gaussian_count):
                raise ValueError("scene gradient shapes must match the scene tensors")
            gradients.grad_means = _ptr(means)
            gradients.grad_covariances = _ptr(covariances)
            gradients.grad_opacities = _ptr(opacities)
            gradients.grad_features = _ptr(features)
        if grad_viewmats is not None:
            if camera_grads is None or (camera_grads.numel() <
                                        POSE_REPLICAS * call.cameras.active_count * CAMERA_GRAD_FLOATS):
                raise ValueError("pose gradients need camera_grads workspace of POSE_REPLICAS x 24 floats per "
                                 "active camera")
            if grad_viewmats.numel() < 16 * call.cameras.batch:
                raise ValueError("grad_viewmats must be [B,4,4]")
            gradients.grad_viewmats = _ptr(grad_viewmats)
            gradients.camera_grads = _ptr(camera_grads)
        with torch.cuda.device(device):
            self._check(self.lib.vgr_render_backward(
                ctypes.byref(call.scene), ctypes.byref(call.cameras), ctypes.byref(call.outputs),
                ctypes.byref(call.workspace), ctypes.byref(call.options), ctypes.byref(gradients), _stream(device)),
                "vgr_render_backward")

    def covariance_backward(self, scales: torch.Tensor, rotations: torch.Tensor, grad_covariances: torch.Tensor,
                            grad_scales: torch.Tensor, grad_rotations: torch.Tensor) -> None:
        count = int(scales.shape[0])
        if count == 0:
            return
        with torch.cuda.device(scales.device):
            self._check(self.lib.vgr_covariance_backward(_ptr(scales), _ptr(rotations), _ptr(grad_covariances), count,
                                                         _ptr(grad_scales), _ptr(grad_rotations),
                                                         _stream(scales.device)), "vgr_covariance_backward")

    def pose_compose(self, twists: torch.Tensor, base_views: torch.Tensor, views: torch.Tensor) -> None:
        count = int(twists.shape[0])
        if count == 0:
            return
        with torch.cuda.device(twists.device):
            self._check(self.lib.vgr_pose_compose(_ptr(twists), _ptr(base_views), count, _ptr(views),
                                                  _stream(twists.device)), "vgr_pose_compose")

    def pose_compose_backward(self, twists: torch.Tensor, base_views: torch.Tensor, grad_views: torch.Tensor,
                              grad_twists: torch.Tensor, grad_base_views: torch.Tensor | None) -> None:
        count = int(twists.shape[0])
        if count == 0:
            return
        with torch.cuda.device(twists.device):
            self._check(self.lib.vgr_pose_compose_backward(
                _ptr(twists), _ptr(base_views), _ptr(grad_views), count, _ptr(grad_twists),
                _ptr(grad_base_views) if grad_base_views is not None else None, _stream(twists.device)),
                "vgr_pose_compose_backward")


class RenderCall:
    """The C structs of one vgr_render call, kept for its backward pass."""

    def __init__(self, scene, cameras, outputs, workspace, options, selector, device, event_array) -> None:
        self.scene = scene
        self.cameras = cameras
        self.outputs = outputs
        self.workspace = workspace
        self.options = options
        self.selector = selector
        self.device = device
        self.options.stage_events = None
        self.options.stage_event_count = 0
        self.options.statistics = None
        del event_array


def _require_dtypes(*items: tuple[torch.Tensor, torch.dtype, str]) -> None:
    for tensor, dtype, name in items:
        if tensor.dtype != dtype:
            raise ValueError(f"{name} has dtype {tensor.dtype}, expected {dtype}")


_MODULE: NativeRenderer | None = None


def load_native_module() -> NativeRenderer:
    """Load (once) and return the native renderer; raises if it is unavailable."""
    global _MODULE
    if _MODULE is None:
        _MODULE = NativeRenderer()
    return _MODULE


def float_bits(value: float) -> int:
    """IEEE-754 binary32 bit pattern of ``value``."""
    return struct.unpack("<I", struct.pack("<f", value))[0]
