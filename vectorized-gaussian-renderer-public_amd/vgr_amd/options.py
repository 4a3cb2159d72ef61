"""Constructor knobs of the gfx950 backend and their validation.

The knob names, defaults and meanings are those of the reference backend's
constructor (``src/isaacsim_gaussian_renderer/cuda_backend.py:94-234``,
Apache-2.0), so code written against it constructs this backend unchanged;
invalid combinations raise ``ValueError`` naming the knob.  Knobs that only
tune this build's kernels (``sh_degree``, ``sort_digit_bits``,
``raster_batch``, ``cameras_per_block``, ``sort_tile_items``,
``lookback_spin_limit``) default to "library default" (``None`` / 0).
"""

from __future__ import annotations

import math
from collections.abc import Callable
from dataclasses import dataclass
from typing import Literal

DEFAULT_GAUSSIAN_SUPPORT_SIGMA = 3.33  # gsplat's support, reference cuda_backend.py:31


@dataclass(frozen=True)
class BackendOptions:
    max_visible_records: int | None = None
    max_intersections: int | None = None
    visible_capacity_per_view: int = 100_000
    intersections_per_visible: float = 4.0
    near_plane: float = 0.01
    far_plane: float = 100.0
    gaussian_support_sigma: float = DEFAULT_GAUSSIAN_SUPPORT_SIGMA
    covariance_epsilon: float = 0.3
    rasterize_mode: Literal["classic", "antialiased"] = "classic"
    semantic_min_alpha: float = 0.01
    ray_gaussian_evaluation: bool = False
    tile_size: int = 16
    depth_bucket_count: int = 4096
    depth_bucket_group_size: int = 64
    compact_projection_cache: bool = False
    materialize_projected_records: bool = False
    enable_projection_cache: bool = False
    output_srgb: bool = True
    deterministic: bool = False
    fixed_capacity_sort: bool = False
    max_physical_views: int | None = None
    adaptive_capacity: bool = True
    capacity_growth_factor: float = 1.25
    capacity_headroom: float = 1.25
    max_capacity_retries: int = 4
    max_workspace_bytes: int | None = None
    sh_degree: int | None = None
    sort_digit_bits: int = 0
    raster_batch: int = 0
    cameras_per_block: int = 0
    sort_tile_items: int = 0
    lookback_spin_limit: int = 0

    def validate(self) -> "BackendOptions":
        for broken, message in _RULES:
            if broken(self):
                raise ValueError(message)
        return self

    @property
    def chunked(self) -> bool:
        return self.max_physical_views is not None


def _non_positive(value: float | None) -> bool:
    return value is not None and value <= 0


def _not_finite_or(value: float, low: float, *, strict: bool) -> bool:
    if not math.isfinite(value):
        return True
    return value <= low if strict else value < low


# (predicate that is True for an invalid configuration, message)
_RULES: tuple[tuple[Callable[[BackendOptions], bool], str], ...] = (
    (lambda o: _non_positive(o.max_visible_records), "max_visible_records must be positive."),
    (lambda o: _non_positive(o.max_intersections), "max_intersections must be positive."),
    (lambda o: o.visible_capacity_per_view <= 0, "visible_capacity_per_view must be positive."),
    (lambda o: o.intersections_per_visible <= 0, "intersections_per_visible must be positive."),
    (lambda o: not 0 < o.near_plane < o.far_plane, "Expected 0 < near_plane < far_plane."),
    (lambda o: not o.gaussian_support_sigma > 0, "gaussian_support_sigma must be positive."),
    (lambda o: _not_finite_or(o.covariance_epsilon, 0.0, strict=False),
     "covariance_epsilon must be finite and non-negative."),
    (lambda o: o.rasterize_mode not in ("classic", "antialiased"),
     "rasterize_mode must be 'classic' or 'antialiased'."),
    (lambda o: not (math.isfinite(o.semantic_min_alpha) and 0.0 <= o.semantic_min_alpha <= 1.0),
     "semantic_min_alpha must be finite and in [0, 1]."),
    (lambda o: not (1 <= o.tile_size <= 16 and o.tile_size & (o.tile_size - 1) == 0),
     "tile_size must be a power of two in [1, 16]."),
    (lambda o: o.depth_bucket_count <= 0, "depth_bucket_count must be positive."),
    (lambda o: o.depth_bucket_group_size <= 0, "depth_bucket_group_size must be positive."),
    (lambda o: o.compact_projection_cache and o.tile_size != 1, "compact_projection_cache requires tile_size=1."),
    (lambda o: o.compact_projection_cache and o.ray_gaussian_evaluation,
     "compact_projection_cache currently supports only the screen-space Gaussian model."),
    (lambda o: o.compact_projection_cache and o.deterministic,
     "compact_projection_cache currently requires deterministic=False."),
    (lambda o: o.materialize_projected_records and not o.compact_projection_cache,
     "materialize_projected_records requires compact_projection_cache=True."),
    (lambda o: o.rasterize_mode == "antialiased" and o.ray_gaussian_evaluation,
     "rasterize_mode='antialiased' is incompatible with ray_gaussian_evaluation=True because the exact-ray path "
     "does not use the screen-space covariance dilation being compensated."),
    (lambda o: o.ray_gaussian_evaluation and o.tile_size not in (8, 16),
     "ray_gaussian_evaluation renders 8x8 or 16x16 tiles on the gfx950 build."),
    (lambda o: o.fixed_capacity_sort and o.deterministic,
     "fixed_capacity_sort applies to the nondeterministic global-radix path only; it cannot be combined with "
     "deterministic=True."),
    (lambda o: _non_positive(o.max_physical_views), "max_physical_views must be positive when set."),
    (lambda o: o.chunked and not o.fixed_capacity_sort,
     "max_physical_views requires fixed_capacity_sort=True: logical chunks must never wait on an emitted count."),
    (lambda o: o.chunked and o.adaptive_capacity,
     "max_physical_views requires adaptive_capacity=False: physical capacities are calibrated and installed "
     "before steady-state logical-batch rendering."),
    (lambda o: o.chunked and o.enable_projection_cache,
     "max_physical_views requires enable_projection_cache=False."),
    (lambda o: _not_finite_or(o.capacity_growth_factor, 1.0, strict=True),
     "capacity_growth_factor must be finite and greater than 1."),
    (lambda o: _not_finite_or(o.capacity_headroom, 1.0, strict=False),
     "capacity_headroom must be finite and at least 1."),
    (lambda o: type(o.max_capacity_retries) is not int or o.max_capacity_retries < 0,
     "max_capacity_retries must be a non-negative integer."),
    (lambda o: _non_positive(o.max_workspace_bytes), "max_workspace_bytes must be positive when set."),
    (lambda o: o.sh_degree is not None and o.sh_degree not in (0, 1, 2, 3),
     "sh_degree must be in [0, 3] or None (inferred from the feature width)."),
    (lambda o: o.sort_digit_bits != 0 and not 4 <= o.sort_digit_bits <= 9,
     "sort_digit_bits must be 0 (library default) or in [4, 9]."),
    (lambda o: o.raster_batch not in (0, 64, 128, 256, 512),
     "raster_batch must be 0 (library default), 64, 128, 256 or 512."),
    (lambda o: not 0 <= o.cameras_per_block <= 64, "cameras_per_block must be 0 (library default) or in [1, 64]."),
    (lambda o: o.sort_tile_items not in (0, 8, 12, 16),
     "sort_tile_items must be 0 (library default), 8, 12 or 16."),
    (lambda o: o.lookback_spin_limit < -1,
     "lookback_spin_limit must be -1 (every wait gives up: tests only), 0 (library default) or positive."),
)
