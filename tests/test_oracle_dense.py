"""The oracle's dense per-pixel path (tile size 1) and small tiles, pinned on the CPU.

* Known answer, per-pixel emission: one isotropic splat (screen variance
  25.3 px^2 = (100 * 0.5 / 10)^2 + 0.3, opacity 0.9) centred on a pixel corner
  emits exactly the pixels whose centres lie within the alpha = 1/255 circle,
  d^2 <= 2 * 25.3 * ln(0.9 * 255) = 275.07 (its radius-17 rectangle does not
  clip it); the count is the lattice-point count of that disc, worked out here
  in float64 with no point within 0.05 of the boundary.
* Known answer, depth cutoff: six opaque copies of a splat at depths 2..7 give
  the four pixels around its centre alpha 0.99 each; the conservative
  thickness of alpha 0.99 is 0.999 * sum_{k<=8} 0.99^k / k = 2.6366, so three
  splats (7.91) stay below the 9.2103 threshold and four (10.55) pass it: those
  pixels keep exactly four intersections (buckets behind the fourth are cut).
* Invariance the mathematics fixes: per-pixel compositing of depth-sorted
  contributions does not depend on the tile size, and the cutoff only removes
  entries behind a transmittance of 1e-4, where compositing has already
  stopped; so images at tile sizes 1, 2, 4, 8, 16, with and without cutoffs of
  several bucket counts, are bitwise equal.

Reference rules: ``renderer_cuda.cu:168-200`` (bucket index, pixel alpha),
``247-257`` (conservative thickness), ``1658-1690`` (cutoff advance),
``1764-1805`` (tile size 1 emission).
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle
from vgr_amd import synthetic

SIZE = 64


def _splats(depths, *, scale, opacity):
    count = len(depths)
    return {
        "means": np.array([[0.0, 0.0, d] for d in depths], np.float32),
        # Scale grows with depth so every copy has the same screen-space footprint.
        "scales": np.array([[scale * d / 10.0, scale * d / 10.0, 0.1] for d in depths], np.float32),
        "rotations": np.array([[1.0, 0.0, 0.0, 0.0]] * count, np.float32),
        "opacities": np.full(count, opacity, np.float32),
        "features": np.full((count, 3), 0.5, np.float32),
        "semantic_ids": np.arange(count, dtype=np.int64),
    }


def _render_one(scene, **kwargs):
    view = np.eye(4, dtype=np.float32)[None]
    k = np.array([[[100.0, 0.0, 32.0], [0.0, 100.0, 32.0], [0.0, 0.0, 1.0]]], np.float32)
    return oracle.render([scene], [0], view, k, [0], height=SIZE, width=SIZE, covariance_epsilon=0.3,
                         tile_size=1, **kwargs)


def test_single_splat_emits_the_pixels_of_its_alpha_disc():
    result = _render_one(_splats([10.0], scale=0.5, opacity=0.9))
    radius2 = 2.0 * 25.3 * math.log(0.9 * 255.0)
    centres = np.arange(SIZE) + 0.5 - 32.0
    d2 = centres[None, :] ** 2 + centres[:, None] ** 2
    assert np.abs(d2 - radius2).min() > 0.05  # no pixel centre near the boundary
    expected = (d2 <= radius2).astype(np.int32).reshape(-1)
    assert int(expected.sum()) == 872
    assert np.array_equal(result["tile_counts"], expected)
    assert int(result["counters"][1]) == 872 and int(result["counters"][0]) == 1


def test_depth_cutoff_keeps_four_opaque_splats_at_the_centre():
    scene = _splats([2.0, 3.0, 4.0, 5.0, 6.0, 7.0], scale=0.5, opacity=1.0)
    full = _render_one(scene)
    cut = _render_one(scene, depth_bucket_count=4096, depth_bucket_group_size=64)
    centre = [31 * SIZE + 31, 31 * SIZE + 32, 32 * SIZE + 31, 32 * SIZE + 32]
    assert full["tile_counts"][centre].tolist() == [6, 6, 6, 6]
    assert cut["tile_counts"][centre].tolist() == [4, 4, 4, 4]
    assert np.all(cut["tile_counts"] <= full["tile_counts"])
    assert int(cut["counters"][1]) == int(cut["tile_counts"].sum()) < int(full["counters"][1])
    for name in ("rgb", "depth", "alpha", "semantic_id"):
        assert np.array_equal(cut[name], full[name]), name


@pytest.mark.parametrize("buckets,group", [(1, 1), (64, 7), (4096, 64), (4096, 4096)])
def test_images_do_not_depend_on_tile_size_or_cutoff(buckets, group):
    scene = synthetic.gaussian_scene(20000, seed=5, sh_degree=1)
    rig = synthetic.orbit_rig(2, 48, 40, 40.0, seed=5, radius=(9.0, 11.0))

    def run(tile_size, **kwargs):
        return oracle.render([scene.as_kwargs()], [0], rig.viewmats, rig.intrinsics, [0, 0], height=40, width=48,
                             tile_size=tile_size, sh_degree=1, **kwargs)

    base = run(16)
    dense = run(1)
    cut = run(1, depth_bucket_count=buckets, depth_bucket_group_size=group)
    assert int(cut["counters"][1]) <= int(dense["counters"][1])
    for other in (run(8), run(4), run(2), dense, cut):
        for name in ("rgb", "depth", "alpha", "semantic_id"):
            assert np.array_equal(other[name], base[name]), name
