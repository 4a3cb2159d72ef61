"""Closed-form known answers for the oracle's tile emission.

Pins the integer outputs the GPU path must reproduce bit for bit (per-tile
intersection counts, the visible and intersection counters, the active-tile
counter) independently of the oracle's own code: every expected radius, tile
range and culling decision below was worked out by hand from the reference's
rules and is written as a literal.

Reference rules (paths relative to the reference repository):
  * radius: r = ceil(min(sigma, sqrt(2 ln(o / (1/255)))) * sqrt(cov2d_ii))
    (``renderer_cuda.cu:769-781``);
  * screen cull: mean_x + r_x <= 0, mean_y + r_y <= 0, mean_x - r_x >= W or
    mean_y - r_y >= H drops the splat (``renderer_cuda.cu:950-956``);
  * tile rectangle: [max(0, floor(m/ts - r/ts)), min(tiles, ceil(m/ts + r/ts)))
    per axis, row-major tiles (``renderer_cuda.cu:1729-1748``).

Set-up shared by every case: one Gaussian at camera-space (0, 0, 10) with an
identity rotation, seen by identity-view cameras with f = 100 whose principal
point places the splat; 64x64 images, 16x16 tiles (4x4 per camera),
covariance epsilon 0.3.  On the optical axis the EWA Jacobian is diag(f/Z) with
no cross term, so cov2d = diag((f s_x / Z)^2, (f s_y / Z)^2) + 0.3:
  s = 0.2 -> 4 + 0.3 = 4.3,   sqrt = 2.073644
  s = 0.5 -> 25 + 0.3 = 25.3, sqrt = 5.029911
  o = 0.9  -> sqrt(2 ln 229.5) = sqrt(10.871806) = 3.297242 (< 3.33)
  o = 0.05 -> sqrt(2 ln 12.75) = sqrt(5.091062)  = 2.256338
so
  (s 0.2, o 0.9, sigma 3.33): r = ceil(6.837307) = 7
  (s 0.5, o 0.9, sigma 3.33): r = ceil(16.584833) = 17
  (s 0.2, o 0.9, sigma 2.0):  r = ceil(4.147288) = 5
  (s 0.5, o 0.9, sigma 2.0):  r = ceil(10.059821) = 11
  (s 0.2, o 0.05, sigma 3.33): r = ceil(4.678843) = 5
  (s 0.5, o 0.05, sigma 3.33): r = ceil(11.349180) = 12
Principal points sit half a pixel either side of a tile boundary minus the
radius, so a radius one too large or one too small (or the two axes swapped)
changes the tile set.  All divisions by 16 below are exact in binary32.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle

WIDTH = HEIGHT = 64
TILES_X = TILES_Y = 4


def _render(points, *, scale_x, scale_y, opacity, sigma):
    scene = {
        "means": np.array([[0.0, 0.0, 10.0]], np.float32),
        "scales": np.array([[scale_x, scale_y, 0.1]], np.float32),
        "rotations": np.array([[1.0, 0.0, 0.0, 0.0]], np.float32),
        "opacities": np.array([opacity], np.float32),
        "features": np.array([[0.5, 0.5, 0.5]], np.float32),
        "semantic_ids": np.array([3], np.int64),
    }
    batch = len(points)
    view = np.tile(np.eye(4, dtype=np.float32), (batch, 1, 1))
    k = np.zeros((batch, 3, 3), np.float32)
    k[:, 0, 0] = k[:, 1, 1] = 100.0
    k[:, 2, 2] = 1.0
    for camera, (cx, cy) in enumerate(points):
        k[camera, 0, 2] = cx
        k[camera, 1, 2] = cy
    return oracle.render([scene], [0], view, k, [0] * batch, height=HEIGHT, width=WIDTH,
                         gaussian_support_sigma=sigma, covariance_epsilon=0.3, tile_size=16)


def _expected_tiles(camera, x_range, y_range):
    """Global tile ids of the literal per-axis tile ranges (row-major)."""
    return [camera * TILES_X * TILES_Y + ty * TILES_X + tx for ty in range(*y_range) for tx in range(*x_range)]


def _check(result, cases):
    expected = np.zeros(len(cases) * TILES_X * TILES_Y, np.int32)
    visible = 0
    for camera, (_point, x_range, y_range) in enumerate(cases):
        if x_range is None:
            continue
        visible += 1
        for tile in _expected_tiles(camera, x_range, y_range):
            expected[tile] += 1
    np.testing.assert_array_equal(result["tile_counts"], expected)
    counters = result["counters"]
    assert int(counters[0]) == visible
    assert int(counters[1]) == int(expected.sum())
    assert int(counters[4]) == int(np.count_nonzero(expected))
    assert int(counters[2]) == 0 and int(counters[3]) == 0


# (principal point, x tile range, y tile range); None = culled.
@pytest.mark.parametrize(("scale_x", "scale_y", "opacity", "sigma", "cases"), [
    # r = (7, 7).  x: 40.5 -> [floor(33.5/16), ceil(47.5/16)) = [2, 3)   (r = 8 would reach 48.5 -> 4)
    #              y: 41.5 -> [floor(34.5/16), ceil(48.5/16)) = [2, 4)   (r = 6 would stop at 47.5 -> 3)
    # tile corner (32, 32): [floor(25/16), ceil(39/16)) = [1, 3) on both axes.
    pytest.param(0.2, 0.2, 0.9, 3.33, [((40.5, 41.5), (2, 3), (2, 4)),
                                      ((32.0, 32.0), (1, 3), (1, 3))], id="isotropic-r7"),
    # r = (7, 17).  (41.5, 31.5): x [2, 4); y [floor(14.5/16), ceil(48.5/16)) = [0, 4)
    #               (40.5, 30.5): x [2, 3); y [floor(13.5/16), ceil(47.5/16)) = [0, 3)
    # Swapped axes would give x [1, 4), y [1, 3) at the first point.
    pytest.param(0.2, 0.5, 0.9, 3.33, [((41.5, 31.5), (2, 4), (0, 4)),
                                      ((40.5, 30.5), (2, 3), (0, 3))], id="anisotropic-r7-r17"),
    # sigma branch, r = (5, 11).  (43.5, 37.5): x [floor(38.5/16), ceil(48.5/16)) = [2, 4);
    #                              y [floor(26.5/16), ceil(48.5/16)) = [1, 4)
    #                             (42.5, 36.5): x [2, 3); y [floor(25.5/16), ceil(47.5/16)) = [1, 3)
    pytest.param(0.2, 0.5, 0.9, 2.0, [((43.5, 37.5), (2, 4), (1, 4)),
                                     ((42.5, 36.5), (2, 3), (1, 3))], id="support-sigma-r5-r11"),
    # opacity branch, r = (5, 12).  (43.5, 36.5): x [2, 4); y [floor(24.5/16), ceil(48.5/16)) = [1, 4)
    #                                (42.5, 35.5): x [2, 3); y [floor(23.5/16), ceil(47.5/16)) = [1, 3)
    pytest.param(0.2, 0.5, 0.05, 3.33, [((43.5, 36.5), (2, 4), (1, 4)),
                                       ((42.5, 35.5), (2, 3), (1, 3))], id="opacity-extent-r5-r12"),
    # image borders, r = (7, 7):
    #   (-2.5, 3.5):  x [max(0, floor(-9.5/16)), ceil(4.5/16)) = [0, 1); y [0, ceil(10.5/16)) = [0, 1)
    #   (-6.75, 20.5): mean_x + r = 0.25 > 0 keeps it; x [0, 1); y [floor(13.5/16), ceil(27.5/16)) = [0, 2)
    #   (-7.0, 20.5): mean_x + r = 0 <= 0 culls it
    #   (70.75, 60.5): mean_x - r = 63.75 < 64 keeps it; x [3, min(4, ceil(77.75/16))) = [3, 4);
    #                  y [floor(53.5/16), min(4, ceil(67.5/16))) = [3, 4)
    #   (71.0, 30.5): mean_x - r = 64 >= 64 culls it
    #   (20.5, -6.5): mean_y + r = 0.5 > 0 keeps it; x [floor(13.5/16), ceil(27.5/16)) = [0, 2); y [0, 1)
    pytest.param(0.2, 0.2, 0.9, 3.33, [((-2.5, 3.5), (0, 1), (0, 1)),
                                      ((-6.75, 20.5), (0, 1), (0, 2)),
                                      ((-7.0, 20.5), None, None),
                                      ((70.75, 60.5), (3, 4), (3, 4)),
                                      ((71.0, 30.5), None, None),
                                      ((20.5, -6.5), (0, 2), (0, 1))], id="image-borders"),
])
def test_single_splat_emission_known_answers(scale_x, scale_y, opacity, sigma, cases):
    result = _render([point for point, _x, _y in cases], scale_x=scale_x, scale_y=scale_y, opacity=opacity,
                     sigma=sigma)
    _check(result, cases)


def test_overlapping_splats_add_per_tile_counts():
    """Two splats sharing tiles: per-tile counts add (2 where the rectangles overlap)."""
    scene = {
        "means": np.array([[0.0, 0.0, 10.0], [0.4, 0.0, 20.0]], np.float32),
        "scales": np.array([[0.2, 0.2, 0.1], [0.4, 0.4, 0.1]], np.float32),
        "rotations": np.array([[1.0, 0.0, 0.0, 0.0]] * 2, np.float32),
        "opacities": np.array([0.9, 0.9], np.float32),
        "features": np.full((2, 3), 0.5, np.float32),
        "semantic_ids": np.array([3, 4], np.int64),
    }
    view = np.eye(4, dtype=np.float32)[None]
    k = np.array([[[100.0, 0.0, 32.0], [0.0, 100.0, 32.0], [0.0, 0.0, 1.0]]], np.float32)
    result = oracle.render([scene], [0], view, k, [0], height=HEIGHT, width=WIDTH, tile_size=16)
    # Splat 0: mean (32, 32), r = 7 -> tiles x [1, 3), y [1, 3).
    # Splat 1: camera-space (0.4, 0, 20): mean_x = 100 * 0.4 / 20 + 32 = 34, mean_y = 32.  Off axis the
    # Jacobian gains j02 = -f x / z^2 = -0.1 next to j00 = f / z = 5, so with s = (0.4, 0.4, 0.1)
    # cov2d_00 = 25 * 0.16 + 0.01 * 0.01 + 0.3 = 4.3001 and cov2d_11 = 25 * 0.16 + 0.3 = 4.3;
    # r_x = ceil(3.297242 * 2.073668) = ceil(6.837386) = 7, r_y = 7
    # -> x [floor(27/16), ceil(41/16)) = [1, 3), y [1, 3): the same four tiles as splat 0.
    expected = np.zeros(TILES_X * TILES_Y, np.int32)
    for tile in _expected_tiles(0, (1, 3), (1, 3)):
        expected[tile] += 2
    np.testing.assert_array_equal(result["tile_counts"], expected)
    assert int(result["counters"][0]) == 2 and int(result["counters"][1]) == 8
