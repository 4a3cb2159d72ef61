"""The oracle's exact-ray evaluation, pinned on the CPU by closed forms.

Set-up: isotropic Gaussians of standard deviation s on the optical axis at
camera-space (0, 0, z), identity cameras with focal length f whose principal
point is a pixel centre, so the pixel at offset (i, j) from it has the
normalized ray r = (i / f, j / f, 1).  For a point mu = (0, 0, z), the
squared distance of the ray's closest approach is
    d^2 = |mu|^2 - (r . mu)^2 / |r|^2 = z^2 rho^2 / (1 + rho^2),  rho^2 = (i^2 + j^2) / f^2,
and the Mahalanobis distance is m = d / s, so (renderer_cuda.cu:203-245)
    alpha = min(1, o exp(-m^2 / 2))  when m <= sigma (the support), else 0.
The support's screen ellipse is the circle rho <= sigma s / sqrt(z^2 - sigma^2 s^2),
i.e. a pixel radius R = ceil(f sigma s / sqrt(z^2 - sigma^2 s^2)); with 16x16
raster tiles the intersections are binned in 32x32 tiles (cuda_backend.py:1973-1985).
Compositing blends every entry with alpha > 0 and stops after the entry that
takes the transmittance to 1e-4 (renderer_cuda.cu:2262-2272), unlike the
screen model, which stops before it.
"""

from __future__ import annotations

import math

import numpy as np

from oracle import oracle
from vgr_amd import synthetic

F = 100.0
SIZE = 128


def _scene(depths, sigma3d, opacity):
    count = len(depths)
    return {
        "means": np.array([[0.0, 0.0, d] for d in depths], np.float32),
        "scales": np.full((count, 3), sigma3d, np.float32),
        "rotations": np.array([[1.0, 0.0, 0.0, 0.0]] * count, np.float32),
        "opacities": np.full(count, opacity, np.float32),
        "features": np.array([[0.2 + 0.1 * k, 0.5, 0.7] for k in range(count)], np.float32),
        "semantic_ids": np.arange(10, 10 + count, dtype=np.int64),
    }


def _render(scene, *, principal, tile_size=16, sigma=3.33):
    view = np.eye(4, dtype=np.float32)[None]
    k = np.array([[[F, 0.0, principal[0]], [0.0, F, principal[1]], [0.0, 0.0, 1.0]]], np.float32)
    return oracle.render([scene], [0], view, k, [0], height=SIZE, width=SIZE, tile_size=tile_size,
                         gaussian_support_sigma=sigma, ray_evaluation=True)


def test_single_splat_alpha_follows_the_ray_distance():
    z, s, o = 10.0, 0.3, 0.8
    result = _render(_scene([z], s, o), principal=(64.5, 64.5))
    alpha = result["alpha"][0, :, :, 0]
    # m^2 is formed in float32 as mu^T P mu - (r^T P mu)^2 / (r^T P r), two terms of
    # about z^2 / s^2 = 1111: a few of their ulps (1.2e-4 each) of cancellation error
    # in m^2, so alpha is good to 0.5 * 5e-4 * o = 2e-4 of the closed form.
    for i, j in ((0, 0), (3, 0), (0, -5), (4, 4), (7, 2)):
        rho2 = (i * i + j * j) / F ** 2
        m2 = z * z * rho2 / (1.0 + rho2) / (s * s)
        expected = min(1.0, o * math.exp(-0.5 * m2)) if m2 <= 3.33 ** 2 else 0.0
        assert abs(float(alpha[64 + j, 64 + i]) - expected) < 2e-4, (i, j)
    # Depth of a single splat is its camera-space depth wherever it blends.
    assert abs(float(result["depth"][0, 64, 64, 0]) - z) < 1e-5


def test_support_ellipse_sets_the_bins():
    # R = ceil(100 * 3.33 * 0.3 / sqrt(100 - 0.999^2)) = ceil(10.04) = 11 pixels around
    # the principal point 64.5: columns [floor(53.5 / 32), ceil(75.5 / 32)) = [1, 3) of 4 bins.
    z, s = 10.0, 0.3
    radius = math.ceil(F * 3.33 * s / math.sqrt(z * z - (3.33 * s) ** 2))
    assert radius == 11
    result = _render(_scene([z], s, 0.8), principal=(64.5, 64.5))
    counts = result["tile_counts"].reshape(4, 4)
    expected = np.zeros((4, 4), np.int32)
    expected[1:3, 1:3] = 1
    np.testing.assert_array_equal(counts, expected)
    assert int(result["counters"][1]) == 4 and int(result["counters"][4]) == 4
    # Beyond the support the alpha is zero even inside the bins.
    rho2 = (12 * 12) / F ** 2
    assert z * z * rho2 / (1.0 + rho2) / (s * s) > 3.33 ** 2
    assert float(result["alpha"][0, 64, 64 + 12, 0]) == 0.0


def test_opaque_splat_stops_after_blending():
    # Opacity 1 at m = 0 gives alpha 1: the first splat is blended in full and the
    # pixel stops (the screen model would have skipped it: 1 - 0.99 never reaches 0).
    scene = _scene([5.0, 8.0], 0.3, 1.0)
    result = _render(scene, principal=(64.5, 64.5))
    assert float(result["alpha"][0, 64, 64, 0]) == 1.0
    assert abs(float(result["depth"][0, 64, 64, 0]) - 5.0) < 1e-6
    np.testing.assert_allclose(result["rgb"][0, 64, 64], [0.2, 0.5, 0.7], atol=1e-6)
    assert int(result["semantic_id"][0, 64, 64, 0]) == 10


def test_images_do_not_depend_on_the_bin_size():
    # Only pixels within rounding of a support boundary (ellipse in double, support test in float)
    # can depend on the bin size; this scene has none.
    scene = synthetic.gaussian_scene(8000, seed=9, sh_degree=1)
    rig = synthetic.orbit_rig(2, 96, 72, 83.0, seed=9, radius=(9.0, 11.0))
    kwargs = dict(height=72, width=96, sh_degree=1, ray_evaluation=True)
    sixteen = oracle.render([scene.as_kwargs()], [0], rig.viewmats, rig.intrinsics, [0, 0], tile_size=16, **kwargs)
    eight = oracle.render([scene.as_kwargs()], [0], rig.viewmats, rig.intrinsics, [0, 0], tile_size=8, **kwargs)
    assert sixteen["tile_counts"].shape[0] == 2 * 3 * 3 and eight["tile_counts"].shape[0] == 2 * 12 * 9
    for name in ("rgb", "depth", "alpha", "semantic_id"):
        assert np.array_equal(sixteen[name], eight[name]), name
    screen = oracle.render([scene.as_kwargs()], [0], rig.viewmats, rig.intrinsics, [0, 0], height=72, width=96,
                           sh_degree=1, tile_size=16)
    assert not np.array_equal(screen["alpha"], sixteen["alpha"])  # a different model
