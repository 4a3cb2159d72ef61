"""Pins the CPU oracle to the reference's known answers and to the mathematics.

Each check is chosen so that a plausible restatement mistake (a dropped term,
a wrong sign, a transposed covariance, an off-by-one tile bound, accumulating
before the transmittance cut-off) fails it:
  * elementary functions against float64 numpy;
  * covariance against R diag(s^2) R^T built from scipy rotations;
  * closed-form projections of on-axis Gaussians (centre alpha, neighbour
    alpha, depth, semantic) and the Mip-Splatting compensation known answers
    of tests/test_antialias_benchmark.py:24-47;
  * two-layer compositing and the break-before-accumulate rule
    (renderer_cuda.cu:2245-2252);
  * the strongest-contribution semantic known answer [10, -1, 4] of
    tests/test_native_exact_semantics_diagnostic.py:15-87;
  * a tile-free float64 brute-force renderer (every Gaussian, every pixel);
  * the real-SH basis against scipy.special.sph_harm_y;
  * ordering invariants and thread-count independence.
"""

from __future__ import annotations

import math

import numpy as np
import pytest
import scipy.special
from scipy.spatial.transform import Rotation

from oracle import oracle
from vgr_amd import synthetic


def scene_dict(means, scales, rotations, opacities, features, semantic_ids):
    return {
        "means": np.asarray(means, np.float32).reshape(-1, 3),
        "scales": np.asarray(scales, np.float32).reshape(-1, 3),
        "rotations": np.asarray(rotations, np.float32).reshape(-1, 4),
        "opacities": np.asarray(opacities, np.float32).reshape(-1),
        "features": np.asarray(features, np.float32).reshape(len(np.asarray(opacities).reshape(-1)), -1),
        "semantic_ids": np.asarray(semantic_ids, np.int64).reshape(-1),
    }


def camera_identity(width, height, focal, cx, cy, batch=1):
    view = np.tile(np.eye(4, dtype=np.float32), (batch, 1, 1))
    k = np.zeros((batch, 3, 3), np.float32)
    k[:, 0, 0] = focal
    k[:, 1, 1] = focal
    k[:, 0, 2] = cx
    k[:, 1, 2] = cy
    k[:, 2, 2] = 1
    return view, k


# ---------------------------------------------------------------- elementary functions

def test_expf_relative_error_bound():
    """2^f polynomial error (1.7e-7) plus rounding of x*log2(e) (|x| * 6e-8 relative)."""
    xs = np.linspace(-20.0, 0.0, 400_001).astype(np.float32)[::3]
    values = np.array([oracle.expf(float(x)) for x in xs], dtype=np.float64)
    rel = np.abs(values / np.exp(xs.astype(np.float64)) - 1.0)
    assert rel.max() < 1.6e-6
    assert oracle.expf(0.0) == 1.0
    # Every alpha that can pass the 1/255 threshold has power > -5.55.
    assert rel[xs > -5.6].max() < 6.0e-7
    wide = np.linspace(-87.0, 88.0, 20_001).astype(np.float32)
    wide_values = np.array([oracle.expf(float(x)) for x in wide], dtype=np.float64)
    wide_rel = np.abs(wide_values / np.exp(wide.astype(np.float64)) - 1.0)
    # argument rounding (2^-24 |x|) plus the binary32 rounding of log2(e) (1.3e-8 |x|)
    assert np.all(wide_rel < 1.7e-7 + np.abs(wide) * 7.5e-8)


def test_expf_nonpositive_is_expf_bit_for_bit():
    """The compositor's variant drops only the upper clamp, a no-op for x <= 0 and NaN."""
    negatives = np.arange(0x80000000, 0xC3000000, 4099, dtype=np.uint64).astype(np.uint32).view(np.float32)
    specials = np.array([0.0, -0.0, -87.33, -87.34, -88.0, -150.0, -1.0e30, -np.inf, np.nan], dtype=np.float32)
    for x in np.concatenate([negatives, specials]):
        a = np.float32(oracle.expf(float(x)))
        b = np.float32(oracle.expf_nonpositive(float(x)))
        assert a.view(np.uint32) == b.view(np.uint32) or (np.isnan(a) and np.isnan(b)), (float(x), a, b)


def test_logf_error_bound():
    xs = np.concatenate([np.geomspace(1.0, 255.0, 50_001), np.geomspace(1e-30, 1e30, 20_001)]).astype(np.float32)
    values = np.array([oracle.logf(float(x)) for x in xs[::5]], dtype=np.float64)
    exact = np.log(xs[::5].astype(np.float64))
    error = np.abs(values - exact)
    assert np.max(error / np.maximum(np.abs(exact), 1.0)) < 4.0e-7


def test_srgb_matches_iec_61966_2_1():
    xs = np.linspace(-0.1, 1.1, 12_001)
    values = np.array([oracle.srgb(float(x)) for x in xs])
    c = np.clip(xs, 0.0, 1.0)
    exact = np.where(c <= 0.0031308, 12.92 * c, 1.055 * np.power(c, 1 / 2.4) - 0.055)
    assert np.max(np.abs(values - exact)) < 2.0e-6


# ------------------------------------------------------------------------ covariance

def test_covariance_identity_quaternion_is_diagonal_scale_squared():
    cov = oracle.covariances([[0.1, 0.2, 0.3]], [[1.0, 0.0, 0.0, 0.0]])
    np.testing.assert_allclose(cov[0], [0.01, 0.0, 0.0, 0.04, 0.0, 0.09], rtol=1e-6, atol=1e-9)


def test_covariance_matches_scipy_rotation_and_ignores_quaternion_norm():
    rng = np.random.default_rng(3)
    quats = rng.normal(size=(500, 4))
    scales = np.exp(rng.normal(-2.0, 0.7, size=(500, 3)))
    cov = oracle.covariances(scales, quats * 3.7)  # unnormalized on purpose
    rot = Rotation.from_quat(quats[:, [1, 2, 3, 0]]).as_matrix()  # scipy is xyzw
    full = rot @ (scales[:, :, None] ** 2 * np.transpose(rot, (0, 2, 1)))
    expected = full[:, [0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]]
    np.testing.assert_allclose(cov, expected, rtol=2e-5, atol=1e-7 * np.max(scales) ** 2)


# ---------------------------------------------------------------- closed-form projection

def test_on_axis_gaussian_closed_form():
    z, focal, s, opacity = 5.0, 50.0, 0.02, 0.8
    scene = scene_dict([[0, 0, z]], [[s, s, s]], [[1, 0, 0, 0]], [opacity], [[0.2, 0.4, 0.6]], [17])
    view, k = camera_identity(64, 64, focal, 31.5, 31.5)
    r = oracle.render([scene], [0], view, k, [0], height=64, width=64)
    assert r["counters"][0] == 1
    # mean2d lands on the centre of pixel (31, 31): power 0, alpha = min(0.99, o).
    assert r["alpha"][0, 31, 31, 0] == pytest.approx(opacity, abs=1e-7)
    assert r["depth"][0, 31, 31, 0] == pytest.approx(z, rel=1e-6)
    assert r["semantic_id"][0, 31, 31, 0] == 17
    np.testing.assert_allclose(r["rgb"][0, 31, 31], opacity * np.array([0.2, 0.4, 0.6]), atol=1e-6)
    # One pixel to the right: conic = 1 / ((f/z)^2 s^2 + eps).
    conic = 1.0 / ((focal / z) ** 2 * s * s + 0.3)
    assert r["alpha"][0, 31, 32, 0] == pytest.approx(opacity * math.exp(-0.5 * conic), rel=2e-6)
    # Diagonal neighbour: both axes contribute equally (no cross term for an isotropic splat).
    assert r["alpha"][0, 32, 32, 0] == pytest.approx(opacity * math.exp(-conic), rel=2e-6)


@pytest.mark.parametrize(("scale_x", "scale_y", "expected_alpha"), [
    (0.2, 0.4, 0.8 * (0.04 * 0.16 / (0.34 * 0.46)) ** 0.5),  # tests/test_antialias_benchmark.py:24-36
    (0.0, 0.0, 0.8 * 0.005),                                  # degenerate floor, :39-47
])
def test_mip_splatting_compensation_known_answers(scale_x, scale_y, expected_alpha):
    # focal / z = 1, so the raw screen covariance is diag(scale_x^2, scale_y^2).
    scene = scene_dict([[0, 0, 5.0]], [[scale_x, scale_y, 0.1]], [[1, 0, 0, 0]], [0.8], [[1, 1, 1]], [3])
    view, k = camera_identity(32, 32, 5.0, 15.5, 15.5)
    r = oracle.render([scene], [0], view, k, [0], height=32, width=32, antialiased=True)
    assert r["alpha"][0, 15, 15, 0] == pytest.approx(expected_alpha, rel=2e-6)


@pytest.mark.parametrize(("mean", "opacity", "visible"), [
    ([0, 0, 5.0], 0.5, 1),
    ([0, 0, 5.0], 1.0 / 256.0, 0),     # below 1/255
    ([0, 0, 0.005], 0.5, 0),           # in front of the near plane
    ([0, 0, 150.0], 0.5, 0),           # beyond the far plane
    ([60.0, 0, 5.0], 0.5, 0),          # bounding box off screen
    ([0, 0, -5.0], 0.5, 0),            # behind the camera
])
def test_cull_cases(mean, opacity, visible):
    scene = scene_dict([mean], [[0.05, 0.05, 0.05]], [[1, 0, 0, 0]], [opacity], [[1, 1, 1]], [1])
    view, k = camera_identity(64, 48, 40.0, 32.0, 24.0)
    r = oracle.render([scene], [0], view, k, [0], height=48, width=64)
    assert r["counters"][0] == visible
    if not visible:
        assert r["counters"][1] == 0 and np.all(r["alpha"] == 0) and np.all(r["semantic_id"] == -1)
        assert np.all(np.isinf(r["depth"]))


# ------------------------------------------------------------------------ compositing

def _stack_at_pixel(alphas, depths, colours, ids, *, pixel=(8.5, 8.5)):
    n = len(alphas)
    means = [[0.0, 0.0, d] for d in depths]
    scene = scene_dict(means, [[1e-4] * 3] * n, [[1, 0, 0, 0]] * n, alphas, colours, ids)
    view, k = camera_identity(16, 16, 10.0, pixel[0], pixel[1])
    return oracle.render([scene], [0], view, k, [0], height=16, width=16)


def test_two_layer_closed_form():
    a1, a2, z1, z2 = 0.6, 0.5, 2.0, 3.0
    r = _stack_at_pixel([a2, a1], [z2, z1], [[0, 0, 1], [1, 0, 0]], [22, 11])  # stored back-to-front
    alpha = 1 - (1 - a1) * (1 - a2)
    assert r["alpha"][0, 8, 8, 0] == pytest.approx(alpha, rel=1e-6)
    np.testing.assert_allclose(r["rgb"][0, 8, 8], [a1, 0.0, (1 - a1) * a2], rtol=1e-6, atol=1e-7)
    assert r["depth"][0, 8, 8, 0] == pytest.approx((a1 * z1 + (1 - a1) * a2 * z2) / alpha, rel=1e-6)
    assert r["semantic_id"][0, 8, 8, 0] == 11


def test_transmittance_cutoff_breaks_before_accumulating():
    # alpha is clamped to 0.99: T goes 1 -> 0.01, and the second layer would
    # take it to 1e-4 <= threshold, so it is not accumulated at all.
    r = _stack_at_pixel([1.0, 1.0, 1.0], [1.0, 2.0, 3.0], [[1, 0, 0], [0, 1, 0], [0, 0, 1]], [1, 2, 3])
    assert r["alpha"][0, 8, 8, 0] == pytest.approx(0.99, abs=1e-7)
    np.testing.assert_allclose(r["rgb"][0, 8, 8], [0.99, 0.0, 0.0], atol=1e-7)
    assert r["depth"][0, 8, 8, 0] == pytest.approx(1.0, rel=1e-6)


def test_strongest_contribution_semantic_known_answer():
    """Pixels with alphas [.35,.30,.40] / none / [.70,.50] give semantics [10, -1, 4]."""
    groups = {0: ([0.35, 0.30, 0.40], [10, 20, 20]), 2: ([0.70, 0.50], [4, 3])}
    means, opac, ids = [], [], []
    for slot, (alphas, sems) in groups.items():
        for order, (a, sem) in enumerate(zip(alphas, sems)):
            # pixel centres 8 + 20*slot + 0.5 on a 64x16 image, focal 10 at z = 1 + order
            z = 1.0 + order
            x_pixel = 8.5 + 20 * slot
            means.append([(x_pixel - 32.0) * z / 10.0, 0.0, z])
            opac.append(a)
            ids.append(sem)
    n = len(opac)
    scene = scene_dict(means, [[1e-4] * 3] * n, [[1, 0, 0, 0]] * n, opac, [[1, 1, 1]] * n, ids)
    view, k = camera_identity(64, 16, 10.0, 32.0, 8.5)
    r = oracle.render([scene], [0], view, k, [0], height=16, width=64, semantic_min_alpha=0.0)
    assert [int(r["semantic_id"][0, 8, 8 + 20 * s, 0]) for s in range(3)] == [10, -1, 4]
    assert np.isinf(r["depth"][0, 8, 28, 0])


# ------------------------------------------------------- float64 tile-free brute force

def brute_force_render(scene, view, k, width, height, eps=0.3, sigma=3.33):
    """Every Gaussian at every pixel, float64, matrix-form EWA projection, no tiles."""
    means = scene["means"].astype(np.float64)
    cov3 = oracle.covariances(scene["scales"], scene["rotations"]).astype(np.float64)
    full = np.zeros((len(means), 3, 3))
    iu = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    for c, (i, j) in enumerate(iu):
        full[:, i, j] = cov3[:, c]
        full[:, j, i] = cov3[:, c]
    rot, trans = view[:3, :3].astype(np.float64), view[:3, 3].astype(np.float64)
    p = means @ rot.T + trans
    fx, fy, cx, cy = (float(k[0, 0]), float(k[1, 1]), float(k[0, 2]), float(k[1, 2]))
    out_rgb = np.zeros((height, width, 3))
    out_alpha = np.zeros((height, width))
    out_sem = np.full((height, width), -1)
    lx = (width - cx) / fx + 0.3 * 0.5 * width / fx, cx / fx + 0.3 * 0.5 * width / fx
    ly = (height - cy) / fy + 0.3 * 0.5 * height / fy, cy / fy + 0.3 * 0.5 * height / fy
    splats = []
    for g in range(len(means)):
        x, y, z = p[g]
        o = float(scene["opacities"][g])
        if o < 1 / 255 or z < 0.01 or z > 100:
            continue
        tx = z * min(lx[0], max(-lx[1], x / z))
        ty = z * min(ly[0], max(-ly[1], y / z))
        jac = np.array([[fx / z, 0, -fx * tx / z ** 2], [0, fy / z, -fy * ty / z ** 2]])
        cov2 = jac @ rot @ full[g] @ rot.T @ jac.T + eps * np.eye(2)
        if np.linalg.det(cov2) <= 0:
            continue
        extend = min(sigma, math.sqrt(2 * math.log(o * 255)))
        mx, my = fx * x / z + cx, fy * y / z + cy
        rx, ry = math.ceil(extend * math.sqrt(cov2[0, 0])), math.ceil(extend * math.sqrt(cov2[1, 1]))
        if mx + rx <= 0 or my + ry <= 0 or mx - rx >= width or my - ry >= height:
            continue
        splats.append((z, g, mx, my, np.linalg.inv(cov2), o))
    splats.sort(key=lambda item: (item[0], item[1]))
    for py in range(height):
        for px in range(width):
            t, best, best_id = 1.0, -1.0, -1
            acc = np.zeros(3)
            for z, g, mx, my, conic, o in splats:
                d = np.array([px + 0.5 - mx, py + 0.5 - my])
                power = -0.5 * d @ conic @ d
                if power > 0 or power < -20:
                    continue
                a = min(0.99, o * math.exp(power))
                if a < 1 / 255:
                    continue
                if t * (1 - a) <= 1e-4:
                    break
                w = t * a
                acc += w * scene["features"][g, :3]
                if w > best:
                    best, best_id = w, int(scene["semantic_ids"][g])
                t *= 1 - a
            out_rgb[py, px] = acc
            out_alpha[py, px] = 1 - t
            out_sem[py, px] = best_id if 1 - t >= 0.01 else -1
    return out_rgb, out_alpha, out_sem


def test_oracle_matches_tile_free_float64_brute_force():
    sc = synthetic.gaussian_scene(300, seed=5)
    sc.means[:] = sc.means * 0.4
    sc.scales[:] = sc.scales * 6.0
    scene = {name: value.numpy() for name, value in sc.as_kwargs().items()}
    rig = synthetic.single_camera(40, 32, 30.0, position=(6.0, 1.0, 0.5))
    view, k = rig.viewmats.numpy()[0], rig.intrinsics.numpy()[0]
    r = oracle.render([scene], [0], rig.viewmats, rig.intrinsics, [0], height=32, width=40)
    rgb, alpha, sem = brute_force_render(scene, view, k, 40, 32)
    assert r["counters"][0] > 50
    np.testing.assert_allclose(r["alpha"][0, ..., 0], alpha, atol=2e-5)
    np.testing.assert_allclose(r["rgb"][0], rgb, atol=2e-5)
    agree = np.mean(r["semantic_id"][0, ..., 0] == sem)
    assert agree >= 0.999, agree


# ------------------------------------------------------------------------- SH basis

def _real_sh_3dgs(l, m, direction):
    """3DGS real SH: sqrt(2) Re/Im of scipy's complex SH (Condon-Shortley phase), m=0 plain."""
    x, y, z = direction
    theta = math.acos(max(-1.0, min(1.0, z)))
    phi = math.atan2(y, x)
    value = scipy.special.sph_harm_y(l, abs(m), theta, phi)
    if m > 0:
        return math.sqrt(2) * value.real
    if m < 0:
        return math.sqrt(2) * value.imag
    return value.real


def test_sh_basis_matches_scipy_for_every_band():
    rng = np.random.default_rng(9)
    for _ in range(20):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        for degree in (1, 2, 3):
            bands = (degree + 1) ** 2
            for k_index in range(1, bands):
                l = int(math.isqrt(k_index))
                m = k_index - l * l - l
                f = np.zeros(3 * bands, np.float32)
                f[:3] = 0.5
                f[3 * k_index + 1] = 0.25  # green channel only
                colour = oracle.sh_colour(f, degree, d * 2.5)
                expected = max(0.5 + 0.25 * _real_sh_3dgs(l, m, d), 0.0)
                assert colour[1] == pytest.approx(expected, abs=2e-6), (degree, l, m)
                assert colour[0] == pytest.approx(0.5) and colour[2] == pytest.approx(0.5)


def test_sh_degree_zero_is_base_colour():
    f = np.arange(48, dtype=np.float32) / 100
    np.testing.assert_array_equal(oracle.sh_colour(f, 0, (1.0, 2.0, 3.0)), f[:3])


# ----------------------------------------------------------------------- invariants

def test_ordering_invariants_and_thread_independence():
    sc = synthetic.gaussian_scene(20_000, seed=77, sh_degree=1)
    rig = synthetic.orbit_rig(3, 96, 64, 83.0, seed=77, radius=(9.0, 12.0))
    kwargs = dict(height=64, width=96, sh_degree=1, want_sorted=True)
    r1 = oracle.render([sc.as_kwargs()], [0], rig.viewmats, rig.intrinsics, [0, 0, 0], threads=1, **kwargs)
    r8 = oracle.render([sc.as_kwargs()], [0], rig.viewmats, rig.intrinsics, [0, 0, 0], threads=8, **kwargs)
    for name in ("rgb", "depth", "alpha", "semantic_id", "counters", "tile_counts", "sorted_keys",
                 "sorted_gaussians"):
        np.testing.assert_array_equal(r1[name], r8[name], err_msg=name)
    keys = r1["sorted_keys"]
    assert np.all(keys[1:] >= keys[:-1])
    ties = keys[1:] == keys[:-1]
    assert np.all(r1["sorted_gaussians"][1:][ties] > r1["sorted_gaussians"][:-1][ties])
    assert int(r1["tile_counts"].sum()) == int(r1["counters"][1]) == keys.shape[0]
    assert int(np.count_nonzero(r1["tile_counts"])) == int(r1["counters"][4])
