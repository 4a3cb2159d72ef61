"""Backward pass on the GPU against the dense FP64 reference.

Reference gradients come from torch autograd of ``oracle/dense_reference.py``
in float64 (whose forward is pinned to the C oracle and whose derivatives are
pinned by central differences in ``tests/test_dense_reference.py``).  Here
the HIP gradients are compared with them element-wise, and directly with
float64 central differences (h = 1e-4) along random directions, on the
1,000-Gaussian / 64x64 subset of the differentiable workload (first 1,000
Gaussians of the 500k SH3 scene, seed 20261204).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import dense_reference as dense
from vgr_amd import synthetic
from vgr_amd.autograd import DifferentiableRenderer, compose_pose, se3_exp

pytestmark = pytest.mark.gpu

SEED = 20261204


def device():
    return torch.device("cuda", 0)


def subset_scene(count: int = 1000, log_scale_shift: float = 0.0, sh_degree: int = 3):
    scene = synthetic.gaussian_scene(500_000, seed=SEED, sh_degree=sh_degree)
    means = scene.means[:count].clone()
    log_scales = torch.log(scene.scales[:count]) + log_scale_shift
    quats = scene.rotations[:count].clone()
    logits = torch.logit(scene.opacities[:count].double()).float()
    features = scene.features[:count].clone()
    return means, log_scales, quats, logits, features


def loss_weights(batch, height, width, seed):
    g = torch.Generator().manual_seed(seed)
    return (torch.randn((batch, height, width, 3), generator=g, dtype=torch.float64),
            torch.randn((batch, height, width, 1), generator=g, dtype=torch.float64),
            torch.randn((batch, height, width, 1), generator=g, dtype=torch.float64))


def linear_loss(out, weights, depth_mask):
    w_rgb, w_alpha, w_depth = weights
    depth = torch.where(depth_mask, out["depth"], torch.zeros_like(out["depth"]))
    return (out["rgb"] * w_rgb).sum() + (out["alpha"] * w_alpha).sum() + (depth * w_depth).sum()


def dense_loss(params, base_views, intrinsics, opt, sh_degree, weights, structure=None, antialiased=False):
    means, log_scales, quats, logits, features, xi = params
    views = se3_exp(xi) @ base_views
    out, structure = dense.render(means, torch.exp(log_scales), quats, torch.sigmoid(logits), features, views,
                                  intrinsics, opt, sh_degree=sh_degree, structure=structure)
    shaped = {"rgb": out["rgb"], "alpha": out["alpha"].unsqueeze(-1), "depth": out["depth"].unsqueeze(-1)}
    mask = torch.isfinite(shaped["depth"])
    return linear_loss(shaped, weights, mask), structure, mask


def hip_gradients(params32, base_views, intrinsics, height, width, sh_degree, weights, depth_mask, **options):
    renderer = DifferentiableRenderer(height=height, width=width, device=device(), **options)
    leaves = [p.to(device()).clone().requires_grad_(True) for p in params32]
    means, log_scales, quats, logits, features, xi = leaves
    views = compose_pose(xi, base_views.to(device()).float().contiguous())
    out = renderer.render(means, log_scales, quats, logits, features, views, intrinsics.to(device()),
                          sh_degree=sh_degree)
    w = [t.to(device()).float() for t in weights]
    loss = linear_loss(out, w, depth_mask.to(device()))
    loss.backward()
    return [leaf.grad.detach().double().cpu() for leaf in leaves], out


NAMES = ("means", "log_scales", "quats", "opacity_logits", "features", "pose")


def compare(hip, ref, label):
    report = {}
    for name, a, b in zip(NAMES, hip, ref):
        scale = float(b.abs().max())
        error = float((a - b).abs().max())
        rel_l2 = float((a - b).norm() / max(b.norm(), 1e-30))
        report[name] = (error / max(scale, 1e-30), rel_l2)
    print(f"[backward] {label}: " + ", ".join(f"{k} max {v[0]:.2e} l2 {v[1]:.2e}" for k, v in report.items()),
          flush=True)
    return report


def setup(count=1000, shift=0.0, sh_degree=3, cameras=1, size=64, seed=SEED):
    params = list(subset_scene(count, shift, sh_degree))
    rig = synthetic.orbit_rig(cameras, size, size, 443.4 * size / 512, seed=seed, radius=(9.0, 11.0))
    xi = torch.zeros((cameras, 6), dtype=torch.float32)
    params.append(xi)
    return params, rig.viewmats, rig.intrinsics


def reference(params, views, intrinsics, size, sh_degree, weights, antialiased=False):
    opt = dense.Options(height=size, width=size, antialiased=antialiased)
    leaves = [p.double().clone().requires_grad_(True) for p in params]
    loss, structure, mask = dense_loss(leaves, views.double(), intrinsics.double(), opt, sh_degree, weights)
    loss.backward()
    return [leaf.grad.detach() for leaf in leaves], structure, mask, opt


# Tile size 16 runs the wave-per-tile backward compositor, tile size 8 the
# one-pixel-per-lane kernel: both are compared with the same references.
@pytest.mark.parametrize("tile_size", [16, 8])
@pytest.mark.parametrize("shift", [0.0, 1.5], ids=["c4-subset", "dense-overlap"])
def test_gradients_match_dense_fp64_autograd(shift, tile_size):
    size, degree = 64, 3
    params, views, intrinsics = setup(shift=shift)
    weights = loss_weights(1, size, size, seed=7)
    ref, _, mask, _ = reference(params, views, intrinsics, size, degree, weights)
    hip, _ = hip_gradients(params, views, intrinsics, size, size, degree, weights, mask, tile_size=tile_size)
    report = compare(hip, ref, f"vs FP64 autograd (log-scale shift {shift}, tile {tile_size})")
    for name, (max_rel, l2) in report.items():
        assert max_rel < 2e-3 and l2 < 1e-3, (name, max_rel, l2)


def test_gradients_match_fp64_central_differences():
    size, degree, h = 64, 3, 1e-4
    params, views, intrinsics = setup(shift=1.0)
    weights = loss_weights(1, size, size, seed=11)
    ref, structure, mask, opt = reference(params, views, intrinsics, size, degree, weights)
    hip, _ = hip_gradients(params, views, intrinsics, size, size, degree, weights, mask)
    g = torch.Generator().manual_seed(12)
    base = [p.double() for p in params]
    for index, name in enumerate(NAMES):
        direction = torch.randn(base[index].shape, generator=g, dtype=torch.float64)
        plus = [b + h * direction if i == index else b for i, b in enumerate(base)]
        minus = [b - h * direction if i == index else b for i, b in enumerate(base)]
        with torch.no_grad():
            lp = dense_loss(plus, views.double(), intrinsics.double(), opt, degree, weights, structure)[0]
            lm = dense_loss(minus, views.double(), intrinsics.double(), opt, degree, weights, structure)[0]
        fd = float((lp - lm) / (2 * h))
        analytic = float((hip[index] * direction).sum())
        print(f"[backward] FD {name}: hip {analytic:.6e} fd {fd:.6e}", flush=True)
        assert abs(analytic - fd) <= 1e-3 * abs(fd) + 1e-6, (name, analytic, fd)


@pytest.mark.parametrize("tile_size", [16, 8])
def test_pose_and_scene_gradients_over_several_cameras(tile_size):
    size, degree = 48, 1
    params, views, intrinsics = setup(count=1500, shift=1.0, sh_degree=degree, cameras=5, size=size, seed=31)
    weights = loss_weights(5, size, size, seed=13)
    ref, _, mask, _ = reference(params, views, intrinsics, size, degree, weights)
    hip, _ = hip_gradients(params, views, intrinsics, size, size, degree, weights, mask, cameras_per_block=2,
                           tile_size=tile_size)
    report = compare(hip, ref, f"5 cameras, 2 per block, tile {tile_size}")
    for name, (max_rel, l2) in report.items():
        assert max_rel < 2e-3 and l2 < 1e-3, (name, max_rel, l2)


def test_antialiased_gradients():
    size, degree = 64, 0
    params, views, intrinsics = setup(shift=1.0, sh_degree=degree)
    weights = loss_weights(1, size, size, seed=17)
    ref, _, mask, _ = reference(params, views, intrinsics, size, degree, weights, antialiased=True)
    hip, _ = hip_gradients(params, views, intrinsics, size, size, degree, weights, mask,
                           rasterize_mode="antialiased")
    report = compare(hip, ref, "antialiased")
    for name, (max_rel, l2) in report.items():
        assert max_rel < 2e-3 and l2 < 1e-3, (name, max_rel, l2)


def test_checkpointed_forward_equals_plain_forward():
    size, degree = 64, 3
    params, views, intrinsics = setup(shift=1.0)
    renderer = DifferentiableRenderer(height=size, width=size, device=device())
    means, log_scales, quats, logits, features, _ = [p.to(device()) for p in params]
    out = renderer.render(means, log_scales, quats, logits, features, views.to(device()), intrinsics.to(device()),
                          sh_degree=degree)
    from oracle import oracle
    # exp / sigmoid as evaluated on the device, so both renders see the same bits.
    scene = dict(means=params[0], scales=torch.exp(log_scales).cpu(), rotations=params[2],
                 opacities=torch.sigmoid(logits).cpu(), features=params[4],
                 semantic_ids=torch.zeros(params[0].shape[0], dtype=torch.int64))
    expected = oracle.render([scene], [0], views, intrinsics, [0], height=size, width=size, sh_degree=degree)
    got = {k: v.detach().cpu().numpy() for k, v in out.items()}
    np.testing.assert_array_equal(got["rgb"], expected["rgb"])
    np.testing.assert_array_equal(got["alpha"], expected["alpha"])
    np.testing.assert_array_equal(got["depth"], expected["depth"])


def test_deferred_capacity_check_fails_loud_and_recovers():
    from vgr_amd.hip_backend import RendererCapacityError

    size, degree = 64, 3
    params, views, intrinsics = setup(shift=1.0)
    renderer = DifferentiableRenderer(height=size, width=size, device=device(), capacity_check="deferred")
    leaves = [p.to(device()).clone().requires_grad_(True) for p in params[:5]]
    views_d, k_d = views.to(device()), intrinsics.to(device())

    def step():
        for leaf in leaves:
            leaf.grad = None
        out = renderer.render(*leaves, views_d, k_d, sh_degree=degree)
        out["rgb"].sum().backward()
        return [leaf.grad.clone() for leaf in leaves]

    reference = step()  # first call: synchronous calibration
    assert all(bool(torch.isfinite(g).all()) for g in reference)
    again = step()      # deferred: no host synchronization, same result
    for a, b in zip(again, reference):
        torch.testing.assert_close(a, b, rtol=1e-5, atol=1e-6)
    renderer.visible_capacity = 16  # buffers far too small for the next render
    poisoned = step()
    assert any(bool(torch.isnan(g).any()) for g in poisoned), "overflowed backward must not be silent"
    with pytest.raises(RendererCapacityError):
        step()
    recovered = step()  # buffers were grown: synchronous again, then correct
    for a, b in zip(recovered, reference):
        torch.testing.assert_close(a, b, rtol=1e-5, atol=1e-6)


def test_compose_pose_matches_fp64_se3_exp_autograd():
    """vgr_pose_compose(+_backward) against torch autograd of se3_exp in float64:
    zero, tiny, series-boundary, moderate and large rotation angles."""
    g = torch.Generator().manual_seed(SEED + 7)
    angles = torch.tensor([0.0, 1e-6, 1e-3, 0.049, 0.051, 0.3, 1.0, 2.5, 3.1], dtype=torch.float64)
    count = angles.numel()
    axes = torch.nn.functional.normalize(torch.randn((count, 3), generator=g, dtype=torch.float64), dim=-1)
    xi = torch.cat((torch.randn((count, 3), generator=g, dtype=torch.float64), axes * angles[:, None]), dim=-1)
    rig = synthetic.orbit_rig(count, 64, 64, 55.0, seed=SEED, radius=(9.0, 11.0))
    base = rig.viewmats.double()
    upstream = torch.randn((count, 4, 4), generator=g, dtype=torch.float64)
    xi64 = xi.clone().requires_grad_(True)
    base64 = base.clone().requires_grad_(True)
    expected = se3_exp(xi64) @ base64
    (expected * upstream).sum().backward()
    xi32 = xi.float().to(device()).requires_grad_(True)
    base32 = base.float().to(device()).requires_grad_(True)
    views = compose_pose(xi32, base32)
    (views * upstream.float().to(device())).sum().backward()
    torch.cuda.synchronize()
    # float32 inputs and outputs around double arithmetic: errors of a few float32 ulps.
    assert torch.allclose(views.double().cpu(), expected.detach(), rtol=0, atol=1e-5 * float(base.abs().max()))
    scale_xi = float(xi64.grad.abs().max())
    assert float((xi32.grad.double().cpu() - xi64.grad).abs().max()) <= 2e-5 * scale_xi
    scale_base = float(base64.grad.abs().max())
    assert float((base32.grad.double().cpu() - base64.grad).abs().max()) <= 2e-5 * scale_base
