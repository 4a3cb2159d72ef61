"""Camera-pose gradients of the batched-persistent renderer (tile-front ordering).

The batched backward walks each tile's front-sorted list back to front from the
checkpointed transmittance and feeds every (Gaussian, tile) pair's ten record
terms through the pair's projection backward into per-camera pose partials.
The projection backward is linear in those terms, so its sum over a pair's tiles
equals the plain training path's per-pair chain up to float rounding; the plain
path's pose gradients are pinned against float64 autograd and central
differences in test_gpu_backward.py.  Here the two are compared on a robotics
rig in a closed room (SH1, 160x120, fx = fy = 139), with images bitwise equal.
"""

from __future__ import annotations

import pytest
import torch

from vgr_amd import synthetic
from vgr_amd.autograd import DifferentiableRenderer, compose_pose
from vgr_amd.batched import BatchedRenderer, CameraSpec

pytestmark = pytest.mark.gpu

WIDTH, HEIGHT, FOCAL = 160, 120, 139.0


def device():
    return torch.device("cuda", 0)


def room_case(count=150_000, seed=20261208, trajectories=2, samples=4):
    scene = synthetic.gaussian_scene(count, seed=seed, sh_degree=1, semantic_classes=256, layout="room")
    rig = synthetic.trajectory_rig(trajectories, samples, WIDTH, HEIGHT, FOCAL, seed=seed)
    dev = device()
    log_scales = torch.log(scene.scales).to(dev)
    logits = torch.logit(scene.opacities.double()).float().to(dev)
    params = dict(means=scene.means.to(dev), log_scales=log_scales, quats=scene.rotations.to(dev),
                  logits=logits, features=scene.features.to(dev), semantic_ids=scene.semantic_ids.to(dev))
    return params, rig


def batched_renderer(params, batch, **kwargs):
    renderer = BatchedRenderer(device=device(), **kwargs)
    # The activations the differentiable renderer applies, so both render the same scene.
    renderer.register_scene(0, means=params["means"], scales=torch.exp(params["log_scales"]),
                            rotations=params["quats"], opacities=torch.sigmoid(params["logits"]),
                            features=params["features"], semantic_ids=params["semantic_ids"])
    renderer.plan([CameraSpec(0, WIDTH, HEIGHT)] * batch)
    return renderer


def weights(batch, seed):
    g = torch.Generator().manual_seed(seed)
    shape = (batch, HEIGHT, WIDTH)
    return [torch.randn(shape + (c,), generator=g).to(device()) for c in (3, 1, 1)]


def loss_of(rgb, depth, alpha, w):
    finite = torch.isfinite(depth)
    depth = torch.where(finite, depth, torch.zeros_like(depth))
    return (rgb * w[0]).sum() + (alpha * w[1]).sum() + (depth * w[2] * finite).sum()


def test_batched_pose_gradients_match_the_plain_training_path():
    params, rig = room_case()
    batch = rig.batch
    viewmats = rig.viewmats.to(device())
    intrinsics = rig.intrinsics.to(device())
    w = weights(batch, 5)

    renderer = batched_renderer(params, batch, pose_gradients=True, ordering="front")
    renderer.calibrate(viewmats, intrinsics)
    assert renderer.ordering == "front" and renderer.chunk_count == 1
    renderer.capture()
    views = viewmats.clone().requires_grad_(True)
    rgb, depth, alpha = renderer.render_pose(views)
    loss_of(rgb.view(batch, HEIGHT, WIDTH, 3), depth.view(batch, HEIGHT, WIDTH, 1),
            alpha.view(batch, HEIGHT, WIDTH, 1), w).backward()
    batched_grad = views.grad.detach().clone()

    trainer = DifferentiableRenderer(height=HEIGHT, width=WIDTH, device=device())
    plain_views = viewmats.clone().requires_grad_(True)
    out = trainer.render(params["means"], params["log_scales"], params["quats"], params["logits"],
                         params["features"], plain_views, intrinsics, semantic_ids=params["semantic_ids"])
    for name, flat in (("rgb", rgb), ("depth", depth), ("alpha", alpha)):
        assert torch.equal(flat.view_as(out[name]), out[name].detach()), name
    loss_of(out["rgb"], out["depth"], out["alpha"], w).backward()
    plain_grad = plain_views.grad.detach()

    assert torch.isfinite(batched_grad).all()
    assert batched_grad[:, 3].abs().max() == 0  # the bottom row of a view matrix carries no gradient
    scale = plain_grad.abs().amax(dim=(1, 2), keepdim=True)
    error = (batched_grad - plain_grad).abs().amax(dim=(1, 2), keepdim=True)
    print(f"[batched pose] max |grad| {float(scale.max()):.4g}, max error / camera scale "
          f"{float((error / scale).max()):.3g}", flush=True)
    # Summation order differs (per tile vs per pair, replica atomics): rounding only.
    assert torch.all(error <= 2e-3 * scale + 1e-6)

    # Replaying the captured forward and the backward again gives the same gradients
    # up to the order of the float atomics.
    views.grad = None
    rgb2, depth2, alpha2 = renderer.render_pose(views)
    loss_of(rgb2.view(batch, HEIGHT, WIDTH, 3), depth2.view(batch, HEIGHT, WIDTH, 1),
            alpha2.view(batch, HEIGHT, WIDTH, 1), w).backward()
    assert torch.equal(rgb2, rgb)
    assert torch.all((views.grad - batched_grad).abs() <= 1e-4 * scale + 1e-7)


def test_pose_checkpoints_do_not_change_the_images():
    params, rig = room_case(count=80_000, seed=11, trajectories=1, samples=6)
    viewmats = rig.viewmats.to(device())
    intrinsics = rig.intrinsics.to(device())
    plain = batched_renderer(params, rig.batch, ordering="front")
    plain.calibrate(viewmats, intrinsics)
    posed = batched_renderer(params, rig.batch, ordering="front", pose_gradients=True)
    posed.calibrate(viewmats, intrinsics)
    a = plain.render()
    b = posed.render()
    for name in ("rgb", "depth", "alpha", "semantic_id"):
        assert torch.equal(a[name], b[name]), name
    # Twist parameterisation on top: dL/d xi through the fused pose composition.
    xi = torch.zeros((rig.batch, 6), device=device(), requires_grad=True)
    rgb, _, alpha = posed.render_pose(compose_pose(xi, viewmats))
    (rgb.mean() + alpha.mean()).backward()
    assert torch.isfinite(xi.grad).all() and xi.grad.abs().sum() > 0


def test_pose_gradients_need_the_tile_front_ordering():
    with pytest.raises(ValueError, match="tile-front"):
        BatchedRenderer(device=device(), pose_gradients=True, ordering="global")
    with pytest.raises(ValueError, match="linear RGB"):
        BatchedRenderer(device=device(), pose_gradients=True, output_srgb=True)
    params, rig = room_case(count=20_000, seed=3, trajectories=1, samples=2)
    renderer = batched_renderer(params, rig.batch)
    renderer.calibrate(rig.viewmats.to(device()), rig.intrinsics.to(device()))
    with pytest.raises(RuntimeError, match="pose_gradients=True"):
        renderer.backward_pose(torch.zeros((renderer.total_pixels, 3), device=device()))
