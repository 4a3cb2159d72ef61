"""Batched-persistent path on the GPU: uneven resolutions over several scenes in
one submission, spatial reordering + chunk culling + persistent compositing,
hipGraph capture.  Checked against the CPU oracle and against the plain path
(bitwise)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from gpu_helpers import compare_images
from oracle import oracle
from vgr_amd import synthetic
from vgr_amd.batched import BatchedRenderer, CameraSpec

pytestmark = pytest.mark.gpu


def device():
    return torch.device("cuda", 0)


def mixed_rig(seed: int, resolutions: list[tuple[int, int]]):
    """Cameras on a sphere with per-camera resolution and 60 degree horizontal FOV."""
    rigs = []
    for i, (w, h) in enumerate(resolutions):
        focal = (w / 2) / np.tan(np.radians(30))
        rigs.append(synthetic.orbit_rig(len(resolutions), w, h, focal, seed=seed, radius=(9.0, 13.0)))
    viewmats = torch.stack([rigs[i].viewmats[i] for i in range(len(resolutions))])
    intrinsics = torch.stack([rigs[i].intrinsics[i] for i in range(len(resolutions))])
    return viewmats.contiguous(), intrinsics.contiguous()


def build(scenes, specs, **kwargs):
    renderer = BatchedRenderer(device=device(), **kwargs)
    for scene_id, scene in scenes:
        renderer.register_scene(scene_id, **scene.to(device()).as_kwargs())
    renderer.plan(specs)
    return renderer


def oracle_flat(scenes, specs, viewmats, intrinsics, sh_degree):
    return oracle.render([s.as_kwargs() for _, s in scenes], [sid for sid, _ in scenes], viewmats, intrinsics,
                         [c.scene_id for c in specs], height=1, width=1,
                         resolutions=[(c.width, c.height) for c in specs], sh_degree=sh_degree)


def flat_numpy(outputs):
    return {name: t.detach().cpu().numpy() for name, t in outputs.items()}


def test_uneven_multi_scene_batch_matches_oracle_and_plain():
    scene_a = synthetic.gaussian_scene(40_000, seed=101, sh_degree=3)
    scene_b = synthetic.gaussian_scene(25_000, seed=102, sh_degree=3)
    resolutions = [(128, 128), (320, 240), (160, 96), (200, 150), (128, 128), (96, 64)]
    specs = [CameraSpec(scene_id=[4, 9][i % 2], width=w, height=h) for i, (w, h) in enumerate(resolutions)]
    scenes = [(4, scene_a), (9, scene_b)]
    viewmats, intrinsics = mixed_rig(103, resolutions)
    renderer = build(scenes, specs)
    counters = renderer.calibrate(viewmats.to(device()), intrinsics.to(device()))
    batched = flat_numpy(renderer.render())
    expected = oracle_flat(scenes, specs, viewmats, intrinsics, 3)
    report = compare_images(batched, expected)
    print(f"[batched] oracle: {report}", flush=True)
    assert counters["tile_intersections"] == int(expected["counters"][1])
    assert counters["visible_gaussians"] == int(expected["counters"][0])
    renderer.render_plain()
    for group in renderer.plain_groups:
        for local, camera in enumerate(group["index_list"]):
            view = renderer.camera_outputs(camera)
            for name in ("rgb", "depth", "alpha", "semantic_id"):
                torch.testing.assert_close(view[name], group["outputs"][name][local], rtol=0, atol=0)


def test_graph_replay_matches_eager_across_camera_changes():
    scene = synthetic.gaussian_scene(30_000, seed=111, sh_degree=1)
    resolutions = [(160, 120)] * 4 + [(96, 64)] * 2
    specs = [CameraSpec(0, w, h) for w, h in resolutions]
    views_a, k_a = mixed_rig(112, resolutions)
    views_b, k_b = mixed_rig(113, resolutions)
    renderer = build([(0, scene)], specs)
    renderer.calibrate(views_a.to(device()), k_a.to(device()))
    renderer.calibrate(views_b.to(device()), k_b.to(device()))  # capacity for both frames
    eager = {}
    for name, (v, k) in {"a": (views_a, k_a), "b": (views_b, k_b)}.items():
        renderer.graph = None
        eager[name] = {n: t.clone() for n, t in renderer.render(v.to(device()), k.to(device())).items()}
    renderer.capture()
    for replay in range(4):
        frame = "ab"[replay % 2]
        v, k = (views_a, k_a) if frame == "a" else (views_b, k_b)
        renderer.set_cameras(v.to(device()), k.to(device()))
        out = renderer.render()
        torch.cuda.synchronize()
        scratch = torch.randn(1 << 16, device=device()).sum()  # unrelated eager work between replays
        assert torch.isfinite(scratch)
        for name in out:
            assert torch.equal(out[name], eager[frame][name]), (replay, name)
    assert renderer.check_capacity()["visible_overflow"] == 0


def test_chunk_culling_drops_work_in_a_room_and_keeps_results():
    scene = synthetic.gaussian_scene(120_000, seed=121, sh_degree=1, layout="room", semantic_classes=256)
    rig = synthetic.trajectory_rig(2, 8, 160, 120, 139.0, seed=121)
    specs = [CameraSpec(0, 160, 120)] * rig.batch
    culled = build([(0, scene)], specs, chunk_culling=True)
    culled.calibrate(rig.viewmats.to(device()), rig.intrinsics.to(device()))
    statistics = torch.zeros(4, dtype=torch.int64, device=device())
    culled._submit(statistics=statistics)
    kept = int(statistics[2])
    chunks = (scene.count + 255) // 256
    groups = (rig.batch + 7) // 8
    print(f"[batched] culling kept {kept} of {chunks * groups} (group, chunk) items", flush=True)
    assert 0 < kept < chunks * groups
    out = flat_numpy(culled.outputs)
    unculled = build([(0, scene)], specs, chunk_culling=False, spatial_order=False, persistent_composite=0)
    unculled.calibrate(rig.viewmats.to(device()), rig.intrinsics.to(device()))
    reference = flat_numpy(unculled.outputs)
    for name in out:
        np.testing.assert_array_equal(out[name], reference[name], err_msg=name)


@pytest.mark.parametrize("count", [48, 400])
def test_spatial_order_keeps_gaussian_id_tie_order(count):
    """Equal depths give runs of equal keys: insertion sort up to 32, heap sort beyond."""
    means = torch.zeros((count, 3))
    means[:, 0] = torch.linspace(-0.05, 0.05, count)  # spread so Morton order differs from ID order
    means[:, 2] = 5.0
    scales = torch.full((count, 3), 0.03)
    rotations = torch.zeros((count, 4))
    rotations[:, 0] = 1.0
    perm = torch.randperm(count, generator=torch.Generator().manual_seed(5))
    means = means[perm].contiguous()
    features = torch.rand((count, 3), generator=torch.Generator().manual_seed(6))
    scene = synthetic.SyntheticScene(means, scales, rotations, torch.full((count,), 0.2), features,
                                     torch.arange(500, 500 + count, dtype=torch.int64))
    specs = [CameraSpec(0, 64, 64)]
    view = torch.eye(4).unsqueeze(0)
    view[0, 2, 3] = -4.0
    k = synthetic.pinhole(1, 64, 64, 55.0)
    renderer = build([(0, scene)], specs, gaussian_support_sigma=3.0)
    renderer.calibrate(view.to(device()), k.to(device()))
    out = flat_numpy(renderer.outputs)
    expected = oracle.render([scene.as_kwargs()], [0], view, k, [0], height=64, width=64,
                             gaussian_support_sigma=3.0)
    for name in ("rgb", "alpha", "depth", "semantic_id"):
        np.testing.assert_array_equal(out[name].reshape(expected[name].shape), expected[name], err_msg=name)


def test_autotuned_plan_matches_oracle(tmp_path):
    from vgr_amd.autotune import Autotuner, tune_batched

    scene = synthetic.gaussian_scene(30_000, seed=131, sh_degree=2)
    resolutions = [(160, 120)] * 3 + [(96, 64)] * 2
    specs = [CameraSpec(0, w, h) for w, h in resolutions]
    views, k = mixed_rig(132, resolutions)
    renderer = build([(0, scene)], specs)
    tuner = Autotuner({"tile_size": (16, 8), "raster_batch": (256, 64), "sort_digit_bits": (8, 6),
                       "cameras_per_block": (8, 2)}, cache_path=tmp_path / "tune.json", repeats=1)
    best = tune_batched(renderer, views.to(device()), k.to(device()), tuner=tuner, force=True)
    print(f"[batched] autotuned {best} from {len(tuner.history)} configurations", flush=True)
    assert (tmp_path / "tune.json").is_file()
    out = flat_numpy(renderer.render())
    expected = oracle.render([scene.as_kwargs()], [0], views, k, [0] * len(specs), height=1, width=1,
                             resolutions=resolutions, sh_degree=2, tile_size=best["tile_size"])
    report = compare_images(out, expected)
    print(f"[batched] tuned vs oracle: {report}", flush=True)
    for name in ("rgb", "alpha", "depth", "semantic_id"):
        np.testing.assert_array_equal(out[name].reshape(expected[name].shape), expected[name], err_msg=name)


def test_split_submissions_match_single_submission():
    scene = synthetic.gaussian_scene(40_000, seed=141, sh_degree=1, layout="room", semantic_classes=256)
    rig = synthetic.trajectory_rig(2, 6, 160, 120, 139.0, seed=141)
    specs = [CameraSpec(0, 160, 120)] * rig.batch
    single = build([(0, scene)], specs)
    counters = single.calibrate(rig.viewmats.to(device()), rig.intrinsics.to(device()))
    assert counters["submissions"] == 1
    reference = flat_numpy(single.outputs)
    limit = max(1, counters["tile_intersections"] // 3)
    split = build([(0, scene)], specs, max_intersections_per_submission=limit)
    split_counters = split.calibrate(rig.viewmats.to(device()), rig.intrinsics.to(device()))
    print(f"[batched] {split_counters['submissions']} submissions for limit {limit}: {split_counters}", flush=True)
    assert split_counters["submissions"] >= 3
    assert split_counters["tile_intersections"] == counters["tile_intersections"]
    assert split_counters["visible_gaussians"] == counters["visible_gaussians"]
    split.capture()
    out = flat_numpy(split.render())
    for name in out:
        np.testing.assert_array_equal(out[name], reference[name], err_msg=name)
    split.render_plain()
    for group in split.plain_groups:
        for local, camera in enumerate(group["index_list"]):
            view = split.camera_outputs(camera)
            for name in ("rgb", "depth", "alpha", "semantic_id"):
                torch.testing.assert_close(view[name], group["outputs"][name][local], rtol=0, atol=0)
