"""Tile-front ordering (gfx950): per-tile buckets, an LDS sort of each tile's
nearest entries, records recomputed by the compositor, and the full-sort
fallback for tiles left open at the end of a truncated list.

Every case renders the same batch with the global radix ordering as well and
requires bitwise-equal outputs (the two orderings must composite the same
entries in the same order); the room case is also checked against the CPU
oracle (RGB/alpha max-abs 1e-4, depth relative 1e-4, semantic IDs exact).
"""

from __future__ import annotations

import math

import pytest
import torch

from gpu_helpers import compare_images
from oracle import oracle
from vgr_amd import synthetic
from vgr_amd.batched import BatchedRenderer, CameraSpec

pytestmark = pytest.mark.gpu


def device():
    return torch.device("cuda", 0)


def render_with(ordering, scene, specs, viewmats, intrinsics, **kwargs):
    renderer = BatchedRenderer(device=device(), ordering=ordering, **kwargs)
    renderer.register_scene(0, **scene.to(device()).as_kwargs())
    renderer.plan(specs)
    counters = renderer.calibrate(viewmats.to(device()), intrinsics.to(device()))
    assert counters["sort_lookback_failures"] == 0, counters
    statistics = torch.zeros(4, dtype=torch.int64, device=device())
    renderer._submit(statistics=statistics)
    out = {name: t.clone() for name, t in renderer.outputs.items()}
    torch.cuda.synchronize()
    return renderer, out, counters, [int(v) for v in statistics.cpu()]


def assert_bitwise(a, b):
    for name in ("rgb", "depth", "alpha", "semantic_id"):
        assert torch.equal(a[name], b[name]), name


def test_room_front_matches_global_and_oracle():
    scene = synthetic.gaussian_scene(1_000_000, seed=301, sh_degree=1, semantic_classes=256, layout="room")
    rig = synthetic.trajectory_rig(4, 32, 160, 120, 139.0, seed=302)
    specs = [CameraSpec(0, 160, 120)] * rig.batch
    front, out_front, counters, stats = render_with("front", scene, specs, rig.viewmats, rig.intrinsics)
    assert front.ordering == "front"
    _, out_global, counters_global, _ = render_with("global", scene, specs, rig.viewmats, rig.intrinsics)
    print(f"[front] room: counters {counters}, fallback tiles {stats[3]}, global counters {counters_global}",
          flush=True)
    assert counters["tile_intersections"] == counters_global["tile_intersections"]
    assert counters["visible_gaussians"] == counters_global["visible_gaussians"]
    assert_bitwise(out_front, out_global)
    subset = [0, 45, 90, 127]
    expected = oracle.render([scene.as_kwargs()], [0], rig.viewmats[subset], rig.intrinsics[subset], [0] * 4,
                             height=120, width=160, sh_degree=1)
    for local, camera in enumerate(subset):
        view = front.camera_outputs(camera, out_front)
        got = {name: t.cpu().numpy() for name, t in view.items()}
        want = {name: expected[name][local] for name in got}
        print(f"[front] room camera {camera}: {compare_images(got, want)}", flush=True)


def faint_cluster(count: int, seed: int, depth_levels: int):
    """Low-opacity splats packed into the first 16x16 tile of a 64x64 camera at the origin
    (looking down +z, f = 100): tile lists far longer than the 4096-entry front and pixels that
    never saturate, so compositing reports the tile open and the fallback sorts it in full.
    Depths take ``depth_levels`` distinct values, so equal-depth runs cross the front threshold."""
    g = torch.Generator().manual_seed(seed)
    level = torch.randint(0, depth_levels, (count,), generator=g).float()
    z = 2.0 + 0.5 * level
    px = 16.0 * torch.rand((count,), generator=g)
    py = 16.0 * torch.rand((count,), generator=g)
    means = torch.stack([(px - 32.0) / 100.0 * z, (py - 32.0) / 100.0 * z, z], dim=1)
    scene = synthetic.gaussian_scene(count, seed=seed, sh_degree=0)
    scene.means = means.float().contiguous()
    scene.scales = torch.full((count, 3), 0.002)
    scene.opacities = torch.full((count,), 0.05)
    view = torch.eye(4).unsqueeze(0)
    intrinsics = synthetic.pinhole(1, 64, 64, 100.0)
    return scene, view, intrinsics


def test_open_tiles_fall_back_to_a_full_sort():
    scene, view, intrinsics = faint_cluster(12_000, seed=311, depth_levels=9)
    specs = [CameraSpec(0, 64, 64)]
    front, out_front, counters, stats = render_with("front", scene, specs, view, intrinsics)
    _, out_global, _, _ = render_with("global", scene, specs, view, intrinsics)
    print(f"[front] faint cluster: counters {counters}, fallback tiles {stats[3]}", flush=True)
    assert stats[3] >= 1, "the long, unsaturated tile must take the full-sort fallback"
    assert_bitwise(out_front, out_global)
    assert float(out_front["alpha"].max()) < 1.0 - 1e-4


def test_front_ties_at_the_threshold_keep_gaussian_id_order():
    # 6000 splats on 2 depth levels: the first level (about 3000 entries) fits the
    # front, the second straddles it; within a level the order is by Gaussian ID.
    scene, view, intrinsics = faint_cluster(6_000, seed=321, depth_levels=2)
    scene.opacities = torch.full((6_000,), 0.3)  # saturating: the tile completes inside the front
    specs = [CameraSpec(0, 64, 64)]
    _, out_front, counters, stats = render_with("front", scene, specs, view, intrinsics)
    _, out_global, _, _ = render_with("global", scene, specs, view, intrinsics)
    print(f"[front] ties: counters {counters}, fallback tiles {stats[3]}", flush=True)
    assert_bitwise(out_front, out_global)


def test_orbit_front_matches_global_with_graph_replay():
    scene = synthetic.gaussian_scene(150_000, seed=331, sh_degree=2)
    rig = synthetic.orbit_rig(12, 192, 128, 166.0, seed=332, radius=(9.0, 12.0))
    specs = [CameraSpec(0, 192, 128)] * rig.batch
    front, out_front, counters, stats = render_with("front", scene, specs, rig.viewmats, rig.intrinsics)
    _, out_global, _, _ = render_with("global", scene, specs, rig.viewmats, rig.intrinsics)
    print(f"[front] orbit SH2: counters {counters}, fallback tiles {stats[3]}", flush=True)
    assert_bitwise(out_front, out_global)
    front.capture()
    for _ in range(3):
        replay = {name: t.clone() for name, t in front.render().items()}
        torch.cuda.synchronize()
        assert_bitwise(replay, out_global)
    assert math.isfinite(float(out_front["rgb"].sum()))
