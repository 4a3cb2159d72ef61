"""GPU parity of the HIP renderer against the CPU oracle (gfx950 only).

Every case renders through RendererService + CustomHipBackend (which calls the
C ABI of libvgr_hip.so) and compares with the oracle on the same seeded
inputs: RGB/alpha max-abs <= 1e-4, depth relative <= 1e-4, semantic IDs,
counters, per-tile intersection counts, sorted keys and per-tile Gaussian
order bit-exact.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from gpu_helpers import compare_images, compare_ordering, device, render_oracle, render_service
from vgr_amd import CustomHipBackend, RendererService, synthetic

pytestmark = pytest.mark.gpu


def _print(name, report):
    print(f"[parity] {name}: {report}", flush=True)


def test_config1_single_camera_100k_matches_oracle():
    """1 camera 640x480, 100k Gaussians, SH0, 16x16 tiles: the CPU-runnable workload in full."""
    scene = synthetic.gaussian_scene(100_000, seed=20261201, sh_degree=0)
    rig = synthetic.single_camera(640, 480, 554.0)
    gpu, backend = render_service([(0, scene)], rig, [0])
    expected = render_oracle([(0, scene)], rig, [0], want_sorted=True)
    _print("config1", compare_images(gpu, expected))
    compare_ordering(backend, expected)


def test_sh3_multi_camera_deterministic_matches_oracle():
    scene = synthetic.gaussian_scene(60_000, seed=20261202, sh_degree=3)
    rig = synthetic.orbit_rig(6, 320, 240, 277.0, seed=20261202, radius=(10.0, 15.0))
    gpu, backend = render_service([(5, scene)], rig, [5] * 6, deterministic=True)
    expected = render_oracle([(5, scene)], rig, [5] * 6, sh_degree=3, want_sorted=True)
    _print("sh3", compare_images(gpu, expected))
    compare_ordering(backend, expected)


@pytest.mark.parametrize("items,digit_bits", [(8, 8), (12, 8), (16, 8), (16, 6), (12, 7), (16, 9), (8, 9)])
def test_sort_tile_sizes_and_digit_widths_keep_order_bit_exact(items, digit_bits):
    """Sort tiles of 512 x {8, 12, 16} keys with 6-9-bit digits: several tiles and a partial
    last one (140,757 intersections), ordering bit-exact against the oracle."""
    scene = synthetic.gaussian_scene(50_000, seed=31, sh_degree=0)
    rig = synthetic.orbit_rig(2, 320, 240, 277.0, seed=31, radius=(9.0, 12.0))
    gpu, backend = render_service([(0, scene)], rig, [0, 0], sort_tile_items=items, sort_digit_bits=digit_bits)
    expected = render_oracle([(0, scene)], rig, [0, 0], want_sorted=True)
    n = backend.read_counters()["tile_intersections"]
    assert n > 3 * 512 * items and n % (512 * items) != 0, n
    _print(f"sort {items}x512 d{digit_bits}", compare_images(gpu, expected))
    compare_ordering(backend, expected)


def test_sh1_tile8_antialiased_matches_oracle():
    scene = synthetic.gaussian_scene(40_000, seed=11, sh_degree=1)
    rig = synthetic.orbit_rig(3, 200, 150, 173.0, seed=11, radius=(9.0, 12.0))
    gpu, backend = render_service([(1, scene)], rig, [1] * 3, tile_size=8, rasterize_mode="antialiased")
    expected = render_oracle([(1, scene)], rig, [1] * 3, sh_degree=1, tile_size=8, antialiased=True,
                             want_sorted=True)
    _print("tile8-aa", compare_images(gpu, expected))
    compare_ordering(backend, expected)


def test_multi_scene_active_subset_and_srgb_matches_oracle():
    scene_a = synthetic.gaussian_scene(30_000, seed=21, sh_degree=2)
    scene_b = synthetic.gaussian_scene(12_345, seed=22, sh_degree=2)
    rig = synthetic.orbit_rig(5, 160, 120, 139.0, seed=23, radius=(8.0, 12.0))
    scene_ids = [42, 7, 42, 7, 42]
    active = [4, 1, 2]
    gpu, backend = render_service([(42, scene_a), (7, scene_b)], rig, scene_ids, active_ids=active,
                                  output_srgb=True)
    expected = render_oracle([(42, scene_a), (7, scene_b)], rig, scene_ids, sh_degree=2, active_ids=active,
                             output_srgb=True)
    _print("multi-scene", compare_images(gpu, expected))
    counters = backend.read_counters()
    assert counters["tile_intersections"] == int(expected["counters"][1])
    # Inactive cameras keep the documented defaults.
    assert np.all(gpu["semantic_id"][[0, 3]] == -1)
    assert np.all(np.isinf(gpu["depth"][[0, 3]]))


def test_rgb_only_output_matches_oracle():
    scene = synthetic.gaussian_scene(25_000, seed=31, sh_degree=0)
    rig = synthetic.orbit_rig(2, 256, 192, 221.0, seed=31, radius=(10.0, 12.0))
    gpu, _ = render_service([(0, scene)], rig, [0, 0], outputs=("rgb",))
    expected = render_oracle([(0, scene)], rig, [0, 0], full_sensor_output=False)
    _print("rgb-only", compare_images(gpu, expected, names=("rgb",)))


def test_equal_depth_order_is_by_gaussian_id_and_lowest_id_wins():
    """All Gaussians at the same depth on the optical axis: ties order by ID (deterministic_cuda_smoke.py:33-132)."""
    count = 64
    dev = device()
    means = torch.zeros((count, 3), dtype=torch.float32)
    means[:, 2] = 5.0
    scales = torch.full((count, 3), 0.03)
    rotations = torch.zeros((count, 4))
    rotations[:, 0] = 1.0
    opacities = torch.full((count,), 0.2)
    features = torch.zeros((count, 3))
    features[:, 0] = torch.linspace(1.0, 0.0, count)
    features[:, 2] = torch.linspace(0.0, 1.0, count)
    semantic = torch.arange(1000, 1000 + count, dtype=torch.int64)
    view = torch.eye(4).unsqueeze(0).repeat(2, 1, 1)
    view[:, 2, 3] = -4.0
    k = synthetic.pinhole(2, 64, 64, 55.0)
    backend = CustomHipBackend(gaussian_support_sigma=3.0, output_srgb=False, deterministic=True,
                               max_visible_records=2 * count, max_intersections=2 * count * 64)
    service = RendererService(backend, height=64, width=64, max_views=2)
    service.initialize(stage=None, device=dev)
    service.load_scene(73, means=means.to(dev), scales=scales.to(dev), rotations=rotations.to(dev),
                       opacities=opacities.to(dev), features=features.to(dev), semantic_ids=semantic.to(dev))
    ids = torch.full((2,), 73, dtype=torch.int64, device=dev)
    outputs = service.render(view.to(dev), k.to(dev), ids)
    first = {name: t.clone() for name, t in outputs.items()}
    for _ in range(7):
        service.render(view.to(dev), k.to(dev), ids, outputs=outputs)
    service.synchronize()
    for name in first:
        assert torch.equal(first[name], outputs[name]), f"{name} not bitwise repeatable"
    tiles_x = 4
    center_tile = (32 // 16) * tiles_x + 32 // 16
    start = int(backend.workspace["tile_starts"][center_tile])
    end = int(backend.workspace["tile_ends"][center_tile])
    values = backend.workspace["values_out"][start:end].long()
    gids = backend.workspace["visible_gaussian_ids"][values].cpu()
    assert torch.equal(gids, torch.arange(count, dtype=torch.int32))
    assert int(outputs["semantic_id"][0, 32, 32, 0]) == 1000


def test_far_scene_renders_empty_with_zero_counters():
    scene = synthetic.gaussian_scene(5_000, seed=41)
    scene.means[:, 2] += 1000.0  # beyond the far plane for a camera at the origin looking +z
    view = torch.eye(4).unsqueeze(0)
    rig = synthetic.CameraRig(view, synthetic.pinhole(1, 96, 64, 80.0), [96], [64])
    gpu, backend = render_service([(0, scene)], rig, [0])
    counters = backend.read_counters()
    assert counters == {"visible_gaussians": 0, "tile_intersections": 0, "visible_overflow": 0,
                        "intersection_overflow": 0, "active_tiles": 0, "sort_lookback_failures": 0}
    assert np.all(gpu["alpha"] == 0) and np.all(gpu["rgb"] == 0)
    assert np.all(np.isinf(gpu["depth"])) and np.all(gpu["semantic_id"] == -1)


def test_capacity_overflow_grows_and_matches_oracle():
    scene = synthetic.gaussian_scene(30_000, seed=51, sh_degree=0)
    rig = synthetic.orbit_rig(2, 160, 120, 139.0, seed=51, radius=(10.0, 12.0))
    gpu, backend = render_service([(0, scene)], rig, [0, 0], max_visible_records=1000, max_intersections=1500)
    expected = render_oracle([(0, scene)], rig, [0, 0], want_sorted=True)
    stats = backend.capacity_stats
    assert stats["totals"]["growth_events"] >= 1
    _print("growth", compare_images(gpu, expected))
    compare_ordering(backend, expected)


def test_projection_cache_and_chunked_batches_match_plain():
    scene = synthetic.gaussian_scene(20_000, seed=61, sh_degree=1)
    rig = synthetic.orbit_rig(5, 128, 96, 110.0, seed=61, radius=(9.0, 12.0))
    plain, _ = render_service([(0, scene)], rig, [0] * 5)
    cached_backend = CustomHipBackend(output_srgb=False, enable_projection_cache=True)
    dev = device()
    service = RendererService(cached_backend, height=96, width=128, max_views=5)
    service.initialize(stage=None, device=dev)
    service.load_scene(0, **scene.to(dev).as_kwargs())
    view, k = rig.viewmats.to(dev), rig.intrinsics.to(dev)
    ids = torch.zeros(5, dtype=torch.int64, device=dev)
    out = service.render(view, k, ids)
    service.render(view, k, ids, outputs=out)
    service.synchronize()
    assert cached_backend.projection_cache_stats["hits"] == 1
    for name in plain:
        np.testing.assert_array_equal(out[name].cpu().numpy(), plain[name])
    counters = cached_backend.read_counters()
    chunked = CustomHipBackend(output_srgb=False, fixed_capacity_sort=True, adaptive_capacity=False,
                               max_physical_views=2, max_visible_records=counters["visible_gaussians"] + 100,
                               max_intersections=counters["tile_intersections"] + 1000)
    service2 = RendererService(chunked, height=96, width=128, max_views=5)
    service2.initialize(stage=None, device=dev)
    service2.load_scene(0, **scene.to(dev).as_kwargs())
    out2 = service2.render(view, k, ids)
    service2.synchronize()
    for name in plain:
        np.testing.assert_array_equal(out2[name].cpu().numpy(), plain[name])
    assert chunked.read_counters()["tile_intersections"] == counters["tile_intersections"]


def test_forced_sort_lookback_give_up_fails_closed():
    """Every look-back wait gives up (host flag -1): the counter reports it, the
    backend re-renders, then raises with poisoned outputs; kernels exit normally."""
    from vgr_amd import RendererCapacityError

    scene = synthetic.gaussian_scene(50_000, seed=31, sh_degree=0)
    rig = synthetic.orbit_rig(2, 320, 240, 277.0, seed=31, radius=(9.0, 12.0))
    dev = device()
    backend = CustomHipBackend(output_srgb=False, lookback_spin_limit=-1, max_capacity_retries=1)
    service = RendererService(backend, height=240, width=320, max_views=2)
    service.initialize(stage=None, device=dev)
    service.load_scene(0, **scene.to(dev).as_kwargs())
    outputs = service.prepare_outputs(2)
    with pytest.raises(RendererCapacityError, match="look-back") as error:
        service.render(rig.viewmats.to(dev), rig.intrinsics.to(dev), torch.zeros(2, dtype=torch.int64, device=dev),
                       outputs=outputs)
    counters = backend.read_counters()
    _print("forced give-up", counters)
    assert counters["sort_lookback_failures"] > 0
    assert error.value.telemetry["reason"] == "sort-lookback-timeout"
    assert bool(torch.isnan(outputs["rgb"]).all()) and bool((outputs["semantic_id"] == -1).all())


def test_small_lookback_spin_limit_is_correct_or_fails_closed():
    """A 1-poll bound on the look-back waits: each render either keeps the exact
    order (no give-up) or reports the give-up and fails closed, never silently wrong."""
    from vgr_amd import RendererCapacityError

    scene = synthetic.gaussian_scene(400_000, seed=20261202, sh_degree=0)
    rig = synthetic.orbit_rig(8, 640, 480, 554.0, seed=20261202, radius=(10.0, 15.0))
    expected = render_oracle([(0, scene)], rig, [0] * 8, want_sorted=False)
    outcomes = []
    for attempt in range(3):
        try:
            gpu, backend = render_service([(0, scene)], rig, [0] * 8, lookback_spin_limit=1, max_capacity_retries=0,
                                          max_visible_records=4_000_000, max_intersections=40_000_000)
        except RendererCapacityError as error:
            assert error.telemetry["reason"] == "sort-lookback-timeout"
            outcomes.append("failed-closed")
            continue
        assert backend.read_counters()["sort_lookback_failures"] == 0
        compare_images(gpu, expected)
        outcomes.append("exact")
    _print("spin limit 1", outcomes)
