"""vgr_render_static_scratch (the allocation-free entry point) against vgr_render.

Both entry points are called through the C ABI on the same seeded inputs: the
existing one with a caller-owned scratch tensor, the new one with the library's
per-device scratch array.  Every output and the counter block must be
bit-identical, on two calls in a row, the second with a smaller input (fewer
Gaussians, fewer cameras, lower resolution), so the second call runs on the
array the first one left behind.  A render whose scratch requirement exceeds
the array is rejected before any launch.
"""

from __future__ import annotations

import pytest
import torch

from vgr_amd import native, synthetic

pytestmark = pytest.mark.gpu


def _inputs(count: int, cameras: int, width: int, height: int, seed: int):
    dev = torch.device("cuda", 0)
    scene = synthetic.gaussian_scene(count, seed=seed, sh_degree=1).to(dev)
    rig = synthetic.orbit_rig(cameras, width, height, 0.9 * width, seed=seed, radius=(9.0, 12.0))
    module = native.load_native_module()
    covariances = torch.empty((count, 6), device=dev)
    module.precompute_covariances(scene.scales, scene.rotations, covariances)
    scene_list = [scene.means, covariances, scene.opacities, scene.features, scene.semantic_ids,
                  torch.zeros(1, dtype=torch.int64, device=dev), torch.tensor([0, count], dtype=torch.int64, device=dev)]
    camera_list = [rig.viewmats.to(dev).contiguous(), rig.intrinsics.to(dev).contiguous(),
                   torch.eye(4, device=dev).repeat(cameras, 1, 1).contiguous(),
                   torch.zeros(cameras, dtype=torch.int64, device=dev), torch.empty(0, dtype=torch.int64, device=dev)]
    return scene_list, camera_list


def _render(scene_list, camera_list, cameras: int, width: int, height: int, *, static_scratch: bool,
            visible_capacity: int = 400_000, intersection_capacity: int = 1_200_000):
    dev = torch.device("cuda", 0)
    module = native.load_native_module()
    tiles = cameras * ((width + 15) // 16) * ((height + 15) // 16)
    outputs = [torch.full((cameras, height, width, 3), 7.0, device=dev),
               torch.full((cameras, height, width, 1), 7.0, device=dev),
               torch.full((cameras, height, width, 1), 7.0, device=dev),
               torch.full((cameras, height, width, 1), 7, dtype=torch.int64, device=dev)]
    scratch_bytes = module.scratch_bytes(max_views=cameras, height=height, width=width,
                                         gaussians=int(scene_list[0].shape[0]),
                                         intersection_capacity=intersection_capacity, tile_size=16,
                                         near_plane=0.01, far_plane=100.0)
    workspace = [torch.empty((visible_capacity, native.RECORD_FLOATS), device=dev),
                 torch.empty(intersection_capacity, dtype=torch.uint64, device=dev),
                 torch.empty(intersection_capacity, dtype=torch.uint64, device=dev),
                 torch.empty(intersection_capacity, dtype=torch.int32, device=dev),
                 torch.empty(intersection_capacity, dtype=torch.int32, device=dev),
                 torch.empty(tiles, dtype=torch.int32, device=dev),
                 torch.empty(tiles, dtype=torch.int32, device=dev),
                 torch.full((native.COUNTER_SLOTS,), -1, dtype=torch.int64, device=dev),
                 (torch.empty(0, dtype=torch.uint8, device=dev) if static_scratch
                  else torch.empty(scratch_bytes, dtype=torch.uint8, device=dev))]
    module.render(scene_list, camera_list, outputs, workspace, height, width, int(scene_list[0].shape[0]), 0.01,
                  100.0, 3.33, 0.3, False, 0.01, False, 16, 4096, 64, False, False, False, False, True, True,
                  False, sh_degree=1, static_scratch=static_scratch)
    torch.cuda.synchronize()
    return [t.cpu() for t in outputs], workspace[7].cpu(), scratch_bytes


@pytest.mark.parametrize("order", ["static_first", "caller_first"])
def test_static_scratch_matches_caller_scratch_on_two_calls(order):
    module = native.load_native_module()
    limit = module.static_scratch_bytes()
    assert limit == 64 << 20
    # Second call smaller in every dimension: it runs on what the first one left in the array.
    for count, cameras, width, height, seed in ((40_000, 4, 192, 128, 11), (9_000, 2, 96, 80, 12)):
        print(f"[no-alloc] {order}: {count} Gaussians, {cameras} cameras {width}x{height}", flush=True)
        scene_list, camera_list = _inputs(count, cameras, width, height, seed)
        runs = {}
        for static in ((True, False) if order == "static_first" else (False, True)):
            runs[static] = _render(scene_list, camera_list, cameras, width, height, static_scratch=static)
        (got, got_counters, needed), (want, want_counters, _) = runs[True], runs[False]
        assert needed <= limit
        assert int(want_counters[1]) > 0 and int(want_counters[5]) == 0
        assert torch.equal(got_counters, want_counters)
        for name, a, b in zip(("rgb", "depth", "alpha", "semantic_id"), got, want):
            assert torch.equal(a, b), name
        assert float(want[2].max()) > 0.5  # the scene is in view


def test_static_scratch_rejects_renders_beyond_its_size():
    module = native.load_native_module()
    scene_list, camera_list = _inputs(1_000, 1, 64, 64, 13)
    # 100M intersections need about 100 MB of sort look-back flags, beyond the 64 MiB array.
    with pytest.raises(RuntimeError, match="static scratch holds"):
        _render(scene_list, camera_list, 1, 64, 64, static_scratch=True, visible_capacity=1_000,
                intersection_capacity=100_000_000)
    assert module.scratch_bytes(max_views=1, height=64, width=64, gaussians=1_000, intersection_capacity=100_000_000,
                                tile_size=16, near_plane=0.01, far_plane=100.0) > module.static_scratch_bytes()
