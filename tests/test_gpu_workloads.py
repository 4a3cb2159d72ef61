"""Oracle parity at the benchmark workloads' own sizes (gfx950 only).

The GPU renders each workload in full; the CPU oracle renders a camera subset
of the same seeded inputs, and the subset is compared:
  * 64 cameras 640x480 around 1M SH3 Gaussians (deterministic plain path):
    cameras 0-7 in full, with per-tile intersection counts, the sorted key
    prefix and the per-tile Gaussian order bit-exact;
  * 256 cameras of 4 resolutions over 4 scenes of 200k-2M SH3 Gaussians
    (batched-persistent path): one camera of every (scene, resolution) pair;
  * 1024 cameras 160x120 in the 5M-Gaussian SH1 room, batched and plain: a
    camera subset spread over the trajectories.
Tolerances: RGB/alpha max-abs 1e-4, depth relative 1e-4, semantic IDs exact.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from gpu_helpers import compare_images, device
from oracle import oracle
from vgr_amd import CustomHipBackend, RendererService, synthetic
from vgr_amd.batched import BatchedRenderer, CameraSpec

pytestmark = pytest.mark.gpu


def _print(name, report):
    print(f"[workload] {name}: {report}", flush=True)


def test_orbit_64_cameras_1m_sh3_full_size_matches_oracle():
    scene, rig = synthetic.orbit_workload()
    dev = device()
    backend = CustomHipBackend(output_srgb=False, deterministic=True, visible_capacity_per_view=1_000_000,
                               intersections_per_visible=2.6)
    service = RendererService(backend, height=480, width=640, max_views=64)
    service.initialize(stage=None, device=dev)
    service.load_scene(0, **scene.to(dev).as_kwargs())
    out = service.render(rig.viewmats.to(dev), rig.intrinsics.to(dev), torch.zeros(64, dtype=torch.int64, device=dev))
    service.synchronize()
    subset = list(range(8))
    expected = oracle.render([scene.as_kwargs()], [0], rig.viewmats[subset], rig.intrinsics[subset], [0] * 8,
                             height=480, width=640, sh_degree=3, want_sorted=True)
    got = {name: tensor[:8].cpu().numpy() for name, tensor in out.items()}
    _print("orbit images", compare_images(got, expected))
    # Cameras 0-7 own global tiles [0, 8 * 1200): their lists open the sorted sequence.
    tiles = expected["tile_counts"].shape[0]
    ws = backend.workspace
    starts = ws["tile_starts"][:tiles].cpu().numpy().astype(np.int64)
    ends = ws["tile_ends"][:tiles].cpu().numpy().astype(np.int64)
    np.testing.assert_array_equal(ends - starts, expected["tile_counts"])
    n = int(expected["counters"][1])
    np.testing.assert_array_equal(ws["keys_out"][:n].cpu().numpy(), expected["sorted_keys"])
    values = ws["values_out"][:n].cpu().numpy().astype(np.int64)
    np.testing.assert_array_equal(ws["visible_gaussian_ids"].cpu().numpy()[values], expected["sorted_gaussians"])
    counters = backend.read_counters()
    assert counters["sort_lookback_failures"] == 0 and counters["intersection_overflow"] == 0
    _print("orbit counters", counters)


def test_uneven_four_scene_batch_matches_oracle():
    batch = synthetic.uneven_workload()
    dev = device()
    renderer = BatchedRenderer(device=dev)
    for scene_id, scene in batch.scenes:
        renderer.register_scene(scene_id, **scene.to(dev).as_kwargs())
    specs = [CameraSpec(sid, w, h) for sid, (w, h) in zip(batch.scene_ids, batch.resolutions)]
    renderer.plan(specs)
    counters = renderer.calibrate(batch.viewmats.to(dev), batch.intrinsics.to(dev))
    renderer.render()
    torch.cuda.synchronize()
    assert counters["sort_lookback_failures"] == 0
    chosen: dict[tuple[int, tuple[int, int]], int] = {}
    for index, (sid, resolution) in enumerate(zip(batch.scene_ids, batch.resolutions)):
        chosen.setdefault((sid, resolution), index)
    subset = sorted(chosen.values())
    assert len(subset) == 16, chosen
    expected = oracle.render([s.as_kwargs() for _, s in batch.scenes], [sid for sid, _ in batch.scenes],
                             batch.viewmats[subset], batch.intrinsics[subset], [batch.scene_ids[i] for i in subset],
                             height=1, width=1, resolutions=[batch.resolutions[i] for i in subset], sh_degree=3)
    offset = 0
    for local, camera in enumerate(subset):
        w, h = batch.resolutions[camera]
        view = renderer.camera_outputs(camera)
        got = {name: t.reshape(h * w, -1).cpu().numpy() for name, t in view.items()}
        want = {name: expected[name][offset:offset + w * h] for name in got}
        report = compare_images(got, want)
        offset += w * h
        _print(f"uneven camera {camera} scene {batch.scene_ids[camera]} {w}x{h}", report)


def test_robotics_room_batched_and_plain_match_oracle():
    scene, rig = synthetic.robotics_workload()
    dev = device()
    renderer = BatchedRenderer(device=dev)
    renderer.register_scene(0, **scene.to(dev).as_kwargs())
    renderer.plan([CameraSpec(0, 160, 120)] * rig.batch)
    counters = renderer.calibrate(rig.viewmats.to(dev), rig.intrinsics.to(dev))
    assert counters["sort_lookback_failures"] == 0
    batched = {name: t.clone() for name, t in renderer.render().items()}
    renderer.render_plain()
    torch.cuda.synchronize()
    plain = renderer.plain_groups[0]["outputs"]
    subset = [0, 37, 300, 511, 702, 1023]
    expected = oracle.render([scene.as_kwargs()], [0], rig.viewmats[subset], rig.intrinsics[subset],
                             [0] * len(subset), height=120, width=160, sh_degree=1)
    pixels = 160 * 120
    for local, camera in enumerate(subset):
        want = {name: expected[name][local] for name in ("rgb", "depth", "alpha", "semantic_id")}
        flat = {name: batched[name][camera * pixels:(camera + 1) * pixels].view(120, 160, -1).cpu().numpy()
                for name in want}
        _print(f"robotics batched camera {camera}", compare_images(flat, want))
        grid = {name: plain[name][camera].cpu().numpy() for name in want}
        _print(f"robotics plain camera {camera}", compare_images(grid, want))
