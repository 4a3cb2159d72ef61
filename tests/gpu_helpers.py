"""Helpers shared by the GPU parity tests: render through the drop-in service
and compare every output with the CPU oracle."""

from __future__ import annotations

import numpy as np
import torch

from oracle import oracle
from vgr_amd import CustomHipBackend, RendererService

RGB_ALPHA_TOLERANCE = 1.0e-4   # max-abs
DEPTH_RELATIVE_TOLERANCE = 1.0e-4


def device() -> torch.device:
    return torch.device("cuda", 0)


def render_service(scenes, rig, camera_scene_ids, *, active_ids=None, outputs=("rgb", "depth", "alpha",
                   "semantic_id"), env_xforms=None, **backend_kwargs):
    """Render ``rig`` through RendererService/CustomHipBackend; returns (outputs, backend)."""
    dev = device()
    backend_kwargs.setdefault("output_srgb", False)
    backend = CustomHipBackend(**backend_kwargs)
    height, width = rig.heights[0], rig.widths[0]
    service = RendererService(backend, height=height, width=width, outputs=outputs, max_views=rig.batch)
    service.initialize(stage=None, device=dev)
    for scene_id, scene in scenes:
        service.load_scene(scene_id, **scene.to(dev).as_kwargs())
    if env_xforms is not None:
        service.update_scene_transforms(torch.as_tensor(camera_scene_ids, dtype=torch.int64, device=dev),
                                        env_xforms.to(dev).contiguous())
    ids = None if active_ids is None else torch.as_tensor(active_ids, dtype=torch.int64, device=dev)
    result = service.render(rig.viewmats.to(dev), rig.intrinsics.to(dev),
                            torch.as_tensor(camera_scene_ids, dtype=torch.int64, device=dev), active_camera_ids=ids)
    service.synchronize()
    return {k: v.detach().cpu().numpy() for k, v in result.items()}, backend


def render_oracle(scenes, rig, camera_scene_ids, **kwargs):
    return oracle.render([s.as_kwargs() for _, s in scenes], [sid for sid, _ in scenes], rig.viewmats,
                         rig.intrinsics, camera_scene_ids, height=rig.heights[0], width=rig.widths[0], **kwargs)


def compare_images(gpu: dict, expected: dict, *, names=("rgb", "depth", "alpha", "semantic_id")) -> dict:
    report = {}
    if "rgb" in names:
        report["rgb_max_abs"] = float(np.abs(gpu["rgb"] - expected["rgb"]).max())
        assert report["rgb_max_abs"] <= RGB_ALPHA_TOLERANCE, report
    if "alpha" in names:
        report["alpha_max_abs"] = float(np.abs(gpu["alpha"] - expected["alpha"]).max())
        assert report["alpha_max_abs"] <= RGB_ALPHA_TOLERANCE, report
    if "depth" in names:
        g, e = gpu["depth"], expected["depth"]
        assert np.array_equal(np.isinf(g), np.isinf(e)), "depth background pixels differ"
        finite = np.isfinite(e)
        rel = np.abs(g[finite] - e[finite]) / np.maximum(np.abs(e[finite]), 1e-12)
        report["depth_max_rel"] = float(rel.max()) if rel.size else 0.0
        assert report["depth_max_rel"] <= DEPTH_RELATIVE_TOLERANCE, report
    if "semantic_id" in names:
        mismatches = int((gpu["semantic_id"] != expected["semantic_id"]).sum())
        report["semantic_mismatches"] = mismatches
        assert mismatches == 0, report
    for name in names:
        if name != "semantic_id":
            report[f"{name}_bitwise_equal_fraction"] = float(np.mean(gpu[name] == expected[name]))
    return report


def compare_ordering(backend, expected: dict) -> None:
    """Tile counts, sorted keys and per-tile Gaussian order must be bit-exact."""
    counters = backend.read_counters()
    assert counters["visible_gaussians"] == int(expected["counters"][0])
    assert counters["tile_intersections"] == int(expected["counters"][1])
    assert counters["active_tiles"] == int(expected["counters"][4])
    assert counters["visible_overflow"] == 0 and counters["intersection_overflow"] == 0
    ws = backend.workspace
    tiles = expected["tile_counts"].shape[0]
    starts = ws["tile_starts"][:tiles].cpu().numpy().astype(np.int64)
    ends = ws["tile_ends"][:tiles].cpu().numpy().astype(np.int64)
    np.testing.assert_array_equal(ends - starts, expected["tile_counts"])
    n = counters["tile_intersections"]
    keys = ws["keys_out"][:n].cpu().numpy()
    values = ws["values_out"][:n].cpu().numpy().astype(np.int64)
    if "sorted_keys" in expected:
        np.testing.assert_array_equal(keys, expected["sorted_keys"])
        gids = ws["visible_gaussian_ids"].cpu().numpy()[values]
        np.testing.assert_array_equal(gids, expected["sorted_gaussians"])
