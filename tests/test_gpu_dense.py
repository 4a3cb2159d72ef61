"""GPU parity of the dense per-pixel path (tile size 1) and small tiles (2, 4).

Every case renders through RendererService + CustomHipBackend with the
reference's knobs (tile_size, deterministic, depth_bucket_count,
depth_bucket_group_size, compact_projection_cache, materialize_projected_records)
and compares with the oracle at the same tile size and effective depth cutoff:
images within the tolerances (bitwise in practice), counters, per-tile (per-pixel
at tile size 1) intersection counts, sorted keys and Gaussian order bit-exact.
The images are also compared with the 16x16-tile render: the cutoff and the tile
size change the intersection lists, never the picture.
"""

from __future__ import annotations

import numpy as np
import pytest

from gpu_helpers import compare_images, compare_ordering, render_oracle, render_service
from vgr_amd import synthetic
from vgr_amd.native import dense_cutoff

pytestmark = pytest.mark.gpu

CASES = [
    # (name, backend knobs)
    ("tile4", {"tile_size": 4}),
    ("tile2", {"tile_size": 2}),
    ("tile1_deterministic", {"tile_size": 1, "deterministic": True}),
    ("tile1_grouped_cutoff", {"tile_size": 1}),
    ("tile1_cutoff_64x7", {"tile_size": 1, "depth_bucket_count": 64, "depth_bucket_group_size": 7}),
    ("tile1_compact", {"tile_size": 1, "compact_projection_cache": True, "materialize_projected_records": True}),
]


@pytest.mark.parametrize("name,knobs", CASES, ids=[c[0] for c in CASES])
def test_dense_and_small_tiles_match_oracle(name, knobs):
    scene = synthetic.gaussian_scene(30_000, seed=20261206, sh_degree=1)
    rig = synthetic.orbit_rig(3, 96, 72, 83.0, seed=20261206, radius=(8.0, 11.0))
    gpu, backend = render_service([(0, scene)], rig, [0, 0, 0], **knobs)
    buckets, group = dense_cutoff(knobs["tile_size"], knobs.get("depth_bucket_count", 4096),
                                  knobs.get("depth_bucket_group_size", 64),
                                  knobs.get("compact_projection_cache", False), knobs.get("deterministic", False))
    expected = render_oracle([(0, scene)], rig, [0, 0, 0], sh_degree=1, tile_size=knobs["tile_size"],
                             depth_bucket_count=buckets, depth_bucket_group_size=group, want_sorted=True)
    report = compare_images(gpu, expected)
    print(f"[dense] {name}: cutoff {buckets}/{group} intersections {int(expected['counters'][1])} {report}",
          flush=True)
    compare_ordering(backend, expected)
    tiled = render_oracle([(0, scene)], rig, [0, 0, 0], sh_degree=1, tile_size=16)
    for key in ("rgb", "depth", "alpha", "semantic_id"):
        assert np.array_equal(expected[key], tiled[key]), key


def test_depth_cutoff_shrinks_the_lists_of_a_closed_scene():
    """A dense room seen from inside: most pixels saturate, so the cutoff removes many
    pixel intersections while the images stay those of the uncut render."""
    scene = synthetic.gaussian_scene(200_000, seed=7, sh_degree=1, layout="room")
    rig = synthetic.trajectory_rig(1, 2, 80, 60, 70.0, seed=7)
    full, full_backend = render_service([(0, scene)], rig, [0] * rig.batch, tile_size=1, deterministic=True)
    cut, cut_backend = render_service([(0, scene)], rig, [0] * rig.batch, tile_size=1)
    full_count = full_backend.read_counters()["tile_intersections"]
    cut_count = cut_backend.read_counters()["tile_intersections"]
    print(f"[dense] room cutoff: {full_count} -> {cut_count} pixel intersections", flush=True)
    assert cut_count < full_count
    for key in ("rgb", "depth", "alpha", "semantic_id"):
        assert np.array_equal(cut[key], full[key]), key
    expected = render_oracle([(0, scene)], rig, [0] * rig.batch, sh_degree=1, tile_size=1,
                             depth_bucket_count=4096, depth_bucket_group_size=64)
    assert int(expected["counters"][1]) == cut_count
