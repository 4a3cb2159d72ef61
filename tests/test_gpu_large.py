"""Full-size robotics batch: the global radix ordering in one submission beyond 2^30
intersections (31-bit look-back prefixes) against the tile-front ordering, bitwise.

1024 cameras 160x120 (fx = fy = 139) on 32 smooth trajectories through a
5M-Gaussian SH1 room: 1.30G tile intersections and 776M visible pairs in one
plain submission (about 70 GB of workspace on one MI355X).  The two orderings
share no sorting code, and both are pinned against the oracle at smaller sizes
(test_gpu_parity.py, test_gpu_batched.py); here they must agree bit for bit at
the size where the radix sort's prefixes pass 2^30.
"""

from __future__ import annotations

import pytest
import torch

from vgr_amd import synthetic
from vgr_amd.batched import BatchedRenderer, CameraSpec

pytestmark = pytest.mark.gpu


def test_global_ordering_beyond_two_to_the_thirty_matches_tile_front():
    dev = torch.device("cuda", 0)
    seed, width, height, focal = 20261205, 160, 120, 139.0
    scene = synthetic.gaussian_scene(5_000_000, seed=seed, sh_degree=1, semantic_classes=256, layout="room")
    rig = synthetic.trajectory_rig(32, 32, width, height, focal, seed=seed)
    renderer = BatchedRenderer(device=dev, max_visible_per_submission=1 << 30)
    renderer.register_scene(0, **scene.to(dev).as_kwargs())
    renderer.plan([CameraSpec(0, width, height)] * rig.batch)
    renderer.calibrate(rig.viewmats.to(dev), rig.intrinsics.to(dev))
    assert renderer.ordering == "front"
    assert renderer.plain_per_submission == rig.batch  # the global calibration fit one submission
    front = {name: tensor.clone() for name, tensor in renderer.render().items()}
    renderer.render_plain()
    counters = renderer.workspace["counters"].cpu().tolist()
    print(f"[large] global submission: {counters[0]} visible, {counters[1]} intersections", flush=True)
    assert counters[1] > 1 << 30 and counters[2] == 0 and counters[3] == 0 and counters[5] == 0
    group = renderer.plain_groups[0]
    for name in ("rgb", "depth", "alpha", "semantic_id"):
        assert torch.equal(front[name].view(rig.batch, height, width, -1), group["outputs"][name]), name
