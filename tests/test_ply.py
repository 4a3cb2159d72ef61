"""PLY ingestion and 3DGS canonicalisation (reference ply_loader.py:16-144).

Known answers of the reference's tests/test_ply_loader.py:16-76 (two-vertex
file: means, identity rotation, semantic ids, feature width; canonicalisation
of exp / sigmoid / quaternion / base colour), plus byte-level fixtures written
by hand with ``struct`` (so the reader is not checked only against this
module's own writer), the f_rest channel-major -> band-major reordering,
ASCII and big-endian bodies, and malformed files.
"""

from __future__ import annotations

import math
import struct
from pathlib import Path

import numpy as np
import pytest
import torch

from vgr_amd.ply import (
    SH_C0,
    canonicalize_3dgs_scene,
    load_ply_to_canonical_gaussians,
    load_ply_to_gaussians,
    ply_vertex_property_names,
    write_3dgs_ply,
)
from vgr_amd.scene import GaussianScene

STANDARD = ["x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2", "opacity", "scale_0", "scale_1", "scale_2",
            "rot_0", "rot_1", "rot_2", "rot_3"]


def write_bytes(path: Path, names: list[str], rows: list[list[float]], *, endian: str = "<",
                extra_int: str | None = None, extra_values: list[int] | None = None, fmt: str | None = None,
                prelude: bytes = b"") -> None:
    fmt = fmt or ("binary_little_endian" if endian == "<" else "binary_big_endian")
    header = ["ply", f"format {fmt} 1.0", "comment written by hand in the test", f"element vertex {len(rows)}"]
    header += [f"property float {n}" for n in names]
    if extra_int:
        header.append(f"property int {extra_int}")
    header.append("end_header")
    body = b""
    for i, row in enumerate(rows):
        if fmt == "ascii":
            words = [repr(float(v)) for v in row] + ([str(extra_values[i])] if extra_int else [])
            body += (" ".join(words) + "\n").encode()
        else:
            body += struct.pack(endian + "f" * len(row), *row)
            if extra_int:
                body += struct.pack(endian + "i", extra_values[i])
    path.write_bytes(("\n".join(header) + "\n").encode() + prelude + body)


def standard_rows():
    rows = [[0.0] * len(STANDARD) for _ in range(2)]
    rows[0][0], rows[1][0] = 1.0, 2.0
    rows[0][STANDARD.index("rot_0")] = rows[1][STANDARD.index("rot_0")] = 1.0
    return rows


def test_load_standard_3dgs_ply_known_answer(tmp_path):
    path = tmp_path / "scene.ply"
    write_bytes(path, STANDARD, standard_rows(), extra_int="semantic_id", extra_values=[4, 5])
    scene = load_ply_to_gaussians(path)
    assert "semantic_id" in ply_vertex_property_names(path)
    torch.testing.assert_close(scene.means[:, 0], torch.tensor([1.0, 2.0]))
    assert scene.features.shape == (2, 3)
    torch.testing.assert_close(scene.rotations[:, 0], torch.ones(2))
    torch.testing.assert_close(scene.semantic_ids, torch.tensor([4, 5]))
    assert scene.semantic_ids.dtype == torch.int64 and scene.means.dtype == torch.float32


def test_missing_semantic_ids_default_to_zero(tmp_path):
    path = tmp_path / "plain.ply"
    write_bytes(path, STANDARD, standard_rows())
    assert load_ply_to_gaussians(path).semantic_ids.tolist() == [0, 0]


def test_sh_rest_is_reordered_from_channel_major_to_band_major(tmp_path):
    # Degree 1: 3 higher bands.  File: f_rest_0..2 = R of bands 1..3, 3..5 = G, 6..8 = B.
    names = STANDARD + [f"f_rest_{i}" for i in range(9)]
    row = [0.0] * len(STANDARD) + [10.0, 11.0, 12.0, 20.0, 21.0, 22.0, 30.0, 31.0, 32.0]
    row[STANDARD.index("f_dc_0"):STANDARD.index("f_dc_2") + 1] = [1.0, 2.0, 3.0]
    path = tmp_path / "sh1.ply"
    write_bytes(path, names, [row])
    features = load_ply_to_gaussians(path).features[0].tolist()
    # features[3k + c]: band-major, channel-minor.
    assert features == [1.0, 2.0, 3.0, 10.0, 20.0, 30.0, 11.0, 21.0, 31.0, 12.0, 22.0, 32.0]


def test_canonicalize_known_answer():
    raw = GaussianScene(
        means=torch.tensor([[1.0, 2.0, 3.0], [4.0, 5.0, 6.0]]),
        scales=torch.tensor([[0.0, 0.1, -0.2], [0.3, -0.4, 0.5]]),
        rotations=torch.tensor([[2.0, 0.0, 0.0, 0.0], [0.0, 0.0, 0.0, 0.0]]),
        opacities=torch.tensor([0.0, 1.0]),
        features=torch.tensor([[0.0, 1.0, -1.0, 9.0], [2.0, -2.0, 0.5, 8.0]]),
        semantic_ids=torch.tensor([4, 5], dtype=torch.int64),
    )
    canonical = canonicalize_3dgs_scene(raw)
    torch.testing.assert_close(canonical.scales, raw.scales.exp())
    torch.testing.assert_close(canonical.opacities, raw.opacities.sigmoid())
    torch.testing.assert_close(canonical.rotations, torch.tensor([[1.0, 0.0, 0.0, 0.0], [1.0, 0.0, 0.0, 0.0]]))
    torch.testing.assert_close(canonical.features[:, :3], (0.5 + SH_C0 * raw.features[:, :3]).clamp(0.0, 1.0))
    assert canonical.features[:, 3].tolist() == [9.0, 8.0]  # higher coefficients kept as stored
    assert canonical.features.is_contiguous()
    # Closed forms: C0 = 1 / (2 sqrt(pi)); dc = 0 gives grey 0.5; dc = 0.5 gives 0.5 + C0 / 2;
    # dc = 2 gives 0.5 + 2 C0 = 1.064 (clamps to 1); dc = -2 gives -0.064 (clamps to 0).
    assert SH_C0 == pytest.approx(1.0 / (2.0 * math.sqrt(math.pi)), rel=1e-15)
    assert canonical.features[0, 0].item() == pytest.approx(0.5)
    assert canonical.features[1, 2].item() == pytest.approx(0.5 + 0.5 * SH_C0, rel=1e-6)
    assert canonical.features[1, 0].item() == 1.0
    assert canonical.features[1, 1].item() == 0.0


def test_canonical_quaternion_is_unit_and_sign_preserving():
    q = torch.tensor([[0.0, 3.0, 0.0, 4.0], [-1.0, 1.0, -1.0, 1.0], [1e-9, 0.0, 0.0, 0.0]])
    raw = GaussianScene(torch.zeros(3, 3), torch.zeros(3, 3), q, torch.zeros(3), torch.zeros(3, 3),
                        torch.zeros(3, dtype=torch.int64))
    out = canonicalize_3dgs_scene(raw).rotations
    torch.testing.assert_close(out[0], torch.tensor([0.0, 0.6, 0.0, 0.8]))
    torch.testing.assert_close(out[1], torch.tensor([-0.5, 0.5, -0.5, 0.5]))
    torch.testing.assert_close(out[2], torch.tensor([1.0, 0.0, 0.0, 0.0]))


@pytest.mark.parametrize("fmt", ["ascii", "binary_big_endian"])
def test_ascii_and_big_endian_bodies(tmp_path, fmt):
    path = tmp_path / f"{fmt}.ply"
    rows = standard_rows()
    rows[1][STANDARD.index("opacity")] = -1.25
    write_bytes(path, STANDARD, rows, endian=">" if fmt == "binary_big_endian" else "<", fmt=fmt,
                extra_int="semantic_id", extra_values=[7, -3])
    scene = load_ply_to_gaussians(path)
    torch.testing.assert_close(scene.means[:, 0], torch.tensor([1.0, 2.0]))
    assert scene.opacities.tolist() == [0.0, -1.25]
    assert scene.semantic_ids.tolist() == [7, -3]


def test_write_then_load_round_trip(tmp_path):
    g = torch.Generator().manual_seed(5)
    count, degree = 257, 3
    scene = GaussianScene(torch.randn(count, 3, generator=g), torch.randn(count, 3, generator=g),
                          torch.randn(count, 4, generator=g), torch.randn(count, generator=g),
                          torch.randn(count, 3 * (degree + 1) ** 2, generator=g),
                          torch.randint(0, 100, (count,), generator=g))
    path = tmp_path / "round.ply"
    write_3dgs_ply(path, scene)
    loaded = load_ply_to_gaussians(path)
    for name in ("means", "scales", "rotations", "opacities", "features", "semantic_ids"):
        assert torch.equal(getattr(loaded, name), getattr(scene, name)), name
    canonical = load_ply_to_canonical_gaussians(path)
    torch.testing.assert_close(canonical.scales, scene.scales.exp())


def test_elements_before_the_vertex_element_are_skipped(tmp_path):
    path = tmp_path / "skip.ply"
    header = ["ply", "format binary_little_endian 1.0", "element meta 2", "property double v",
              "element vertex 1"] + [f"property float {n}" for n in STANDARD] + ["end_header"]
    row = [5.0] + [0.0] * (len(STANDARD) - 1)
    body = struct.pack("<dd", 1.0, 2.0) + struct.pack("<" + "f" * len(STANDARD), *row)
    path.write_bytes(("\n".join(header) + "\n").encode() + body)
    assert load_ply_to_gaussians(path).means[0, 0].item() == 5.0


@pytest.mark.parametrize("case", ["magic", "truncated", "list", "no-end", "rest-not-rgb"])
def test_malformed_files_are_rejected(tmp_path, case):
    path = tmp_path / f"{case}.ply"
    if case == "magic":
        path.write_bytes(b"plx\nformat ascii 1.0\nend_header\n")
    elif case == "truncated":
        write_bytes(path, STANDARD, standard_rows())
        path.write_bytes(path.read_bytes()[:-5])
    elif case == "list":
        path.write_bytes(b"ply\nformat binary_little_endian 1.0\nelement vertex 1\n"
                         b"property list uchar int idx\nend_header\n\x01\x00\x00\x00\x00")
    elif case == "no-end":
        path.write_bytes(b"ply\nformat ascii 1.0\nelement vertex 0\n")
    else:
        write_bytes(path, STANDARD + ["f_rest_0", "f_rest_1"], [[0.0] * (len(STANDARD) + 2)])
    with pytest.raises(ValueError):
        load_ply_to_gaussians(path)


def test_canonical_scene_renders_like_the_in_memory_scene(tmp_path):
    """A scene written to disk and read back canonicalises to bit-identical render inputs."""
    from oracle import oracle
    from vgr_amd import synthetic

    g = torch.Generator().manual_seed(9)
    count = 2000
    raw = GaussianScene(3.0 * torch.randn(count, 3, generator=g), -4.0 + 0.5 * torch.randn(count, 3, generator=g),
                        torch.randn(count, 4, generator=g), torch.randn(count, generator=g),
                        torch.randn(count, 12, generator=g), torch.randint(0, 64, (count,), generator=g))
    path = tmp_path / "scene.ply"
    write_3dgs_ply(path, raw)
    from_disk = load_ply_to_canonical_gaussians(path)
    in_memory = canonicalize_3dgs_scene(raw)
    rig = synthetic.orbit_rig(2, 64, 48, 55.0, seed=9, radius=(9.0, 11.0))
    renders = [oracle.render([s.tensors()], [0], rig.viewmats, rig.intrinsics, [0, 0], height=48, width=64,
                             sh_degree=1) for s in (from_disk, in_memory)]
    assert int(renders[0]["counters"][0]) > 0
    for name in ("rgb", "alpha", "depth", "semantic_id"):
        np.testing.assert_array_equal(renders[0][name], renders[1][name])
