"""INRIA 3D Gaussian Splatting PLY scenes: reading, writing, canonicalisation.

Same public functions and results as the reference loader
(``src/isaacsim_gaussian_renderer/ply_loader.py:16-144``): the vertex element's
``x y z``, ``f_dc_0..2``, ``f_rest_*``, ``opacity``, ``scale_*``, ``rot_*`` and
optional ``semantic_id`` properties become a raw :class:`GaussianScene`, and
:func:`canonicalize_3dgs_scene` activates it (``exp`` scales, ``sigmoid``
opacity, unit WXYZ quaternions with the identity for a zero quaternion, base RGB
``clamp(0.5 + C0 * dc, 0, 1)``).  SH coefficients are stored band-major,
channel-minor: ``features[g, 3 k + c]``; the file keeps ``f_rest`` channel-major
(all bands of R, then G, then B), so coefficient ``c * K' + (k - 1)`` of
``f_rest`` moves to ``features[g, 3 k + c]`` (K' = bands - 1).

The reference parses with the ``plyfile`` package; this module has its own
parser: the header is read as text, a binary body (little or big endian) is
mapped straight into a NumPy structured array (no per-vertex Python work, so
multi-million-splat scenes load at disk speed), and ASCII bodies are parsed
line by line.  Only scalar properties are supported in the vertex element
(3DGS files have no list properties there).
"""

from __future__ import annotations

import mmap
from dataclasses import dataclass
from pathlib import Path

import numpy as np
import torch

from .scene import GaussianScene

SH_C0 = 0.28209479177387814  # real SH basis constant of band 0: 1 / (2 sqrt(pi))
_HEADER_LIMIT = 1 << 20

_SCALAR_TYPES = {
    "char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1",
    "short": "i2", "int16": "i2", "ushort": "u2", "uint16": "u2",
    "int": "i4", "int32": "i4", "uint": "u4", "uint32": "u4",
    "float": "f4", "float32": "f4", "double": "f8", "float64": "f8",
}


@dataclass
class _Element:
    name: str
    count: int
    properties: list[tuple[str, str | None]]  # (name, numpy scalar code, or None for a list property)


@dataclass
class _Header:
    format: str            # "ascii", "binary_little_endian" or "binary_big_endian"
    elements: list[_Element]
    body_offset: int


def _read_header(path: Path) -> _Header:
    with path.open("rb") as stream:
        if stream.readline().strip() != b"ply":
            raise ValueError("PLY header must begin with the 'ply' magic line.")
        fmt = None
        elements: list[_Element] = []
        while stream.tell() <= _HEADER_LIMIT:
            raw = stream.readline()
            if not raw:
                break
            try:
                words = raw.decode("ascii").split()
            except UnicodeDecodeError as error:
                raise ValueError("PLY header is not ASCII.") from error
            if not words or words[0] in ("comment", "obj_info"):
                continue
            keyword = words[0]
            if keyword == "end_header":
                if fmt is None:
                    raise ValueError("PLY header has no format line.")
                return _Header(fmt, elements, stream.tell())
            if keyword == "format":
                if len(words) != 3 or words[1] not in ("ascii", "binary_little_endian", "binary_big_endian"):
                    raise ValueError(f"Unsupported PLY format line: {raw!r}")
                fmt = words[1]
            elif keyword == "element":
                if len(words) != 3:
                    raise ValueError(f"Malformed PLY element line: {raw!r}")
                elements.append(_Element(words[1], int(words[2]), []))
            elif keyword == "property":
                if not elements:
                    raise ValueError("PLY property precedes every element.")
                if len(words) == 3 and words[1] in _SCALAR_TYPES:
                    elements[-1].properties.append((words[2], _SCALAR_TYPES[words[1]]))
                elif len(words) == 5 and words[1] == "list":
                    elements[-1].properties.append((words[4], None))
                else:
                    raise ValueError(f"Malformed PLY property line: {raw!r}")
            else:
                raise ValueError(f"Unknown PLY header keyword {keyword!r}.")
    raise ValueError("PLY header lacks end_header within the first MiB.")


def ply_vertex_property_names(path: str | Path) -> tuple[str, ...]:
    """Vertex property names from the header (the body is not read)."""
    header = _read_header(Path(path))
    for element in header.elements:
        if element.name == "vertex":
            return tuple(name for name, _ in element.properties)
    return ()


def _vertex_table(path: Path) -> np.ndarray:
    """The vertex element as a structured array (fields named after the properties)."""
    header = _read_header(path)
    names = [e.name for e in header.elements]
    if "vertex" not in names:
        raise ValueError("PLY file has no vertex element.")
    index = names.index("vertex")
    vertex = header.elements[index]
    if any(code is None for _, code in vertex.properties):
        raise ValueError("List properties in the vertex element are not supported.")
    if header.format == "ascii":
        return _ascii_vertices(path, header, index)
    order = "<" if header.format == "binary_little_endian" else ">"
    offset = header.body_offset
    for before in header.elements[:index]:
        if any(code is None for _, code in before.properties):
            raise ValueError(f"Cannot skip element {before.name!r}: it has list properties.")
        offset += before.count * np.dtype([(n, order + c) for n, c in before.properties]).itemsize
    dtype = np.dtype([(n, order + c) for n, c in vertex.properties])
    needed = offset + vertex.count * dtype.itemsize
    if path.stat().st_size < needed:
        raise ValueError(f"PLY body is truncated: {path.stat().st_size} bytes, the header describes {needed}.")
    with path.open("rb") as stream, mmap.mmap(stream.fileno(), 0, access=mmap.ACCESS_READ) as mapped:
        table = np.frombuffer(mapped, dtype=dtype, count=vertex.count, offset=offset).copy()
    return table


def _ascii_vertices(path: Path, header: _Header, index: int) -> np.ndarray:
    vertex = header.elements[index]
    dtype = np.dtype([(n, c) for n, c in vertex.properties])
    skip = sum(e.count for e in header.elements[:index])
    rows = []
    with path.open("rb") as stream:
        stream.seek(header.body_offset)
        for _ in range(skip):
            stream.readline()
        for _ in range(vertex.count):
            words = stream.readline().split()
            if len(words) != len(vertex.properties):
                raise ValueError("ASCII PLY vertex line does not match the header.")
            rows.append(tuple(float(w) if dtype[i].kind == "f" else int(w) for i, w in enumerate(words)))
    return np.array(rows, dtype=dtype)


def _numbered(names: list[str], prefix: str) -> list[str]:
    return sorted((n for n in names if n.startswith(prefix)), key=lambda n: int(n.rsplit("_", 1)[1]))


def load_ply_to_gaussians(path: str | Path) -> GaussianScene:
    """Raw 3DGS parameters of a PLY file as CPU tensors (log scales, opacity logits, raw SH)."""
    table = _vertex_table(Path(path))
    names = list(table.dtype.names or ())
    count = table.shape[0]

    def columns(selected: list[str]) -> np.ndarray:
        if not selected:
            return np.zeros((count, 0), dtype=np.float32)
        return np.stack([table[name].astype(np.float32) for name in selected], axis=1)

    rest = _numbered(names, "f_rest_")
    if len(rest) % 3:
        raise ValueError(f"Expected an RGB-multiple number of f_rest fields, got {len(rest)}.")
    higher = len(rest) // 3
    # File order: channel-major [3, higher]; renderer order: band-major [higher, 3].
    rest_values = columns(rest).reshape(count, 3, higher).transpose(0, 2, 1).reshape(count, 3 * higher)
    features = np.concatenate((columns(["f_dc_0", "f_dc_1", "f_dc_2"]), rest_values), axis=1)
    semantic = table["semantic_id"].astype(np.int64) if "semantic_id" in names else np.zeros(count, np.int64)
    return GaussianScene(
        means=torch.from_numpy(np.ascontiguousarray(columns(["x", "y", "z"]))),
        scales=torch.from_numpy(np.ascontiguousarray(columns(_numbered(names, "scale_")))),
        rotations=torch.from_numpy(np.ascontiguousarray(columns(_numbered(names, "rot_")))),
        opacities=torch.from_numpy(np.ascontiguousarray(table["opacity"].astype(np.float32))),
        features=torch.from_numpy(np.ascontiguousarray(features)),
        semantic_ids=torch.from_numpy(np.ascontiguousarray(semantic)),
    )


def canonicalize_3dgs_scene(scene: GaussianScene, *, device: torch.device | str | None = None) -> GaussianScene:
    """Activate raw 3DGS parameters for rendering.

    Scales become ``exp(log_scale)``, opacities ``sigmoid(logit)``, rotations
    unit WXYZ quaternions (a quaternion with norm <= float32 epsilon becomes the
    identity), and ``features[:, :3]`` the linear base colour
    ``clamp(0.5 + C0 * dc, 0, 1)``; higher SH coefficients are kept as stored.
    """
    target = torch.device(device) if device is not None else scene.means.device

    def as_f32(tensor: torch.Tensor) -> torch.Tensor:
        return tensor.to(device=target, dtype=torch.float32)

    features = as_f32(scene.features).clone()
    if features.shape[1] < 3:
        raise ValueError("Raw 3DGS scene must contain the three DC SH channels.")
    features[:, :3] = torch.clamp(SH_C0 * features[:, :3] + 0.5, 0.0, 1.0)
    quats = as_f32(scene.rotations)
    eps = torch.finfo(torch.float32).eps
    norms = torch.linalg.vector_norm(quats, dim=1, keepdim=True)
    identity = torch.tensor([1.0, 0.0, 0.0, 0.0], device=target)
    unit = torch.where(norms > eps, quats / torch.clamp(norms, min=eps), identity)
    return GaussianScene(
        means=as_f32(scene.means).contiguous(),
        scales=torch.exp(as_f32(scene.scales)).contiguous(),
        rotations=unit.contiguous(),
        opacities=torch.sigmoid(as_f32(scene.opacities)).contiguous(),
        features=features.contiguous(),
        semantic_ids=scene.semantic_ids.to(device=target, dtype=torch.int64).contiguous(),
    )


def load_ply_to_canonical_gaussians(path: str | Path, *, device: torch.device | str | None = None) -> GaussianScene:
    """A standard 3DGS PLY as renderer-ready canonical tensors."""
    return canonicalize_3dgs_scene(load_ply_to_gaussians(path), device=device)


def write_3dgs_ply(path: str | Path, scene: GaussianScene, *, semantic: bool = True) -> None:
    """Write raw 3DGS parameters as a binary little-endian PLY (inverse of :func:`load_ply_to_gaussians`)."""
    means = scene.means.detach().cpu().numpy().astype("<f4")
    count = means.shape[0]
    features = scene.features.detach().cpu().numpy().astype("<f4")
    if features.shape[1] % 3:
        raise ValueError("features must hold whole RGB triples")
    higher = features.shape[1] // 3 - 1
    rest = features[:, 3:].reshape(count, higher, 3).transpose(0, 2, 1).reshape(count, 3 * higher)
    columns: list[tuple[str, np.ndarray, str]] = [
        *((axis, means[:, i], "f4") for i, axis in enumerate("xyz")),
        *((f"f_dc_{i}", features[:, i], "f4") for i in range(3)),
        *((f"f_rest_{i}", rest[:, i], "f4") for i in range(3 * higher)),
        ("opacity", scene.opacities.detach().cpu().numpy().astype("<f4"), "f4"),
        *((f"scale_{i}", scene.scales.detach().cpu().numpy()[:, i].astype("<f4"), "f4") for i in range(3)),
        *((f"rot_{i}", scene.rotations.detach().cpu().numpy()[:, i].astype("<f4"), "f4") for i in range(4)),
    ]
    if semantic:
        columns.append(("semantic_id", scene.semantic_ids.detach().cpu().numpy().astype("<i4"), "i4"))
    table = np.empty(count, dtype=[(name, "<" + code) for name, _, code in columns])
    for name, values, _ in columns:
        table[name] = values
    ply_type = {"f4": "float", "i4": "int"}
    header = ["ply", "format binary_little_endian 1.0", f"element vertex {count}"]
    header += [f"property {ply_type[code]} {name}" for name, _, code in columns]
    header.append("end_header")
    with Path(path).open("wb") as stream:
        stream.write(("\n".join(header) + "\n").encode("ascii"))
        stream.write(table.tobytes())
