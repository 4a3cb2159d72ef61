"""Device workspace as one arena allocation.

Every buffer a render submission needs (visible records, the double-buffered
intersection keys and values, tile ranges, the counter block, kernel scratch,
identity transforms, chunking tables) lives in a single ``uint8`` device
allocation, carved into 256-byte-aligned views.  One allocation per capacity
change keeps the caching allocator out of the steady state, keeps the
buffers of one submission contiguous in HBM, and makes the footprint exact:
``nbytes`` is what the device holds for the workspace.

A :class:`WorkspacePlan` is the layout for one pair of capacities; it is pure
host arithmetic, so capacity policies can price a plan before allocating it.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import torch

ALIGNMENT = 256


@dataclass(frozen=True)
class BufferSpec:
    shape: tuple[int, ...]
    dtype: torch.dtype

    @property
    def nbytes(self) -> int:
        return math.prod(self.shape) * torch.empty((), dtype=self.dtype).element_size()


@dataclass(frozen=True)
class WorkspacePlan:
    """Named buffers and their offsets in one arena, plus the sort facts behind them."""

    buffers: dict[str, BufferSpec]
    offsets: dict[str, int]
    nbytes: int
    facts: dict[str, int | str]

    @classmethod
    def build(cls, buffers: dict[str, BufferSpec], facts: dict[str, int | str]) -> "WorkspacePlan":
        offsets: dict[str, int] = {}
        cursor = 0
        # Largest buffers first: the big streams start on their own 256-byte boundaries
        # and the small tables pack behind them.
        for name in sorted(buffers, key=lambda n: -buffers[n].nbytes):
            offsets[name] = cursor
            cursor += -(-buffers[name].nbytes // ALIGNMENT) * ALIGNMENT
        return cls(dict(buffers), offsets, max(cursor, ALIGNMENT), dict(facts))


class WorkspaceArena:
    """One device allocation holding every buffer of a :class:`WorkspacePlan`."""

    def __init__(self, plan: WorkspacePlan, device: torch.device) -> None:
        self.plan = plan
        self.device = device
        self.storage = torch.empty((plan.nbytes,), dtype=torch.uint8, device=device)

    def views(self) -> dict[str, torch.Tensor]:
        out: dict[str, torch.Tensor] = {}
        for name, spec in self.plan.buffers.items():
            start = self.plan.offsets[name]
            raw = self.storage.narrow(0, start, spec.nbytes)
            out[name] = raw.view(spec.dtype).view(spec.shape)
        return out

    @property
    def nbytes(self) -> int:
        return self.plan.nbytes
