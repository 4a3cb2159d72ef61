"""Bounded adaptive capacity: growth arithmetic and telemetry.

Growth rule (the reference's, ``src/isaacsim_gaussian_renderer/cuda_backend.py:
1536-1657``, Apache-2.0, pinned by its tests (4, 8) -> (15, 150) and
(100, 100) -> (100, 127)):

* demand = what the device counters prove is needed: for each buffer,
  max(counted, capacity + overflow);
* a buffer that overflowed grows to max(capacity + 1, ceil(capacity x growth),
  ceil(demand x headroom));
* when the visible buffer grows and the native counters only cover stored
  records, the intersection demand is extrapolated by
  tile_intersections x visible_demand / stored_records;
* the intersection buffer grows whenever it overflowed or the visible buffer
  grows;
* under a workspace-byte ceiling the headroom is dropped first (capacities
  fall back to the observed demand); if even that does not fit, the growth
  fails closed at installation.

:class:`CapacityLedger` keeps the provenance that ``capacity_stats``
publishes under the ``renderer-capacity/v1`` schema.
"""

from __future__ import annotations

import copy
import math
from collections.abc import Callable
from dataclasses import dataclass, field
from typing import Any

COUNTER_NAMES = (
    "visible_gaussians",
    "tile_intersections",
    "visible_overflow",
    "intersection_overflow",
    "active_tiles",
)
SORT_FAILURES = "sort_lookback_failures"


@dataclass(frozen=True)
class Capacity:
    visible_records: int
    tile_intersections: int

    def as_dict(self) -> dict[str, int]:
        return {"visible_records": self.visible_records, "tile_intersections": self.tile_intersections}


@dataclass(frozen=True)
class Demand:
    """Minimum capacities the counters of one render prove necessary."""

    visible: int
    intersections: int
    visible_short: bool
    intersections_short: bool

    @classmethod
    def from_counters(cls, counters: dict[str, int], current: Capacity) -> "Demand":
        return cls(
            visible=max(counters["visible_gaussians"], current.visible_records + counters["visible_overflow"]),
            intersections=max(counters["tile_intersections"],
                              current.tile_intersections + counters["intersection_overflow"]),
            visible_short=counters["visible_overflow"] > 0,
            intersections_short=counters["intersection_overflow"] > 0,
        )


@dataclass(frozen=True)
class GrowthPlan:
    target: Capacity
    ceiling_adjustment: dict[str, Any] | None = None


@dataclass(frozen=True)
class CapacityPolicy:
    growth_factor: float
    headroom: float
    max_workspace_bytes: int | None = None

    def _grown(self, capacity: int, demand: int, *, with_headroom: bool) -> int:
        if not with_headroom:
            return max(capacity + 1, demand)
        return max(capacity + 1, math.ceil(capacity * self.growth_factor), math.ceil(demand * self.headroom))

    def _target(self, current: Capacity, demand: Demand, counters: dict[str, int], *, full_demand: bool,
                with_headroom: bool) -> Capacity:
        visible = current.visible_records
        if demand.visible_short:
            visible = self._grown(current.visible_records, demand.visible, with_headroom=with_headroom)
        intersection_demand = demand.intersections
        if visible > current.visible_records and not full_demand:
            # Counters cover stored records only: scale the intersections of the
            # stored records up to the full visible demand.
            stored = max(1, min(counters["visible_gaussians"], current.visible_records))
            intersection_demand = max(intersection_demand,
                                      math.ceil(counters["tile_intersections"] * demand.visible / stored))
        intersections = current.tile_intersections
        if demand.intersections_short or visible > current.visible_records:
            intersections = self._grown(current.tile_intersections, intersection_demand, with_headroom=with_headroom)
        return Capacity(visible, intersections)

    def propose(self, current: Capacity, counters: dict[str, int], *, full_demand: bool,
                workspace_bytes: Callable[[Capacity], int | None]) -> GrowthPlan:
        """Growth after an overflowing render; ``workspace_bytes`` prices a capacity (None: beyond native limits)."""
        demand = Demand.from_counters(counters, current)
        target = self._target(current, demand, counters, full_demand=full_demand, with_headroom=True)
        if self.max_workspace_bytes is None:
            return GrowthPlan(target)
        target_bytes = workspace_bytes(target)
        if target_bytes is None or target_bytes <= self.max_workspace_bytes:
            return GrowthPlan(target)
        floor = self._target(current, demand, counters, full_demand=full_demand, with_headroom=False)
        floor_bytes = workspace_bytes(floor)
        fits = floor_bytes is not None and floor_bytes <= self.max_workspace_bytes
        return GrowthPlan(floor, {
            "policy": "relax-headroom-to-observed-demand",
            "target_capacity": target.as_dict(),
            "target_workspace_bytes": target_bytes,
            "minimum_observed_demand_capacity": floor.as_dict(),
            "minimum_observed_demand_workspace_bytes": floor_bytes,
            "max_workspace_bytes": self.max_workspace_bytes,
            "status": "headroom-relaxed" if fits else "minimum-observed-demand-exceeds-limit",
        })


def headroom_ratios(capacity: Capacity, counters: dict[str, int] | None) -> dict[str, float | None] | None:
    if counters is None:
        return None
    visible, intersections = counters["visible_gaussians"], counters["tile_intersections"]
    return {
        "visible_records_ratio": capacity.visible_records / visible if visible else None,
        "tile_intersections_ratio": capacity.tile_intersections / intersections if intersections else None,
    }


def overflowed(counters: dict[str, int]) -> bool:
    return bool(counters["visible_overflow"] or counters["intersection_overflow"])


@dataclass
class CapacityLedger:
    """Running provenance of validations, growth and allocations."""

    device_validations: int = 0
    validated_input_reuses: int = 0
    growth_events: int = 0
    retries: int = 0
    last_render: dict[str, Any] | None = None
    last_growth: dict[str, Any] | None = None
    last_allocation: dict[str, Any] | None = None
    attempts: list[dict[str, Any]] = field(default_factory=list)

    def reset(self) -> None:
        self.__init__()  # type: ignore[misc]

    def totals(self) -> dict[str, int]:
        return {"device_validations": self.device_validations, "validated_input_reuses": self.validated_input_reuses,
                "growth_events": self.growth_events, "retries": self.retries}

    def begin(self) -> None:
        self.attempts = []

    def attempt(self, number: int, capacity: dict[str, int], counters: dict[str, int], *, source: str,
                synchronized: bool) -> dict[str, Any]:
        record = {
            "attempt": number + 1,
            "requested_capacity": dict(capacity),
            "actual_demand": dict(counters),
            "counter_source": source,
            "synchronized": synchronized,
            "overflow": overflowed(counters),
        }
        if counters.get(SORT_FAILURES):
            record["sort_lookback_failures"] = counters[SORT_FAILURES]
        self.attempts.append(record)
        return record

    def outcome(self, status: str, validation: str, *, capacity: dict[str, int], counters: dict[str, int] | None,
                headroom: dict[str, float | None] | None, retry_count: int, failure_reason: str | None = None,
                attempts: list[dict[str, Any]] | None = None) -> dict[str, Any]:
        record: dict[str, Any] = {
            "status": status,
            "validation": validation,
            "synchronized": any(a["synchronized"] for a in (self.attempts if attempts is None else attempts)),
            "retry_count": retry_count,
            "attempts": copy.deepcopy(self.attempts if attempts is None else attempts),
            "final_capacity": dict(capacity),
            "final_counters": None if counters is None else dict(counters),
            "headroom": headroom,
        }
        if failure_reason is not None:
            record["failure_reason"] = failure_reason
        self.last_render = record
        if any("growth_request" in a for a in record["attempts"]):
            self.last_growth = copy.deepcopy(record)
        return record
