"""Launch-parameter autotuner for the batched camera path.

Picks the intersection ordering (global radix sort or tile-front), the tile
size, the Gaussians staged per LDS batch of the compositor (``raster_batch``),
the radix digit width of the sort (``sort_digit_bits``) and the cameras
sharing one projection workgroup (``cameras_per_block``) by measurement on the
device, and caches the winner per (device, workload
signature) in a JSON file so later processes start tuned.

Search is coordinate descent over the knobs in a fixed order (each knob's
best value is kept before the next knob is swept), which needs about a dozen
timed configurations instead of the full product.  It starts from the library
defaults (the first value of every knob) and moves only for a gain above
SWITCH_MARGIN.  Every configuration is
measured as the median of a few timed renders after one untimed render that
also sizes its capacities.

There is no reference counterpart: the reference fixes these at launch
(tile size 16, 256-thread composite blocks, CUB's digit width;
``renderer_cuda.cu:1973-2305``, ``2463-2534``).
"""

from __future__ import annotations

import json
import os
import statistics
from pathlib import Path
from typing import Callable, Mapping, Sequence

import torch

DEFAULT_SPACE: dict[str, tuple[int, ...]] = {
    "front_ordering": (1, 0),  # 1: tile-front ordering where the plan allows it, 0: global radix sort
    "tile_size": (16, 8),
    "raster_batch": (0, 256, 128, 512, 64),  # 0: the library default
    "sort_digit_bits": (0, 8, 7, 6),  # 0: the library default (8, or 9 where that saves a pass)
    "cameras_per_block": (0, 32, 16, 8, 64),  # 0: the library default
}
# A candidate replaces the best configuration only when it is faster by more than
# this fraction, so timing noise does not move the search off the defaults.
SWITCH_MARGIN = 0.01

DEFAULT_CACHE = Path(os.environ.get("VGR_AUTOTUNE_CACHE", Path.home() / ".cache" / "vgr_amd" / "autotune.json"))


def workload_signature(*, cameras: int, pixels: int, gaussians: int, sh_degree: int, scenes: int = 1) -> str:
    """Coarse key: counts rounded to powers of two so near-identical workloads share an entry."""

    def bucket(value: int) -> int:
        return 1 << max(0, int(value - 1).bit_length()) if value > 0 else 0

    return f"cams{bucket(cameras)}-px{bucket(pixels)}-g{bucket(gaussians)}-sh{sh_degree}-s{scenes}"


def device_key(device: torch.device | None = None) -> str:
    if device is None or device.type != "cuda":
        return "cpu"
    props = torch.cuda.get_device_properties(device)
    return f"{props.name}-{props.multi_processor_count}cu"


class Autotuner:
    """Coordinate-descent search with a persistent cache.

    ``measure(config) -> milliseconds`` renders with ``config`` (a dict of
    knob values) and returns a time, or raises ``RuntimeError`` /
    ``ValueError`` for an unusable configuration (skipped).
    """

    def __init__(self, space: Mapping[str, Sequence[int]] | None = None, cache_path: Path | str | None = None,
                 repeats: int = 3) -> None:
        self.space = {k: tuple(v) for k, v in (space or DEFAULT_SPACE).items()}
        for name, values in self.space.items():
            if not values:
                raise ValueError(f"knob {name!r} has no candidate values")
        self.cache_path = Path(cache_path) if cache_path is not None else DEFAULT_CACHE
        self.repeats = max(1, int(repeats))
        self.history: list[tuple[dict[str, int], float]] = []

    # -- cache ---------------------------------------------------------------------------
    def _load(self) -> dict:
        try:
            return json.loads(self.cache_path.read_text())
        except (OSError, ValueError):
            return {}

    def _store(self, key: str, entry: dict) -> None:
        data = self._load()
        data[key] = entry
        try:
            self.cache_path.parent.mkdir(parents=True, exist_ok=True)
            tmp = self.cache_path.with_suffix(".tmp")
            tmp.write_text(json.dumps(data, indent=1, sort_keys=True))
            tmp.replace(self.cache_path)
        except OSError:
            pass  # a read-only home only loses the cache, not the result

    def cached(self, key: str) -> dict[str, int] | None:
        entry = self._load().get(key)
        if not isinstance(entry, dict) or not isinstance(entry.get("config"), dict):
            return None
        config = {k: int(v) for k, v in entry["config"].items() if k in self.space}
        return config if len(config) == len(self.space) else None

    # -- search --------------------------------------------------------------------------
    def _time(self, measure: Callable[[dict[str, int]], float], config: dict[str, int]) -> float:
        for seen, value in self.history:
            if seen == config:
                return value
        try:
            samples = [float(measure(dict(config))) for _ in range(self.repeats)]
            value = statistics.median(samples)
        except (RuntimeError, ValueError):
            value = float("inf")
        self.history.append((dict(config), value))
        return value

    def tune(self, key: str, measure: Callable[[dict[str, int]], float], *, force: bool = False,
             start: Mapping[str, int] | None = None) -> dict[str, int]:
        """Best configuration for ``key`` (from the cache unless ``force``)."""
        if not force:
            hit = self.cached(key)
            if hit is not None:
                return hit
        self.history = []
        best = {name: (start or {}).get(name, values[0]) for name, values in self.space.items()}
        best_time = self._time(measure, best)
        for name, values in self.space.items():
            for value in values:
                if value == best[name]:
                    continue
                candidate = dict(best)
                candidate[name] = value
                elapsed = self._time(measure, candidate)
                if elapsed < best_time * (1.0 - SWITCH_MARGIN):
                    best, best_time = candidate, elapsed
        if best_time == float("inf"):
            raise RuntimeError("no candidate configuration could be measured")
        self._store(key, {"config": best, "ms": best_time,
                          "history": [{"config": c, "ms": t} for c, t in self.history]})
        return best


def tune_batched(renderer, viewmats: torch.Tensor, intrinsics: torch.Tensor, *, tuner: Autotuner | None = None,
                 force: bool = False, capture: bool = True) -> dict[str, int]:
    """Tune a planned :class:`~vgr_amd.batched.BatchedRenderer` on one frame and leave it
    re-planned (and captured) with the winning configuration."""
    tuner = tuner or Autotuner(repeats=1)
    specs = list(renderer.cameras)
    packed = renderer._pack()
    key = "|".join((device_key(renderer.device), "batched", workload_signature(
        cameras=len(specs), pixels=sum(s.width * s.height for s in specs), gaussians=int(packed["means"].shape[0]),
        sh_degree=int(renderer.sh_degree), scenes=len({s.scene_id for s in specs}))))

    def replan(config: dict[str, int]) -> None:
        renderer.configure(**config)
        renderer.plan(specs, visible_capacity=getattr(renderer, "visible_capacity", None),
                      intersection_capacity=getattr(renderer, "intersection_capacity", None))
        renderer.calibrate(viewmats, intrinsics)

    def measure(config: dict[str, int]) -> float:
        replan(config)
        samples = []
        for _ in range(3):
            start = torch.cuda.Event(enable_timing=True)
            stop = torch.cuda.Event(enable_timing=True)
            start.record()
            renderer.render()
            stop.record()
            torch.cuda.synchronize(renderer.device)
            samples.append(start.elapsed_time(stop))
        return statistics.median(samples)

    best = tuner.tune(key, measure, force=force)
    replan(best)
    if capture:
        renderer.capture()
    return best
