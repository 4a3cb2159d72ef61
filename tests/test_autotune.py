"""Autotuner search and cache logic (CPU; the measurement is a synthetic cost)."""

from __future__ import annotations

import json

import pytest

from vgr_amd.autotune import Autotuner, workload_signature

SPACE = {"tile_size": (16, 8), "raster_batch": (256, 128, 512, 64), "sort_digit_bits": (8, 7, 6),
         "cameras_per_block": (8, 16, 4, 32)}


def cost(config):
    # separable cost with its minimum at (8, 128, 7, 16); 512-batches are unusable
    if config["raster_batch"] == 512:
        raise RuntimeError("does not fit")
    return ({16: 3.0, 8: 1.0}[config["tile_size"]] + {256: 2.0, 128: 0.5, 64: 1.0}[config["raster_batch"]]
            + {8: 1.0, 7: 0.2, 6: 0.7}[config["sort_digit_bits"]]
            + {8: 1.0, 16: 0.1, 4: 2.0, 32: 0.3}[config["cameras_per_block"]])


def test_coordinate_descent_finds_separable_minimum_and_caches(tmp_path):
    calls = []

    def measure(config):
        calls.append(dict(config))
        return cost(config)

    tuner = Autotuner(SPACE, cache_path=tmp_path / "cache.json", repeats=1)
    best = tuner.tune("key", measure)
    assert best == {"tile_size": 8, "raster_batch": 128, "sort_digit_bits": 7, "cameras_per_block": 16}
    # one start point plus every other value of every knob, each measured once
    assert len(calls) == 1 + sum(len(v) - 1 for v in SPACE.values())
    stored = json.loads((tmp_path / "cache.json").read_text())
    assert stored["key"]["config"] == best
    assert any(entry["ms"] == float("inf") for entry in stored["key"]["history"])
    calls.clear()
    assert Autotuner(SPACE, cache_path=tmp_path / "cache.json").tune("key", measure) == best
    assert calls == []  # cache hit: nothing measured
    assert Autotuner(SPACE, cache_path=tmp_path / "cache.json", repeats=1).tune("key", measure, force=True) == best
    assert calls


def test_all_unusable_configurations_raise(tmp_path):
    def measure(config):
        raise ValueError("no")

    with pytest.raises(RuntimeError):
        Autotuner(SPACE, cache_path=tmp_path / "c.json").tune("k", measure)


def test_signature_buckets_counts():
    a = workload_signature(cameras=1000, pixels=19_000_000, gaussians=4_900_000, sh_degree=1)
    b = workload_signature(cameras=1024, pixels=19_660_800, gaussians=5_000_000, sh_degree=1)
    assert a == b
    assert a != workload_signature(cameras=1024, pixels=19_660_800, gaussians=5_000_000, sh_degree=3)
