"""Host-side contract of RendererService + CustomHipBackend, driven by fake
native modules on the CPU (the reference's test strategy,
tests/test_cuda_backend.py:19-297 and tests/test_renderer_service.py).

Covers: output defaults and contract errors, render cadence, scene packing,
argument forwarding, inactive-camera defaults, bounded grow/retry arithmetic
((4,8)->(15,150) and (100,100)->(100,127) as tests/test_cuda_backend.py:1615-1840),
fail-closed poisoning, allocator-failure restoration, validated-input reuse,
projection-cache hits, DoubleBuffer selector normalisation and chunked
logical batches.
"""

from __future__ import annotations

import json

import pytest
import torch

from vgr_amd import (
    COUNTER_NAMES,
    CustomCudaBackend,
    CustomHipBackend,
    RendererCapacityError,
    RendererService,
)
from vgr_amd.native import WORKSPACE_TENSORS


def scene_tensors(count: int, feature_width: int = 3, seed: int = 0) -> dict[str, torch.Tensor]:
    g = torch.Generator().manual_seed(seed)
    rotations = torch.zeros((count, 4))
    rotations[:, 0] = 1.0
    return {
        "means": torch.randn((count, 3), generator=g),
        "scales": torch.full((count, 3), 0.05),
        "rotations": rotations,
        "opacities": torch.full((count,), 0.5),
        "features": torch.rand((count, feature_width), generator=g),
        "semantic_ids": torch.arange(count, dtype=torch.int64),
    }


class FakeNativeModule:
    """Records every forwarded argument and writes recognisable outputs."""

    def __init__(self, counters: list[dict[str, int]] | None = None, selector: int = 1) -> None:
        self.render_calls = 0
        self.calls: list[dict] = []
        self.capacities: list[tuple[int, int]] = []
        self.output_pointers: list[int] = []
        self.reuse_projection_flags: list[bool] = []
        self.counter_sequence = list(counters or [])
        self.selector = selector
        self.chunk_prepares: list[tuple] = []
        self.aggregations = 0

    def scratch_bytes(self, **kwargs) -> int:
        return 1024

    def sort_temp_bytes(self, n_items: int, n_segments: int) -> int:
        return 1024

    def render(self, scene, cameras, outputs, workspace, height, width, max_scene_gaussians, near, far, sigma, eps,
               antialiased, semantic_min_alpha, ray, tile_size, dbc, dbg, compact, materialize, reuse, srgb,
               deterministic, full_sensor, fixed_capacity, **kwargs):
        assert len(scene) == 7 and len(cameras) == 5 and len(outputs) == 4
        assert len(workspace) == len(WORKSPACE_TENSORS)
        self.render_calls += 1
        records = workspace[0]
        keys = workspace[1]
        self.capacities.append((int(records.shape[0]), int(keys.numel())))
        self.output_pointers.append(outputs[0].data_ptr())
        self.reuse_projection_flags.append(bool(reuse))
        self.calls.append(dict(height=height, width=width, tile_size=tile_size, deterministic=deterministic,
                               full_sensor=full_sensor, eps=eps, antialiased=antialiased, srgb=srgb,
                               semantic_min_alpha=semantic_min_alpha, max_scene_gaussians=max_scene_gaussians,
                               fixed_capacity=fixed_capacity, cameras=[c.clone() for c in cameras],
                               scene=[s.clone() for s in scene], kwargs=dict(kwargs)))
        active = cameras[4]
        outputs[0][:] = 0.2 if active.numel() == 0 else outputs[0]
        if active.numel():
            for camera in active.tolist():
                if camera >= 0:
                    outputs[0][camera] = 0.2
        if full_sensor:
            outputs[1].fill_(3.0) if active.numel() == 0 else None
        counters = workspace[WORKSPACE_TENSORS.index("counters")]
        values = self.counter_sequence.pop(0) if self.counter_sequence else {
            "visible_gaussians": 3, "tile_intersections": 5, "visible_overflow": 0, "intersection_overflow": 0,
            "active_tiles": 1}
        block = [values[name] for name in COUNTER_NAMES] + [values.get("sort_lookback_failures", 0)]
        counters.copy_(torch.tensor(block, dtype=torch.int64))
        return self.selector

    def prepare_chunked_active_ids(self, active_ids, mask, chunk_ids, logical, physical):
        self.chunk_prepares.append((active_ids.clone(), logical, physical))
        mask.zero_()
        mask[active_ids] = 1
        slots = torch.arange(chunk_ids.numel())
        flat = chunk_ids.view(-1)
        valid = (slots < logical) & (mask[slots.clamp(max=logical - 1)] != 0)
        flat.copy_(torch.where(valid, slots % physical, torch.full_like(slots, -1)))

    def aggregate_chunk_counters(self, chunk_counters, counters, maxima, count):
        self.aggregations += 1
        counters.copy_(chunk_counters[:count].sum(dim=0))
        maxima.copy_(chunk_counters[:count].max(dim=0).values)


def make_service(native, *, batch=1, height=4, width=5, outputs=("rgb", "depth", "alpha", "semantic_id"),
                 **kwargs):
    backend = CustomHipBackend(native_module=native, allow_cpu_for_tests=True, **kwargs)
    service = RendererService(backend, height=height, width=width, max_views=batch, outputs=outputs,
                              allow_cpu_for_tests=True)
    service.initialize(stage=None, device="cpu")
    return service, backend


def cameras(batch: int):
    return (torch.eye(4).repeat(batch, 1, 1).contiguous(), torch.eye(3).repeat(batch, 1, 1).contiguous())


# -------------------------------------------------------------------- service contract

def test_alias_is_the_same_backend():
    assert CustomCudaBackend is CustomHipBackend


def test_service_requires_gpu_unless_allowed():
    service = RendererService(CustomHipBackend(native_module=FakeNativeModule()), height=2, width=2)
    with pytest.raises(ValueError, match="CUDA device"):
        service.initialize(stage=None, device="cpu")


def test_owned_outputs_default_values_and_reuse():
    service, _ = make_service(FakeNativeModule(), batch=2)
    outputs = service.prepare_outputs(2)
    assert torch.all(outputs["rgb"] == 0) and torch.all(outputs["depth"] == 0)
    assert torch.all(outputs["alpha"] == 0) and torch.all(outputs["semantic_id"] == -1)
    assert outputs["semantic_id"].dtype == torch.int64
    assert service.prepare_outputs(2) is outputs


@pytest.mark.parametrize("bad", ["dtype", "shape", "contiguity"])
def test_camera_contract_errors(bad):
    service, _ = make_service(FakeNativeModule(), batch=2)
    service.load_scene(1, **scene_tensors(3))
    view, k = cameras(2)
    if bad == "dtype":
        view = view.double()
    elif bad == "shape":
        k = torch.eye(3).repeat(3, 1, 1)
    else:
        view = view.transpose(1, 2)
    with pytest.raises(ValueError):
        service.render(view, k, torch.tensor([1, 1]))


def test_unregistered_scene_and_active_range_are_rejected_on_host():
    service, _ = make_service(FakeNativeModule(), batch=2)
    service.load_scene(1, **scene_tensors(3))
    view, k = cameras(2)
    with pytest.raises(ValueError, match="Unregistered"):
        service.render(view, k, torch.tensor([1, 9]))
    with pytest.raises(ValueError, match="outside"):
        service.render(view, k, torch.tensor([1, 1]), active_camera_ids=torch.tensor([2]))


def test_render_cadence_returns_previous_outputs():
    native = FakeNativeModule()
    backend = CustomHipBackend(native_module=native, allow_cpu_for_tests=True)
    service = RendererService(backend, height=4, width=5, max_views=1, render_every_n=3, allow_cpu_for_tests=True)
    service.initialize(stage=None, device="cpu")
    service.load_scene(1, **scene_tensors(3))
    view, k = cameras(1)
    first = service.render(view, k, torch.tensor([1]))
    assert service.render(view, k, torch.tensor([1])) is first
    assert service.render(view, k, torch.tensor([1])) is first
    service.render(view, k, torch.tensor([1]))
    assert native.render_calls == 2


# ------------------------------------------------------------------ backend forwarding

def test_packs_arbitrary_scene_ids_and_submits_once():
    native = FakeNativeModule()
    service, backend = make_service(native, batch=2, tile_size=8, deterministic=True, covariance_epsilon=0.1,
                                    semantic_min_alpha=0.05, output_srgb=False, sh_degree=0)
    service.load_scene(42, **scene_tensors(3))
    service.load_scene(7, **scene_tensors(5, seed=1))
    view, k = cameras(2)
    service.render(view, k, torch.tensor([7, 42]))
    assert native.render_calls == 1
    call = native.calls[0]
    registered, offsets = call["scene"][5], call["scene"][6]
    assert registered.tolist() == [42, 7] and offsets.tolist() == [0, 3, 8]
    assert call["max_scene_gaussians"] == 5
    assert (call["tile_size"], call["deterministic"], call["eps"], call["semantic_min_alpha"], call["srgb"]) == \
           (8, True, pytest.approx(0.1), pytest.approx(0.05), False)
    assert call["kwargs"]["sh_degree"] == 0
    # identity environment transforms and "all cameras" active list
    assert torch.equal(call["cameras"][2], torch.eye(4).repeat(2, 1, 1))
    assert call["cameras"][4].numel() == 0
    # covariances of identity rotations are diag(s^2)
    assert torch.allclose(call["scene"][1][:, [0, 3, 5]], torch.full((8, 3), 0.0025))


def test_feature_width_must_match_and_sh_degree_is_inferred():
    native = FakeNativeModule()
    service, backend = make_service(native)
    service.load_scene(1, **scene_tensors(3, feature_width=48))
    assert backend.resolved_sh_degree == 3
    with pytest.raises(ValueError, match="feature width"):
        service.load_scene(2, **scene_tensors(3, feature_width=12))


def test_rgb_only_and_partial_output_contracts():
    native = FakeNativeModule()
    service, _ = make_service(native, outputs=("rgb",))
    service.load_scene(1, **scene_tensors(3))
    view, k = cameras(1)
    service.render(view, k, torch.tensor([1]))
    assert native.calls[0]["full_sensor"] is False
    partial_service, _ = make_service(FakeNativeModule(), outputs=("rgb", "depth"))
    partial_service.load_scene(1, **scene_tensors(3))
    with pytest.raises(ValueError, match="RGB-only"):
        partial_service.render(view, k, torch.tensor([1]))


def test_active_subset_defines_inactive_outputs():
    native = FakeNativeModule()
    service, _ = make_service(native, batch=3)
    service.load_scene(1, **scene_tensors(3))
    view, k = cameras(3)
    outputs = service.prepare_outputs(3)
    for tensor in outputs.values():
        tensor.fill_(7)
    service.render(view, k, torch.tensor([1, 1, 1]), outputs=outputs, active_camera_ids=torch.tensor([1]))
    assert torch.all(outputs["rgb"][0] == 0) and torch.all(outputs["rgb"][1] == 0.2)
    assert torch.all(torch.isinf(outputs["depth"][0])) and torch.all(outputs["alpha"][2] == 0)
    assert torch.all(outputs["semantic_id"][0] == -1)
    assert native.calls[0]["cameras"][4].tolist() == [1]


def test_environment_transforms_are_forwarded_and_must_align():
    native = FakeNativeModule()
    service, _ = make_service(native, batch=2)
    service.load_scene(1, **scene_tensors(3))
    env = torch.eye(4).repeat(2, 1, 1).contiguous()
    env[:, 0, 3] = 2.5
    service.update_scene_transforms(torch.tensor([1, 1]), env)
    view, k = cameras(2)
    service.render(view, k, torch.tensor([1, 1]))
    assert torch.equal(native.calls[0]["cameras"][2], env)
    with pytest.raises(ValueError, match="one transform per render camera"):
        service.render(view[:1].contiguous(), k[:1].contiguous(), torch.tensor([1]))


def test_unsupported_reference_modes_are_rejected():
    with pytest.raises(ValueError, match="incompatible with ray_gaussian_evaluation"):
        CustomHipBackend(ray_gaussian_evaluation=True, rasterize_mode="antialiased")
    with pytest.raises(ValueError, match="8x8 or 16x16"):
        CustomHipBackend(ray_gaussian_evaluation=True, tile_size=4)
    assert CustomHipBackend(ray_gaussian_evaluation=True).bin_tile_size == 32
    assert CustomHipBackend(ray_gaussian_evaluation=True, tile_size=8).bin_tile_size == 8
    assert CustomHipBackend().bin_tile_size == 16
    with pytest.raises(ValueError, match="fixed_capacity_sort"):
        CustomHipBackend(fixed_capacity_sort=True, deterministic=True)


def test_dense_path_knobs_follow_the_reference_rules():
    """cuda_backend.py:156-191: tile sizes are powers of two in [1, 16]; the compact
    projection cache needs tile size 1, the screen-space model and nondeterministic
    mode; materialized records need the compact cache."""
    for tile_size in (1, 2, 4, 8, 16):
        CustomHipBackend(tile_size=tile_size)
    CustomHipBackend(tile_size=1, compact_projection_cache=True, materialize_projected_records=True)
    for kwargs, message in (({"tile_size": 3}, "power of two"), ({"tile_size": 32}, "power of two"),
                            ({"compact_projection_cache": True, "tile_size": 4}, "requires tile_size=1"),
                            ({"compact_projection_cache": True, "tile_size": 1, "deterministic": True},
                             "requires deterministic=False"),
                            ({"materialize_projected_records": True, "tile_size": 1},
                             "requires compact_projection_cache=True"),
                            ({"depth_bucket_count": 0}, "depth_bucket_count"),
                            ({"depth_bucket_group_size": 0}, "depth_bucket_group_size")):
        with pytest.raises(ValueError, match=message):
            CustomHipBackend(**kwargs)


def test_dense_cutoff_follows_the_reference_modes():
    """The cutoff runs at tile size 1 in nondeterministic mode (grouped) or with the
    compact cache (one pass over every bucket), renderer_cuda.cu:2888-3144."""
    from vgr_amd.native import dense_cutoff
    assert dense_cutoff(1, 4096, 64, False, False) == (4096, 64)
    assert dense_cutoff(1, 4096, 64, True, False) == (4096, 4096)
    assert dense_cutoff(1, 4096, 64, False, True) == (0, 0)
    assert dense_cutoff(1, 32, 64, False, False) == (32, 32)
    assert dense_cutoff(2, 4096, 64, False, False) == (0, 0)
    assert dense_cutoff(16, 4096, 64, True, False) == (0, 0)


@pytest.mark.parametrize("kwargs", [
    {"capacity_growth_factor": 1.0}, {"capacity_headroom": 0.9}, {"max_capacity_retries": -1},
    {"max_capacity_retries": True}, {"max_workspace_bytes": 0}, {"semantic_min_alpha": 1.5},
    {"covariance_epsilon": float("nan")}, {"near_plane": 0.0}, {"sort_digit_bits": 10}, {"sort_digit_bits": 3}, {"raster_batch": 96},
    {"cameras_per_block": 65}, {"sh_degree": 4},
])
def test_parameters_are_strictly_validated(kwargs):
    with pytest.raises(ValueError):
        CustomHipBackend(**kwargs)


# ------------------------------------------------------------------ adaptive capacity

def overflow(visible, intersections, visible_overflow, intersection_overflow):
    return {"visible_gaussians": visible, "tile_intersections": intersections, "visible_overflow": visible_overflow,
            "intersection_overflow": intersection_overflow, "active_tiles": 2}


def test_growth_arithmetic_and_validated_reuse():
    native = FakeNativeModule([overflow(12, 40, 8, 32), overflow(12, 40, 0, 0)])
    service, backend = make_service(native, max_visible_records=4, max_intersections=8,
                                    enable_projection_cache=True)
    service.load_scene(17, **scene_tensors(3))
    identity = id(service.workspace)
    view, k = cameras(1)
    scene_ids = torch.tensor([17])
    outputs = service.render(view, k, scene_ids)
    assert native.capacities == [(4, 8), (15, 150)]
    assert native.output_pointers[0] == native.output_pointers[1]
    assert id(service.workspace) == identity and service.workspace is backend.workspace
    stats = backend.capacity_stats
    assert stats["initial_capacity"] == {"visible_records": 4, "tile_intersections": 8}
    assert stats["current_capacity"]["visible_records"] == 15
    assert stats["current_capacity"]["tile_intersections"] == 150
    assert stats["totals"] == {"device_validations": 1, "validated_input_reuses": 0, "growth_events": 1,
                               "retries": 1}
    assert stats["last_allocation"]["allocation_strategy"] == "release-old-workspace-first"
    json.dumps(stats)
    reused = service.render(view, k, scene_ids, outputs=outputs)
    assert reused is outputs
    assert native.reuse_projection_flags == [False, False, True]
    assert backend.capacity_stats["last_render"]["validation"] == "validated-identical-input"
    view[0, 0, 3] = 1.0
    service.render(view, k, scene_ids, outputs=outputs)
    assert native.reuse_projection_flags[-1] is False
    assert backend.capacity_stats["totals"]["device_validations"] == 2


def test_growth_does_not_double_a_near_full_workspace():
    native = FakeNativeModule([overflow(80, 101, 0, 1), overflow(80, 101, 0, 0)])
    service, backend = make_service(native, max_visible_records=100, max_intersections=100)
    service.load_scene(17, **scene_tensors(3))
    view, k = cameras(1)
    service.render(view, k, torch.tensor([17]))
    assert native.capacities == [(100, 100), (100, 127)]
    assert backend.capacity_stats["last_growth"]["attempts"][0]["growth_request"] == {
        "visible_records": 100, "tile_intersections": 127}


def test_full_demand_counters_skip_intersection_extrapolation():
    native = FakeNativeModule([overflow(12, 40, 8, 32), overflow(12, 40, 0, 0)])
    native.counts_full_intersection_demand = True
    service, _ = make_service(native, max_visible_records=4, max_intersections=8)
    service.load_scene(17, **scene_tensors(3))
    view, k = cameras(1)
    service.render(view, k, torch.tensor([17]))
    assert native.capacities == [(4, 8), (15, 50)]


def test_retry_limit_fails_closed_with_poisoned_outputs():
    native = FakeNativeModule([overflow(50, 60, 46, 0)] * 10)
    service, backend = make_service(native, max_visible_records=4, max_intersections=8, max_capacity_retries=2)
    service.load_scene(17, **scene_tensors(3))
    view, k = cameras(1)
    outputs = service.prepare_outputs(1)
    with pytest.raises(RendererCapacityError, match="bounded retries") as error:
        service.render(view, k, torch.tensor([17]), outputs=outputs)
    assert native.render_calls == 3
    assert torch.all(torch.isnan(outputs["rgb"])) and torch.all(torch.isnan(outputs["depth"]))
    assert torch.all(outputs["semantic_id"] == -1)
    assert error.value.telemetry["reason"] == "retry-limit"
    assert backend.capacity_stats["last_render"]["status"] == "failure"


def test_workspace_ceiling_fails_closed():
    probe_native = FakeNativeModule()
    _, probe = make_service(probe_native, max_visible_records=4, max_intersections=8)
    limit = probe.workspace_bytes + 64
    native = FakeNativeModule([overflow(5000, 9000, 4996, 8992)])
    service, _ = make_service(native, max_visible_records=4, max_intersections=8, max_workspace_bytes=limit)
    service.load_scene(17, **scene_tensors(3))
    view, k = cameras(1)
    with pytest.raises(RendererCapacityError) as error:
        service.render(view, k, torch.tensor([17]))
    assert error.value.telemetry["reason"] == "growth-allocation-failed"


def test_allocator_failure_restores_released_workspace(monkeypatch):
    native = FakeNativeModule([overflow(12, 40, 8, 32)])
    service, backend = make_service(native, max_visible_records=4, max_intersections=8)
    service.load_scene(17, **scene_tensors(3))
    real_empty = torch.empty
    state = {"fail": True, "installed": backend.workspace_bytes}

    def failing_empty(shape, *args, **kwargs):
        # The grown workspace arena is the first allocation larger than the installed one.
        if (state["fail"] and isinstance(shape, tuple) and len(shape) == 1 and kwargs.get("dtype") == torch.uint8
                and shape[0] > state["installed"]):
            state["fail"] = False
            raise torch.OutOfMemoryError("injected")
        return real_empty(shape, *args, **kwargs)

    monkeypatch.setattr(torch, "empty", failing_empty)
    view, k = cameras(1)
    with pytest.raises(RendererCapacityError, match="allocator") as error:
        service.render(view, k, torch.tensor([17]))
    growth = error.value.telemetry["capacity"]["last_allocation"]
    assert growth["failure"] == "allocator-out-of-memory"
    assert growth["restoration"]["status"] == "restored-old-capacity"
    assert backend.current_visible_capacity == 4 and backend.workspace["visible_records"].shape[0] == 4


def test_selector_zero_swaps_sorted_buffers_without_copy():
    native = FakeNativeModule(selector=0)
    service, backend = make_service(native)
    service.load_scene(1, **scene_tensors(3))
    keys_in, keys_out = backend.workspace["keys_in"], backend.workspace["keys_out"]
    values_in = backend.workspace["values_in"]
    view, k = cameras(1)
    service.render(view, k, torch.tensor([1]))
    assert backend.workspace["keys_out"] is keys_in and backend.workspace["keys_in"] is keys_out
    assert backend.workspace["values_out"] is values_in


def test_chunked_logical_batch_preserves_order_and_active_ids():
    native = FakeNativeModule()
    service, backend = make_service(native, batch=5, max_physical_views=2, fixed_capacity_sort=True,
                                    adaptive_capacity=False)
    service.load_scene(1, **scene_tensors(3))
    view, k = cameras(5)
    for i in range(5):
        view[i, 0, 3] = float(i)
    service.render(view, k, torch.ones(5, dtype=torch.int64), active_camera_ids=torch.tensor([4, 1]))
    assert native.render_calls == 3
    assert [c["cameras"][0].shape[0] for c in native.calls] == [2, 2, 1]
    assert [c["cameras"][0][:, 0, 3].tolist() for c in native.calls] == [[0, 1], [2, 3], [4]]
    assert [c["cameras"][4].tolist() for c in native.calls] == [[-1, 1], [-1, -1], [0]]
    assert native.aggregations == 1
    assert backend.capacity_stats["execution_schedule"]["maximum_logical_chunks"] == 3
    with pytest.raises(ValueError):
        CustomHipBackend(max_physical_views=2, fixed_capacity_sort=True)  # adaptive capacity still on


# ------------------------------------------------------------ sort look-back give-ups

def sort_failure(count):
    return {"visible_gaussians": 3, "tile_intersections": 5, "visible_overflow": 0, "intersection_overflow": 0,
            "active_tiles": 1, "sort_lookback_failures": count}


def test_sort_lookback_failure_is_rendered_again_without_growth():
    native = FakeNativeModule([sort_failure(2), sort_failure(0)])
    service, backend = make_service(native, max_visible_records=4, max_intersections=8)
    service.load_scene(17, **scene_tensors(3))
    view, k = cameras(1)
    service.render(view, k, torch.tensor([17]))
    assert native.render_calls == 2
    assert native.capacities == [(4, 8), (4, 8)]  # re-rendered, not grown
    stats = backend.capacity_stats
    assert stats["totals"]["growth_events"] == 0 and stats["totals"]["retries"] == 1
    assert stats["last_render"]["status"] == "success"
    assert stats["last_render"]["attempts"][0]["sort_lookback_failures"] == 2


def test_persistent_sort_lookback_failure_fails_closed():
    native = FakeNativeModule([sort_failure(1)] * 10)
    service, backend = make_service(native, max_visible_records=4, max_intersections=8, max_capacity_retries=2)
    service.load_scene(17, **scene_tensors(3))
    view, k = cameras(1)
    outputs = service.prepare_outputs(1)
    with pytest.raises(RendererCapacityError, match="look-back") as error:
        service.render(view, k, torch.tensor([17]), outputs=outputs)
    assert native.render_calls == 3
    assert error.value.telemetry["reason"] == "sort-lookback-timeout"
    assert torch.all(torch.isnan(outputs["rgb"])) and torch.all(torch.isnan(outputs["alpha"]))
    assert torch.all(outputs["semantic_id"] == -1)


def test_manual_capacity_check_reports_sort_lookback_failure():
    native = FakeNativeModule([sort_failure(1)])
    service, backend = make_service(native, adaptive_capacity=False)
    service.load_scene(17, **scene_tensors(3))
    view, k = cameras(1)
    service.render(view, k, torch.tensor([17]))
    assert backend.read_counters()["sort_lookback_failures"] == 1
    with pytest.raises(RendererCapacityError, match="sort_lookback_failures=1"):
        backend.check_capacity()


def test_workspace_is_one_arena_of_aligned_views():
    native = FakeNativeModule()
    service, backend = make_service(native, batch=2, max_visible_records=7, max_intersections=33)
    arena = backend._arena
    base = arena.storage.data_ptr()
    for name, spec in arena.plan.buffers.items():
        view = backend.workspace[name]
        offset = view.data_ptr() - base if view.numel() else arena.plan.offsets[name]
        assert offset % 256 == 0 and offset + spec.nbytes <= arena.nbytes, name
    assert backend.workspace_bytes == arena.storage.numel()
    spans = sorted((arena.plan.offsets[n], arena.plan.offsets[n] + s.nbytes) for n, s in arena.plan.buffers.items()
                   if s.nbytes)
    assert all(a_end <= b_start for (_, a_end), (b_start, _) in zip(spans, spans[1:]))  # no overlaps
    assert backend.workspace["counters"].numel() == 6 and int(backend.workspace["counters"].abs().sum()) == 0
