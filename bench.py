"""Benchmark of the batched camera path on one MI355X.

Default workload (the metric's GPU configuration): plain GPU forward of 64
cameras 640x480 (fx = fy = 554) on a Fibonacci sphere of radius 10-15 around
1M synthetic Gaussians, SH degree 3, FP32, all four outputs, deterministic
ordering.  One step = one RendererService.render() of the whole batch
(projection + culling, tile emission, radix ordering, tile ranges, fused
compositing), inputs resident in HBM.

Prints one JSON line: Mpix/s = sum of W*H over cameras / 1e6 / mean step time,
the roofline of the dominant kernel measured live with HIP events on the render
stream, and the CPU oracle timed on a bounded sample of the same workload.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import tempfile
import time
from pathlib import Path

# torch is imported before the repository root goes on sys.path: bench processes
# that imported it afterwards (the second and later on one host) measured
# 1.4-2.4 ms of idle GPU time between back-to-back renders, with the same
# per-render GPU spans; importing it first removed the idle time.
import torch

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "vectorized-gaussian-renderer-public_amd"))
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBPS = 8000.0
FP32_VALU_PEAK_TFLOPS = 157.3
COMPOSITE_FLOPS_PER_EVALUATION = 20
RECORD_BYTES = 48
STAGES = ("projection+emission", "sort", "tile_ranges", "composite")


def profiled_traffic(kernel: str) -> int | None:
    """HBM bytes per launch of ``kernel`` (per render for a ``stage:<name>`` entry) from the committed
    PMC summary of this workload."""
    path = ROOT / "profiles" / "c2_plain_forward_hbm_bytes.md"
    try:
        for line in path.read_text().splitlines():
            fields = line.split()
            if len(fields) == 3 and fields[0] == "traffic_bytes" and fields[1] == kernel:
                return int(fields[2])
    except OSError:
        return None
    return None


def parse_args() -> argparse.Namespace:
    parser = argparse.ArgumentParser(description=__doc__)
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=50)
    parser.add_argument("--warmup", type=int, default=5)
    parser.add_argument("--cameras", type=int, default=64)
    parser.add_argument("--gaussians", type=int, default=1_000_000)
    parser.add_argument("--cpu-cameras", type=int, default=16, help="camera sample for the CPU baseline")
    parser.add_argument("--no-cpu-baseline", action="store_true")
    parser.add_argument("--sort-digit-bits", type=int, default=0)
    parser.add_argument("--raster-batch", type=int, default=0)
    parser.add_argument("--cameras-per-block", type=int, default=0)
    parser.add_argument("--sort-tile-items", type=int, default=0)
    parser.add_argument("--workloads", default="c4,c5,c3,dense,ray",
                        help="extra workloads reported beside the main line: c4 (fwd+bwd), c5 (robotics batch, "
                             "batched-persistent vs plain, pose backward), c3 (uneven multi-scene batch), dense "
                             "(tile size 1 with the depth cutoff), ray (exact-ray evaluation); '' for none")
    parser.add_argument("--extra-steps", type=int, default=10, help="timed steps of each extra workload")
    parser.add_argument("--skip-main", action="store_true",
                        help="run only the extra workloads (profiling aid); prints their JSON")
    parser.add_argument("--copy-calibration-mib", type=int, default=0,
                        help="also run one device-to-device copy of this size (PMC byte-counter calibration)")
    return parser.parse_args()


def main() -> int:
    args = parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): cameras would shard across GPUs "
              "with no exchange in the forward pass, and that layout is not measured.", file=sys.stderr)
        return 2
    import numpy as np
    import torch

    from vgr_amd import CustomHipBackend, RendererService, synthetic

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 1
    device = torch.device("cuda", 0)
    if args.skip_main:
        extras = {}
        for name in [w for w in args.workloads.split(",") if w]:
            runner = {"c4": bench_c4_fwd_bwd, "c5": bench_c5_robotics, "c3": bench_c3_uneven,
                      "dense": bench_dense, "ray": bench_ray}[name]
            extras[name] = runner(args, torch, device)
            print(f"[bench] workload {name}: {json.dumps(extras[name])}", file=sys.stderr, flush=True)
        print(json.dumps({"workloads": extras}), flush=True)
        return 0
    width, height, focal, seed = 640, 480, 554.0, 20261202
    scene = synthetic.gaussian_scene(args.gaussians, seed=seed, sh_degree=3)
    rig = synthetic.orbit_rig(args.cameras, width, height, focal, seed=seed, radius=(10.0, 15.0))
    backend = CustomHipBackend(output_srgb=False, deterministic=True, visible_capacity_per_view=1_000_000,
                               intersections_per_visible=2.6, sort_digit_bits=args.sort_digit_bits,
                               raster_batch=args.raster_batch, cameras_per_block=args.cameras_per_block,
                               sort_tile_items=args.sort_tile_items)
    service = RendererService(backend, height=height, width=width, max_views=args.cameras)
    service.initialize(stage=None, device=device)
    service.load_scene(0, **scene.to(device).as_kwargs())
    viewmats = rig.viewmats.to(device)
    intrinsics = rig.intrinsics.to(device)
    scene_ids = torch.zeros(args.cameras, dtype=torch.int64, device=device)
    outputs = service.prepare_outputs(args.cameras)
    print(f"[bench] setup done: {args.cameras} cameras, {args.gaussians} Gaussians", file=sys.stderr, flush=True)
    if args.copy_calibration_mib > 0:
        # A plain copy of known size: its FETCH_SIZE / WRITE_SIZE counters give the
        # scale between those counters and bytes moved.
        source = torch.ones(args.copy_calibration_mib * (1 << 20) // 4, dtype=torch.float32, device=device)
        target = torch.empty_like(source)
        target.copy_(source)
        torch.cuda.synchronize()
        del source, target

    for _ in range(args.warmup):
        service.render(viewmats, intrinsics, scene_ids, outputs=outputs)
    service.synchronize()
    counters = backend.read_counters()
    print(f"[bench] warmup done: counters {counters}", file=sys.stderr, flush=True)

    # Evaluation count for the compositor roofline (one instrumented render, untimed).
    statistics = torch.zeros(4, dtype=torch.int64, device=device)
    backend.statistics = statistics
    service.render(viewmats, intrinsics, scene_ids, outputs=outputs)
    service.synchronize()
    backend.statistics = None
    evaluations = int(statistics[0].item())
    spin_overflows = int(statistics[1].item())

    events = [[torch.cuda.Event(enable_timing=True) for _ in range(5)] for _ in range(args.steps)]
    for step_events in events:
        for event in step_events:
            event.record()  # materialize the HIP event handles
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    wall_start = time.perf_counter()
    start.record()
    for step in range(args.steps):
        backend.stage_events = events[step]
        service.render(viewmats, intrinsics, scene_ids, outputs=outputs)
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall_start
    backend.stage_events = None
    gpu_ms = start.elapsed_time(stop) / args.steps
    stage_ms = {name: 0.0 for name in STAGES}
    for step_events in events:
        for index, name in enumerate(STAGES):
            stage_ms[name] += step_events[index].elapsed_time(step_events[index + 1]) / args.steps
    ms_per_step = max(gpu_ms, wall * 1000.0 / args.steps)
    # Distribution over the timed renders: each render's span from its first stage
    # event to its last (HIP events on the render stream), beside the bracketing
    # event pair and the synchronized wall clock.
    spans = sorted(step_events[0].elapsed_time(step_events[-1]) for step_events in events)
    span_mean = sum(spans) / len(spans)
    timing = {"renders": len(spans), "gpu_ms_mean": round(gpu_ms, 4), "wall_ms_mean": round(wall * 1000.0 / args.steps, 4),
              "span_ms": {"mean": round(span_mean, 4),
                          "std": round(math.sqrt(sum((x - span_mean) ** 2 for x in spans) / len(spans)), 4),
                          "p50": round(spans[len(spans) // 2], 4),
                          "p95": round(spans[min(len(spans) - 1, math.ceil(0.95 * len(spans)) - 1)], 4),
                          "min": round(spans[0], 4), "max": round(spans[-1], 4)}}
    pixels = args.cameras * width * height
    value = pixels / 1.0e6 / (ms_per_step / 1000.0)

    # Algorithmic work per launch (one render of the batch).
    visible = counters["visible_gaussians"]
    intersections = counters["tile_intersections"]
    bands = 16
    last_allocation = backend.capacity_stats["last_allocation"] or {}
    passes = int(last_allocation.get("sort_radix_passes_expected", 6))
    global_tiles = args.cameras * math.ceil(width / 16) * math.ceil(height / 16)
    # Minimum HBM traffic of each stage: the scene is read once (mean, opacity,
    # covariance, SH), every record and key/value written once; repeated reads of a
    # Gaussian by the cameras of a group are cache traffic, not algorithmic bytes.
    work_bytes = {
        "projection+emission": args.gaussians * (16 + 24 + 12 * bands) + visible * RECORD_BYTES
        + intersections * 12,
        "sort": intersections * (8 + 24 * passes),
        "tile_ranges": intersections * 8 + global_tiles * 8,
    }
    composite_flops = COMPOSITE_FLOPS_PER_EVALUATION * evaluations
    dominant = max(stage_ms, key=stage_ms.get)
    default_workload = (args.cameras, args.gaussians) == (64, 1_000_000)
    if dominant == "composite":
        achieved = composite_flops / (stage_ms[dominant] / 1000.0) / 1.0e12
        # VALU-bound: FP32 vector ALU, no dense contraction and hence no MFMA instructions.
        roofline = {"bound": "valu", "kernel": "composite_kernel<16,256,full>", "achieved": round(achieved, 3),
                    "peak": FP32_VALU_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / FP32_VALU_PEAK_TFLOPS, 4),
                    "traffic": profiled_traffic("composite_kernel<16,256,full>") if default_workload else None,
                    "traffic_source": "profiles/c2_plain_forward_hbm_bytes.md (FETCH_SIZE, WRITE_SIZE, copy-scaled)"}
    else:
        achieved = work_bytes[dominant] / (stage_ms[dominant] / 1000.0) / 1.0e9
        roofline = {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
                    "traffic": profiled_traffic(f"stage:{dominant}") if default_workload else None,
                    "traffic_source": "profiles/c2_plain_forward_hbm_bytes.md (FETCH_SIZE, WRITE_SIZE, copy-scaled)"}
    stages = {}
    for name in STAGES:
        entry = {"ms": round(stage_ms[name], 4)}
        if name in work_bytes:
            entry["algorithmic_bytes"] = work_bytes[name]
            entry["achieved_GBps"] = round(work_bytes[name] / (stage_ms[name] / 1000.0) / 1e9, 1)
            entry["hbm_frac"] = round(entry["achieved_GBps"] / HBM_PEAK_GBPS, 4)
        else:
            entry["evaluations"] = evaluations
            entry["achieved_TFLOPs"] = round(composite_flops / (stage_ms[name] / 1000.0) / 1e12, 3)
            entry["valu_frac"] = round(entry["achieved_TFLOPs"] / FP32_VALU_PEAK_TFLOPS, 4)
        stages[name] = entry

    cpu_baseline = None
    parity = None
    if not args.no_cpu_baseline:
        from oracle import oracle

        sample = list(range(min(args.cpu_cameras, args.cameras)))
        threads = cpu_threads()
        cpu_start = time.perf_counter()
        expected = oracle.render([scene.as_kwargs()], [0], rig.viewmats[sample], rig.intrinsics[sample],
                                 [0] * len(sample), height=height, width=width, sh_degree=3, threads=threads)
        cpu_seconds = time.perf_counter() - cpu_start
        one = sample[:2]
        one_start = time.perf_counter()
        oracle.render([scene.as_kwargs()], [0], rig.viewmats[one], rig.intrinsics[one], [0] * len(one),
                      height=height, width=width, sh_degree=3, threads=1)
        one_seconds = time.perf_counter() - one_start
        cpu_baseline = {"value": round(len(sample) * width * height / 1e6 / cpu_seconds, 4), "unit": "Mpix/s",
                        "cores": threads, "kind": "port",
                        "sample": f"first {len(sample)} of {args.cameras} cameras over the full "
                                  f"{args.gaussians}-Gaussian SH3 scene ({cpu_seconds:.1f} s)",
                        "nproc": os.cpu_count(), "cpu_model": cpu_model(),
                        "threads_note": "16 threads: the host-CPU share for one GPU",
                        "one_core": {"value": round(len(one) * width * height / 1e6 / one_seconds, 4),
                                     "sample": f"first {len(one)} cameras, 1 thread ({one_seconds:.1f} s)"},
                        "config1_full": cpu_config1(threads)}
        got = {name: outputs[name][sample].cpu().numpy() for name in outputs}
        finite = np.isfinite(expected["depth"])
        depth_rel = np.abs(got["depth"][finite] - expected["depth"][finite]) / np.abs(expected["depth"][finite])
        parity = {
            "cameras_checked": len(sample),
            "rgb_max_abs": float(np.abs(got["rgb"] - expected["rgb"]).max()),
            "alpha_max_abs": float(np.abs(got["alpha"] - expected["alpha"]).max()),
            "depth_max_rel": float(depth_rel.max()) if depth_rel.size else 0.0,
            "depth_background_equal": bool(np.array_equal(np.isinf(got["depth"]), np.isinf(expected["depth"]))),
            "semantic_mismatches": int((got["semantic_id"] != expected["semantic_id"]).sum()),
        }
        parity["pass"] = bool(parity["rgb_max_abs"] <= 1e-4 and parity["alpha_max_abs"] <= 1e-4
                              and parity["depth_max_rel"] <= 1e-4 and parity["depth_background_equal"]
                              and parity["semantic_mismatches"] == 0 and spin_overflows == 0)

    del outputs, backend, service
    torch.cuda.empty_cache()
    extras = {}
    for name in [w for w in args.workloads.split(",") if w]:
        runner = {"c4": bench_c4_fwd_bwd, "c5": bench_c5_robotics, "c3": bench_c3_uneven,
                      "dense": bench_dense, "ray": bench_ray}[name]
        print(f"[bench] workload {name}", file=sys.stderr, flush=True)
        try:
            extras[name] = runner(args, torch, device)
        except (RuntimeError, ValueError) as error:  # reported, never hides the main line
            extras[name] = {"error": f"{type(error).__name__}: {error}"[:500]}
        print(f"[bench] workload {name}: {json.dumps(extras[name])}", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()

    line = {
        "metric": "Mpix/s fwd and fwd+bwd, batched-persistent vs plain, 1 MI355X, at tolerance",
        "value": round(value, 2),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "timing": timing,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded torch.Generator scene and Fibonacci-sphere cameras)",
        "config": {"workload": f"plain GPU forward, {args.cameras} cameras {width}x{height}, "
                               f"{args.gaussians} Gaussians SH3 FP32, all four outputs, deterministic order",
                   "mode": "plain", "tile_size": 16, "counters": counters},
        "roofline": roofline,
        "stages": stages,
        "cpu_baseline": cpu_baseline,
        "parity": parity,
        "workloads": extras,
    }
    print(json.dumps(line), flush=True)
    return 0


def oracle_parity(scene, rig, cameras: list[int], width: int, height: int, sh_degree: int, gpu_image,
                  keep: dict | None = None, **oracle_options) -> dict:
    """Worst deviation of the GPU images of ``cameras`` from the CPU oracle; ``gpu_image(i)`` gives
    camera i's outputs as numpy [H,W,C] arrays.  Tolerances: RGB/alpha max-abs 1e-4, depth relative
    1e-4, semantic IDs exact.  ``oracle_options`` (tile_size, ray_evaluation, depth buckets) select
    the oracle's matching mode; ``keep``, when given, receives the oracle's images."""
    import numpy as np
    from oracle import oracle

    expected = oracle.render([scene.as_kwargs()], [0], rig.viewmats[cameras], rig.intrinsics[cameras],
                             [0] * len(cameras), height=height, width=width, sh_degree=sh_degree,
                             threads=cpu_threads(), **oracle_options)
    if keep is not None:
        keep.update(expected)
    report = {"cameras_checked": cameras, "rgb_max_abs": 0.0, "alpha_max_abs": 0.0, "depth_max_rel": 0.0,
              "depth_background_equal": True, "semantic_mismatches": 0}
    for local, camera in enumerate(cameras):
        got = gpu_image(camera)
        want = {name: expected[name][local] for name in ("rgb", "depth", "alpha", "semantic_id")}
        report["rgb_max_abs"] = max(report["rgb_max_abs"], float(np.abs(got["rgb"] - want["rgb"]).max()))
        report["alpha_max_abs"] = max(report["alpha_max_abs"], float(np.abs(got["alpha"] - want["alpha"]).max()))
        finite = np.isfinite(want["depth"])
        report["depth_background_equal"] &= bool(np.array_equal(np.isinf(got["depth"]), np.isinf(want["depth"])))
        if finite.any():
            rel = np.abs(got["depth"][finite] - want["depth"][finite]) / np.abs(want["depth"][finite])
            report["depth_max_rel"] = max(report["depth_max_rel"], float(rel.max()))
        report["semantic_mismatches"] += int((got["semantic_id"] != want["semantic_id"]).sum())
    report["pass"] = bool(report["rgb_max_abs"] <= 1e-4 and report["alpha_max_abs"] <= 1e-4
                          and report["depth_max_rel"] <= 1e-4 and report["depth_background_equal"]
                          and report["semantic_mismatches"] == 0)
    return report


def autotune_cache() -> Path:
    """Per-run autotuner cache (the measurement itself is what the bench reports)."""
    return Path(tempfile.gettempdir()) / f"vgr_autotune_bench_{os.getpid()}.json"


def cpu_threads() -> int:
    """Host threads for the CPU baseline: the host-CPU share for one GPU (16), or fewer if the host has fewer."""
    return max(1, min(16, os.cpu_count() or 1))


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_config1(threads: int) -> dict:
    """1 camera 640x480 (fx = fy = 554) at radius 12 around 100k SH0 Gaussians, all outputs,
    rendered in full by the CPU oracle with ``threads`` threads and with one."""
    from oracle import oracle
    from vgr_amd import synthetic

    scene, rig = synthetic.single_camera_workload()
    result = {"workload": "1 camera 640x480, 100k Gaussians SH0, all four outputs (in full)"}
    for label, count in (("threads", threads), ("one_core", 1)):
        start = time.perf_counter()
        oracle.render([scene.as_kwargs()], [0], rig.viewmats, rig.intrinsics, [0], height=480, width=640,
                      sh_degree=0, threads=count)
        seconds = time.perf_counter() - start
        result[label] = {"threads": count, "seconds": round(seconds, 3),
                         "mpix_s": round(640 * 480 / 1e6 / seconds, 4)}
    return result


def timed(torch, fn, steps: int, warmup: int) -> tuple[float, float]:
    """(mean GPU ms per call from events on the current stream, wall ms per call)."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    wall = time.perf_counter()
    start.record()
    for _ in range(steps):
        fn()
    stop.record()
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - wall) * 1000.0 / steps
    return start.elapsed_time(stop) / steps, wall_ms


def bench_c4_fwd_bwd(args, torch, device) -> dict:
    """16 cameras 512x512 around 500k SH3 Gaussians; L1 + depth-L1 loss against targets
    rendered from a perturbed copy; one step = forward + loss + backward of every parameter."""
    from vgr_amd import synthetic
    from vgr_amd.autograd import DifferentiableRenderer, compose_pose

    seed, size, focal, cameras = 20261204, 512, 443.4, 16
    scene = synthetic.gaussian_scene(500_000, seed=seed, sh_degree=3)
    rig = synthetic.orbit_rig(cameras, size, size, focal, seed=seed, radius=(10.0, 15.0))
    g = torch.Generator().manual_seed(seed + 1)
    logits = torch.logit(scene.opacities.double()).float()
    params = [scene.means, torch.log(scene.scales), scene.rotations, logits, scene.features]
    params = [p.to(device).contiguous() for p in params]
    target_means = (scene.means + 0.05 * torch.randn(scene.means.shape, generator=g)).to(device)
    target_logits = (logits + 0.3 * torch.randn(logits.shape, generator=g)).to(device)
    viewmats = rig.viewmats.to(device)
    intrinsics = rig.intrinsics.to(device)
    renderer = DifferentiableRenderer(height=size, width=size, device=device, capacity_check="deferred")
    with torch.no_grad():
        target = renderer.render(target_means, params[1], params[2], target_logits, params[4], viewmats, intrinsics)
        target = {k: v.clone() for k, v in target.items()}
    mask = (target["alpha"] > 0.5) & torch.isfinite(target["depth"])
    # Target-side terms of the loss are fixed: prepared once, like a data loader would.
    depth_target = torch.where(mask, target["depth"], 0.0)
    depth_count = mask.sum().clamp(min=1).float()
    leaves = [p.clone().requires_grad_(True) for p in params]
    # Camera pose as a 6-DoF twist on the base views: the step also differentiates the pose.
    xi = torch.zeros((cameras, 6), device=device, requires_grad=True)

    def loss_of(out):
        depth = torch.where(mask, out["depth"], 0.0)
        depth_term = (depth - depth_target).abs().sum() / depth_count
        return (out["rgb"] - target["rgb"]).abs().mean() + depth_term

    def forward_only():
        with torch.no_grad():
            renderer.render(*leaves, viewmats, intrinsics)

    def step():
        for leaf in leaves:
            leaf.grad = None
        xi.grad = None
        views = compose_pose(xi, viewmats)
        loss_of(renderer.render(*leaves, views, intrinsics)).backward()

    fwd_ms, fwd_wall = timed(torch, forward_only, args.extra_steps, 2)
    step_ms, step_wall = timed(torch, step, args.extra_steps, 2)
    pixels = cameras * size * size
    fwd = max(fwd_ms, fwd_wall)
    both = max(step_ms, step_wall)
    finite = all(bool(torch.isfinite(leaf.grad).all()) for leaf in leaves) and bool(torch.isfinite(xi.grad).all())
    # Forward images of the training path against the CPU oracle on two cameras.
    with torch.no_grad():
        images = renderer.render(*params, viewmats, intrinsics, semantic_ids=scene.semantic_ids.to(device))
        # The oracle takes the activated parameters exactly as the renderer computed them.
        used = synthetic.SyntheticScene(params[0].cpu(), torch.exp(params[1]).cpu(), params[2].cpu(),
                                        torch.sigmoid(params[3]).cpu(), params[4].cpu(), scene.semantic_ids)
    parity = oracle_parity(used, rig, [0, cameras - 1], size, size, 3,
                           lambda i: {k: v[i].cpu().numpy() for k, v in images.items()})
    return {"workload": f"{cameras} cameras {size}x{size}, 500k Gaussians SH3 FP32, L1 + depth-L1 loss, grads of "
                        "means/log-scales/quats/opacity logits/SH and the 6-DoF camera pose (se(3) twist on the base views)",
            "fwd_ms": round(fwd, 3), "fwd_bwd_ms": round(both, 3), "fwd_bwd_over_fwd": round(both / fwd, 3),
            "gpu_and_wall_ms": {"fwd": [round(fwd_ms, 3), round(fwd_wall, 3)],
                                "fwd_bwd": [round(step_ms, 3), round(step_wall, 3)]},
            "fwd_mpix_s": round(pixels / 1e6 / (fwd / 1e3), 1),
            "fwd_bwd_mpix_s": round(pixels / 1e6 / (both / 1e3), 1),
            "counters": renderer.last_counters, "grads_finite": finite, "oracle_parity": parity,
            "note": "capacity_check='deferred': counters are verified one call later without a host "
                    "synchronization; an overflowed render would give NaN gradients and raise at the next call"}


def bench_c5_robotics(args, torch, device) -> dict:
    """1024 cameras 160x120 along 32 smooth trajectories in a 5M-Gaussian SH1 room:
    batched-persistent (graph replay) vs the plain single submission, plus a pose-only backward."""
    from vgr_amd import synthetic
    from vgr_amd.autograd import DifferentiableRenderer
    from vgr_amd.batched import BatchedRenderer, CameraSpec

    seed, width, height, focal = 20261205, 160, 120, 139.0
    scene = synthetic.gaussian_scene(5_000_000, seed=seed, sh_degree=1, semantic_classes=256, layout="room")
    rig = synthetic.trajectory_rig(32, 32, width, height, focal, seed=seed)
    viewmats = rig.viewmats.to(device)
    intrinsics = rig.intrinsics.to(device)
    gpu_scene = scene.to(device)
    renderer = BatchedRenderer(device=device)
    renderer.register_scene(0, **gpu_scene.as_kwargs())
    specs = [CameraSpec(0, width, height)] * rig.batch
    renderer.plan(specs)
    counters = renderer.calibrate(viewmats, intrinsics)
    renderer.capture()
    batched_ms, batched_wall = timed(torch, renderer.render, args.extra_steps, 2)
    plain_ms, plain_wall = timed(torch, renderer.render_plain, args.extra_steps, 1)
    renderer.render()
    torch.cuda.synchronize()
    group = renderer.plain_groups[0]
    flat = renderer.outputs
    equal = all(torch.equal(flat[name].view(rig.batch, height, width, -1), group["outputs"][name])
                for name in ("rgb", "depth", "alpha", "semantic_id"))
    statistics = torch.zeros(4, dtype=torch.int64, device=device)
    renderer.graph = None
    renderer._submit(statistics=statistics)
    torch.cuda.synchronize()
    kept = int(statistics[2])
    parity = oracle_parity(scene, rig, [0, rig.batch - 1], width, height, 1,
                           lambda i: {k: v.cpu().numpy() for k, v in renderer.camera_outputs(i).items()})
    pixels = rig.batch * width * height
    batched = max(batched_ms, batched_wall)
    plain = max(plain_ms, plain_wall)
    result = {"workload": "1024 cameras 160x120 (fx=fy=139) on 32 Catmull-Rom trajectories in a 5M-Gaussian SH1 "
                          "room, rgb+depth+alpha+semantic",
              "batched_graph_ms": round(batched, 3), "plain_ms": round(plain, 3),
              "batched_mpix_s": round(pixels / 1e6 / (batched / 1e3), 1),
              "plain_mpix_s": round(pixels / 1e6 / (plain / 1e3), 1),
              "batched_over_plain": round(plain / batched, 3), "outputs_bitwise_equal": equal,
              "batched_ordering": renderer.ordering, "fully_sorted_fallback_tiles": int(statistics[3]),
              "oracle_parity": parity, "counters": counters, "kept_cull_items": kept}
    del renderer
    torch.cuda.empty_cache()
    # Pose-only backward on the plain training path.
    trainer = DifferentiableRenderer(height=height, width=width, device=device, cameras_per_block=16)
    logits = torch.logit(gpu_scene.opacities.double()).float()
    frozen = [gpu_scene.means, torch.log(gpu_scene.scales), gpu_scene.rotations, logits, gpu_scene.features]
    xi = torch.zeros((rig.batch, 6), device=device, requires_grad=True)
    from vgr_amd.autograd import compose_pose

    def pose_step():
        xi.grad = None
        views = compose_pose(xi, viewmats)
        out = trainer.render(*frozen, views, intrinsics, semantic_ids=gpu_scene.semantic_ids)
        (out["rgb"].mean() + out["alpha"].mean()).backward()

    def forward_only():
        with torch.no_grad():
            trainer.render(*frozen, viewmats, intrinsics, semantic_ids=gpu_scene.semantic_ids)

    fwd_ms, fwd_wall = timed(torch, forward_only, args.extra_steps, 1)
    pose_ms, pose_wall = timed(torch, pose_step, args.extra_steps, 1)
    plain_pose = max(pose_ms, pose_wall)
    result.update({"pose_fwd_ms": round(max(fwd_ms, fwd_wall), 3),
                   "pose_fwd_bwd_ms": round(plain_pose, 3),
                   "pose_fwd_bwd_over_fwd": round(plain_pose / max(fwd_ms, fwd_wall), 3),
                   "pose_grads_finite": bool(torch.isfinite(xi.grad).all())})
    plain_grad = xi.grad.detach().clone()
    del trainer
    torch.cuda.empty_cache()
    # Pose-only backward in batched-persistent mode: checkpointed tile-front forward
    # (graph replay), backward fused into the compositing walk; the scene is given
    # with the trainer's activations so both paths render the same images.
    posed = BatchedRenderer(device=device, pose_gradients=True)
    posed.register_scene(0, means=frozen[0], scales=torch.exp(frozen[1]), rotations=frozen[2],
                         opacities=torch.sigmoid(frozen[3]), features=frozen[4],
                         semantic_ids=gpu_scene.semantic_ids)
    posed.plan(specs)
    posed.calibrate(viewmats, intrinsics)
    posed.capture()
    xi_batched = torch.zeros((rig.batch, 6), device=device, requires_grad=True)

    def batched_pose_step():
        xi_batched.grad = None
        rgb, _, alpha = posed.render_pose(compose_pose(xi_batched, viewmats))
        (rgb.mean() + alpha.mean()).backward()

    posed_fwd_ms, posed_fwd_wall = timed(torch, posed.render, args.extra_steps, 1)
    bpose_ms, bpose_wall = timed(torch, batched_pose_step, args.extra_steps, 1)
    batched_pose = max(bpose_ms, bpose_wall)
    scale = float(plain_grad.abs().max())
    result.update({"batched_pose_fwd_ms": round(max(posed_fwd_ms, posed_fwd_wall), 3),
                   "batched_pose_fwd_bwd_ms": round(batched_pose, 3),
                   "batched_pose_fwd_bwd_over_fwd": round(batched_pose / max(posed_fwd_ms, posed_fwd_wall), 3),
                   "batched_over_plain_pose_step": round(plain_pose / batched_pose, 3),
                   "batched_pose_ordering": posed.ordering,
                   "pose_grad_max_error_over_max": round(float((xi_batched.grad - plain_grad).abs().max()) / scale, 6)
                   if scale > 0 else None})
    del posed
    torch.cuda.empty_cache()
    return result


def service_renderer(torch, device, scene, batch: int, width: int, height: int, **knobs):
    """RendererService + CustomHipBackend over one scene with the given backend knobs."""
    from vgr_amd import CustomHipBackend, RendererService

    backend = CustomHipBackend(output_srgb=False, **knobs)
    service = RendererService(backend, height=height, width=width, max_views=batch)
    service.initialize(stage=None, device=device)
    service.load_scene(0, **scene.as_kwargs())
    return backend, service, service.prepare_outputs(batch)


def bench_dense(args, torch, device) -> dict:
    """Dense per-pixel path (tile size 1) with the depth cutoff, without it, and 16x16 tiles, on
    8 cameras 128x128 (fx = fy = 111) along 2 trajectories in a 5M-Gaussian SH1 room."""
    from vgr_amd import synthetic

    seed, size, focal = 20261205, 128, 111.0
    scene = synthetic.gaussian_scene(5_000_000, seed=seed, sh_degree=1, semantic_classes=256, layout="room")
    rig = synthetic.trajectory_rig(2, 4, size, size, focal, seed=seed)
    gpu_scene = scene.to(device)
    viewmats, intrinsics = rig.viewmats.to(device), rig.intrinsics.to(device)
    ids = torch.zeros(rig.batch, dtype=torch.int64, device=device)
    result = {"workload": f"{rig.batch} cameras {size}x{size} (fx=fy={focal:g}) in a 5M-Gaussian SH1 room, "
                          "rgb+depth+alpha+semantic"}
    images = {}
    for name, knobs in (("tile1_cutoff", {"tile_size": 1}), ("tile1_no_cutoff", {"tile_size": 1, "deterministic": True}),
                        ("tile16", {"tile_size": 16})):
        backend, service, outputs = service_renderer(torch, device, gpu_scene, rig.batch, size, size, **knobs)

        def render():
            service.render(viewmats, intrinsics, ids, outputs=outputs)

        ms, wall = timed(torch, render, args.extra_steps, 2)
        both = max(ms, wall)
        images[name] = {k: v.clone() for k, v in outputs.items()}
        result[name] = {"ms": round(both, 3), "mpix_s": round(rig.batch * size * size / 1e6 / (both / 1e3), 1),
                        "intersections": backend.read_counters()["tile_intersections"]}
        del service, backend, outputs
        torch.cuda.empty_cache()
    result["images_bitwise_equal"] = all(torch.equal(images["tile1_cutoff"][k], images[n][k])
                                         for n in ("tile1_no_cutoff", "tile16") for k in images[n])
    result["oracle_parity"] = oracle_parity(
        scene, rig, [0], size, size, 1, lambda i: {k: v[i].cpu().numpy() for k, v in images["tile1_cutoff"].items()},
        tile_size=1, depth_bucket_count=4096, depth_bucket_group_size=64)
    return result


def bench_ray(args, torch, device) -> dict:
    """Exact-ray evaluation on the main workload (64 cameras 640x480, 1M SH3): 16x16 raster tiles
    over 32x32 bins, and 8x8 tiles."""
    from vgr_amd import synthetic

    width, height, focal, seed = 640, 480, 554.0, 20261202
    scene = synthetic.gaussian_scene(1_000_000, seed=seed, sh_degree=3)
    rig = synthetic.orbit_rig(64, width, height, focal, seed=seed, radius=(10.0, 15.0))
    gpu_scene = scene.to(device)
    viewmats, intrinsics = rig.viewmats.to(device), rig.intrinsics.to(device)
    ids = torch.zeros(rig.batch, dtype=torch.int64, device=device)
    result = {"workload": "64 cameras 640x480, 1M Gaussians SH3, exact-ray evaluation, all four outputs"}
    images = {}
    for name, tile in (("ray16", 16), ("ray8", 8)):
        backend, service, outputs = service_renderer(torch, device, gpu_scene, rig.batch, width, height,
                                                     ray_gaussian_evaluation=True, tile_size=tile,
                                                     deterministic=True, visible_capacity_per_view=1_000_000,
                                                     intersections_per_visible=2.6)

        def render():
            service.render(viewmats, intrinsics, ids, outputs=outputs)

        ms, wall = timed(torch, render, args.extra_steps, 2)
        both = max(ms, wall)
        images[name] = {k: v.clone() for k, v in outputs.items()}
        result[name] = {"ms": round(both, 3), "mpix_s": round(rig.batch * width * height / 1e6 / (both / 1e3), 1),
                        "intersections": backend.read_counters()["tile_intersections"]}
        del service, backend, outputs
        torch.cuda.empty_cache()
    # Bins come from the support ellipse (double, integer radius) while the support test at a pixel
    # runs in float, so a pixel within rounding of a support boundary can take an entry under one
    # bin size and not the other, as in the reference.  Both bin sizes are checked against the
    # oracle on the camera where they differ most (camera 0 when they agree), and the oracle's own
    # images at the two bin sizes are compared on that camera.
    differing = (images["ray16"]["rgb"] != images["ray8"]["rgb"]).flatten(1).sum(1)
    camera = int(differing.argmax()) if int(differing.max()) > 0 else 0
    result["ray16_vs_ray8"] = {
        "rgb_max_abs": float((images["ray16"]["rgb"] - images["ray8"]["rgb"]).abs().max()),
        "alpha_max_abs": float((images["ray16"]["alpha"] - images["ray8"]["alpha"]).abs().max()),
        "cameras_differing": int((differing > 0).sum()), "rgb_values_differing": int(differing.sum())}
    expected = {}
    for name, tile in (("ray16", 16), ("ray8", 8)):
        expected[name] = {}
        result["oracle_parity_" + name] = oracle_parity(
            scene, rig, [camera], width, height, 3,
            lambda i, name=name: {k: v[i].cpu().numpy() for k, v in images[name].items()},
            keep=expected[name], tile_size=tile, ray_evaluation=True)
    result["oracle_ray16_vs_ray8_rgb_values_differing"] = int(
        (expected["ray16"]["rgb"] != expected["ray8"]["rgb"]).sum())
    return result


def bench_c3_uneven(args, torch, device) -> dict:
    """256 cameras with resolutions from {128^2, 320x240, 640x480, 1280x720} over 4 scenes
    (200k/500k/1M/2M SH3), round-robin: batched-persistent graph replay vs plain per-group path."""
    import numpy as np

    from oracle import oracle
    from vgr_amd import synthetic
    from vgr_amd.autotune import Autotuner, tune_batched
    from vgr_amd.batched import BatchedRenderer, CameraSpec

    seed = 20261203
    sizes = (200_000, 500_000, 1_000_000, 2_000_000)
    resolutions = ((128, 128), (320, 240), (640, 480), (1280, 720))
    g = torch.Generator().manual_seed(seed)
    choice = torch.randint(0, 4, (256,), generator=g).tolist()
    renderer = BatchedRenderer(device=device)
    for k, count in enumerate(sizes):
        renderer.register_scene(k, **synthetic.gaussian_scene(count, seed=seed + k, sh_degree=3).to(device).as_kwargs())
    specs, views, ks = [], [], []
    rig = synthetic.orbit_rig(256, 1, 1, 1.0, seed=seed, radius=(10.0, 15.0))
    for i in range(256):
        w, h = resolutions[choice[i]]
        specs.append(CameraSpec(i % 4, w, h))
        views.append(rig.viewmats[i])
        ks.append(synthetic.pinhole(1, w, h, (w / 2) / math.tan(math.radians(30)))[0])
    viewmats = torch.stack(views).to(device)
    intrinsics = torch.stack(ks).to(device)
    renderer.plan(specs)
    counters = renderer.calibrate(viewmats, intrinsics)
    renderer.capture()
    default_ms, default_wall = timed(torch, renderer.render, max(3, args.extra_steps // 2), 1)
    default_config = dict(renderer.tuning, tile_size=renderer.options["tile_size"], ordering=renderer.ordering)
    # Autotuned launch parameters (tile size, LDS batch, digit width, camera group, ordering),
    # measured on this frame; the tuned plan is then timed like the default one.
    tuned = tune_batched(renderer, viewmats, intrinsics, tuner=Autotuner(repeats=1, cache_path=autotune_cache()),
                         force=True)
    counters = renderer.read_counters()
    batched_ms, batched_wall = timed(torch, renderer.render, max(3, args.extra_steps // 2), 1)
    plain_ms, plain_wall = timed(torch, renderer.render_plain, max(3, args.extra_steps // 2), 1)
    renderer.render()
    torch.cuda.synchronize()
    equal = True
    for group in renderer.plain_groups:
        for local, camera in enumerate(group["index_list"]):
            view = renderer.camera_outputs(camera)
            for name in ("rgb", "depth", "alpha", "semantic_id"):
                equal = equal and torch.equal(view[name], group["outputs"][name][local])
    pixels = sum(s.width * s.height for s in specs)
    batched = max(batched_ms, batched_wall)
    default = max(default_ms, default_wall)
    plain = max(plain_ms, plain_wall)
    # Oracle check on one camera of every resolution (the scenes cycle with the cameras).
    subset = sorted({(specs[i].width, specs[i].height): i for i in range(len(specs))}.values())
    expected = oracle.render([synthetic.gaussian_scene(count, seed=seed + k, sh_degree=3).as_kwargs()
                              for k, count in enumerate(sizes)], list(range(4)), viewmats[subset].cpu(),
                             intrinsics[subset].cpu(), [specs[i].scene_id for i in subset], height=1, width=1,
                             resolutions=[(specs[i].width, specs[i].height) for i in subset], sh_degree=3,
                             threads=cpu_threads())
    parity = {"cameras_checked": subset, "rgb_max_abs": 0.0, "alpha_max_abs": 0.0, "depth_max_rel": 0.0,
              "semantic_mismatches": 0}
    offset = 0
    for i in subset:
        w, h = specs[i].width, specs[i].height
        got = {k: v.reshape(h * w, -1).cpu().numpy() for k, v in renderer.camera_outputs(i).items()}
        want = {k: expected[k][offset:offset + w * h] for k in got}
        offset += w * h
        parity["rgb_max_abs"] = max(parity["rgb_max_abs"], float(np.abs(got["rgb"] - want["rgb"]).max()))
        parity["alpha_max_abs"] = max(parity["alpha_max_abs"], float(np.abs(got["alpha"] - want["alpha"]).max()))
        finite = np.isfinite(want["depth"])
        if finite.any():
            rel = np.abs(got["depth"][finite] - want["depth"][finite]) / np.abs(want["depth"][finite])
            parity["depth_max_rel"] = max(parity["depth_max_rel"], float(rel.max()))
        parity["semantic_mismatches"] += int((got["semantic_id"] != want["semantic_id"]).sum())
    parity["pass"] = bool(parity["rgb_max_abs"] <= 1e-4 and parity["alpha_max_abs"] <= 1e-4
                          and parity["depth_max_rel"] <= 1e-4 and parity["semantic_mismatches"] == 0)
    return {"workload": "256 cameras, resolutions U{128^2, 320x240, 640x480, 1280x720}, 4 scenes 200k/500k/1M/2M "
                        "SH3 round-robin",
            "batched_graph_ms": round(batched, 3), "plain_ms": round(plain, 3),
            "batched_mpix_s": round(pixels / 1e6 / (batched / 1e3), 1),
            "plain_mpix_s": round(pixels / 1e6 / (plain / 1e3), 1),
            "batched_over_plain": round(plain / batched, 3),
            "default_config": default_config, "default_graph_ms": round(default, 3),
            "tuned_config": tuned, "tuned_over_default": round(default / batched, 3),
            "outputs_bitwise_equal": equal, "oracle_parity": parity, "counters": counters,
            "plain_groups": len(renderer.plain_groups)}


if __name__ == "__main__":
    sys.exit(main())
