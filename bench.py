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
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "vectorized-gaussian-renderer-public_amd"))
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBPS = 8000.0
FP32_VALU_PEAK_TFLOPS = 157.3
COMPOSITE_FLOPS_PER_EVALUATION = 20
RECORD_BYTES = 48
STAGES = ("projection+emission", "sort", "tile_ranges", "composite")


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
    parser.add_argument("--copy-calibration-mib", type=int, default=0,
                        help="also run one device-to-device copy of this size (PMC byte-counter calibration)")
    return parser.parse_args()


def main() -> int:
    args = parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: every box here holds one MI355X (asked for --gpus {args.gpus}). "
              "The multi-GPU layout is described in DESIGN.md, not measured.", file=sys.stderr)
        return 2
    import numpy as np
    import torch

    from vgr_amd import CustomHipBackend, RendererService, synthetic

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 1
    device = torch.device("cuda", 0)
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
    pixels = args.cameras * width * height
    value = pixels / 1.0e6 / (ms_per_step / 1000.0)

    # Algorithmic work per launch (one render of the batch).
    pairs = args.cameras * args.gaussians
    visible = counters["visible_gaussians"]
    intersections = counters["tile_intersections"]
    bands = 16
    last_allocation = backend.capacity_stats["last_allocation"] or {}
    passes = int(last_allocation.get("sort_radix_passes_expected", 6))
    global_tiles = args.cameras * math.ceil(width / 16) * math.ceil(height / 16)
    work_bytes = {
        "projection+emission": pairs * 16 + visible * (24 + 12 * bands) + visible * RECORD_BYTES
        + visible * 20 + intersections * 12,
        "sort": intersections * (8 + 24 * passes),
        "tile_ranges": intersections * 8 + global_tiles * 8,
    }
    composite_flops = COMPOSITE_FLOPS_PER_EVALUATION * evaluations
    dominant = max(stage_ms, key=stage_ms.get)
    if dominant == "composite":
        achieved = composite_flops / (stage_ms[dominant] / 1000.0) / 1.0e12
        roofline = {"bound": "valu", "kernel": "composite_kernel<16,256,full>", "achieved": round(achieved, 3),
                    "peak": FP32_VALU_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / FP32_VALU_PEAK_TFLOPS, 4),
                    "traffic": None}
    else:
        achieved = work_bytes[dominant] / (stage_ms[dominant] / 1000.0) / 1.0e9
        roofline = {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None}
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
        threads = min(16, os.cpu_count() or 1)
        cpu_start = time.perf_counter()
        expected = oracle.render([scene.as_kwargs()], [0], rig.viewmats[sample], rig.intrinsics[sample],
                                 [0] * len(sample), height=height, width=width, sh_degree=3, threads=threads)
        cpu_seconds = time.perf_counter() - cpu_start
        cpu_baseline = {"value": round(len(sample) * width * height / 1e6 / cpu_seconds, 4), "unit": "Mpix/s",
                        "cores": threads, "kind": "port",
                        "sample": f"first {len(sample)} of {args.cameras} cameras over the full "
                                  f"{args.gaussians}-Gaussian SH3 scene ({cpu_seconds:.1f} s)"}
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

    line = {
        "metric": "Mpix/s fwd and fwd+bwd, batched-persistent vs plain, 1 MI355X, at tolerance",
        "value": round(value, 2),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
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
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
