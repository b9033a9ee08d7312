#!/usr/bin/env python3
"""Benchmark: full all-layer peel on one MI355X, as graph replays.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config NAME] [--extras]

A step is one replay of the captured extraction graph (validation, welding,
projection, topology, hierarchies, subdivision, ranks, noding, fans, lift) on
synthetic float64 input already resident in HBM; upload, allocation, capture
and readback are outside the timed region (the reference's protocol,
scripts/benchmark_events.py:77-93). The headline workload is the 65,536-face
(2,3) torus-knot tube; --extras adds the 717,740-face nested shells and the
65,536-layer identical-footprint stack.

Prints one JSON line. ``roofline`` prices the dominant kernel's algorithmic
bytes per launch against its own event-timed duration; ``cpu_baseline`` times
the CPU FP64 restatement of the same route (oracle) on a bounded sample.
"""

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for p in (str(ROOT), str(ROOT / "exact-depth-peeling_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="torus-knot-65536")
    ap.add_argument("--extras", action="store_true", help="also time nested shells and the stack")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument(
        "--copy-calibration",
        type=int,
        metavar="MIB",
        help="only run a plain device copy of MIB mebibytes (scales rocprofv3 FETCH_SIZE/WRITE_SIZE)",
    )
    return ap.parse_args()


def make_peeler(name):
    import torch

    from hip_peel import EventPeeler, scenes
    from hip_peel.io import upload

    host = scenes.CONFIGS[name]()
    arrays = upload(*host, device="cuda:0")
    torch.cuda.synchronize()
    return EventPeeler(*arrays), host


def time_replays(peeler, steps, warmup):
    import torch

    peeler.capture()
    for _ in range(warmup):
        peeler.run()
    peeler.synchronize()
    torch.cuda.synchronize()
    events = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    wall = time.perf_counter()
    for start, end in events:
        start.record(peeler.stream)
        peeler.run()
        end.record(peeler.stream)
    peeler.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall
    per = sorted(s.elapsed_time(e) for s, e in events)
    return {
        "ms_per_step": 1e3 * wall / steps,
        "median_ms": per[len(per) // 2],
        "min_ms": per[0],
        "replays_ms": per,
    }


def algorithmic_bytes(peeler, host, info):
    """Minimum HBM bytes of one extraction: inputs once, output polygons and
    triangles once (the per-unit figures: 24 B per vertex, 12 B per face,
    176 B camera + settings, 16 B per polygon vertex + 8 B per polygon,
    128 B per triangle: uv 48 + world 72 + source 4 + layer 4)."""
    vertices, faces = host[0], host[1]
    P = info["polygons"]
    nverts = int(peeler._output_n[:P].sum().item()) if P else 0
    total = 24 * len(vertices) + 12 * len(faces) + 176 + 128 * info["triangles"] + 16 * nverts + 8 * P
    return total, nverts


def kernel_bytes(name, host, info, nverts):
    """Algorithmic bytes of one launch of a hot kernel (inputs it must read,
    outputs it must write, each once)."""
    F = len(host[1])
    P, T = info["polygons"], info["triangles"]
    if name == "k_mesh_edges":
        # polygon rings + ids + layers in, triangles out
        return 16 * nverts + 12 * P + 128 * T
    if name == "k_subdivide":
        # source rings + planes in (8 vertices of 16 B, 32 B plane, 4 B count), final rings out
        return (16 * 3 + 32 + 4) * F + 16 * nverts + 12 * P
    if name == "k_ray_hits":
        # samples in, hits out (depth 8 + source 4 + group 4)
        return 16 * P + 16 * info["ray_hits"]
    return None


def run_config(name, steps, warmup):
    peeler, host = make_peeler(name)
    timing = time_replays(peeler, steps, warmup)
    info = peeler.read_status()
    stages = peeler.profile()
    peeler.read_status()
    total, nverts = algorithmic_bytes(peeler, host, info)
    return peeler, host, info, stages, timing, total, nverts


def cpu_baseline(name, seconds):
    """CPU FP64 restatement of the device route (oracle), timed on this host."""
    try:
        import oracle
        from hip_peel import scenes
    except Exception as e:  # pragma: no cover
        return {"value": None, "error": str(e)}
    if not hasattr(oracle, "cpu_event_peel"):
        return None
    host = scenes.CONFIGS[name]()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return oracle.cpu_baseline(host, seconds=seconds, threads=threads, workload=name)


def main():
    args = parse()
    if args.gpus != 1:
        print(
            json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design only (DESIGN.md)",
                        "n_gpus": args.gpus})
        )
        sys.exit(2)
    import torch

    assert torch.cuda.is_available(), "bench.py needs a ROCm GPU"
    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": 8 * n, "launches": 3}))
        return
    peeler, host, info, stages, timing, total, nverts = run_config(args.config, args.steps, args.warmup)
    hot = {k: v for k, v in stages.items() if k.startswith("k_")}
    dominant = max(hot, key=hot.get)
    kbytes = kernel_bytes(dominant, host, info, nverts)
    achieved = kbytes / (hot[dominant] * 1e-3) / 1e9 if kbytes else None
    result = {
        "metric": "ms/full peel @65K & 720K faces: FP32-filtered vs plain FP64 HIP, exact ranks",
        "value": round(timing["median_ms"], 4),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(timing["ms_per_step"], 4),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": args.config,
            "faces": len(host[1]),
            "vertices": len(host[0]),
            "route": "events",
            "predicates": "fp64",
            "polygons": info["polygons"],
            "triangles": info["triangles"],
            "layers": info["layers"],
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dominant,
            "kernel_ms": round(hot[dominant], 4),
            "achieved": None if achieved is None else round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": None if achieved is None else round(achieved / HBM_PEAK_GBS, 5),
            "traffic": None,
            "pipeline_algorithmic_bytes": total,
            "pipeline_GBps": round(total / (timing["median_ms"] * 1e-3) / 1e9, 2),
        },
        "stages_ms": {k: round(v, 4) for k, v in stages.items()},
        "counters": {k: info[k] for k in ("splits", "ray_groups", "ray_hits", "unique_contours", "rounds",
                                          "target_visits", "contour_tests", "ray_candidate_tests")},
        "cpu_baseline": None,
    }
    del peeler
    if args.extras:
        extra = {}
        for name in ("nested-shells-717740", "stack-65536"):
            p, h, i, st, t, tot, nv = run_config(name, max(3, args.steps // 2), 1)
            extra[name] = {"median_ms": round(t["median_ms"], 3), "layers": i["layers"],
                           "polygons": i["polygons"], "triangles": i["triangles"],
                           "stages_ms": {k: round(v, 3) for k, v in st.items() if k.startswith("k_")}}
            del p
        result["extra"] = extra
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(args.config, args.cpu_seconds)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
