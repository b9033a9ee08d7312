#!/usr/bin/env python3
"""Benchmark: full all-layer peel on one MI355X, as graph replays.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config NAME]
                    [--no-large] [--no-extras] [--batch-graph] [--no-cpu]

A step is one replay of the captured extraction graph (validation, welding,
projection, topology, hierarchies, subdivision, ranks, noding, fans, lift) on
synthetic float64 input already resident in HBM; upload, allocation, capture
and readback are outside the timed region (the reference's protocol,
scripts/benchmark_events.py:77-93). The headline workload is the 65,536-face
(2,3) torus-knot tube, timed with binary32-filtered ray predicates (`value`)
and with plain binary64 predicates; the 717,740-face nested shells are timed
both ways as well. Extras (default on): the 65,536-layer identical-footprint
stack, the 256 uneven icospheres peeled as one merged workspace, and the
layered route (DevicePeeler) on the headline workload; --batch-graph adds the
icospheres as one graph over 256 workspaces.

Before timing, each workload's launch tuning is autotuned (hip_peel.tune;
outputs do not depend on it; --no-tune keeps the defaults), as a user tunes a
workload once before replaying it.

Prints one JSON line. ``roofline`` prices the dominant kernel's algorithmic
bytes per launch against its own event-timed duration; ``cpu_baseline`` times
the restatement of the reference's fast CPU method (fast-bvh-peel
--resolved-mesh) on the same workload with the reference's protocol, and the
CPU restatement of the device's own route as a second figure.
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
# the CPU baseline's OpenMP threads yield when idle, as in the reference's
# protocol; set before any OpenMP runtime starts
os.environ.setdefault("OMP_WAIT_POLICY", "PASSIVE")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="torus-knot-65536")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the 65,536-layer stack, the 256 merged icospheres and the layered route")
    ap.add_argument("--batch-graph", action="store_true",
                    help="also time the 256 icospheres as one graph over 256 workspaces")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-large", action="store_true", help="skip the 717,740-face nested shells")
    ap.add_argument("--no-tune", action="store_true", help="time the default launch tuning only")
    ap.add_argument("--cpu-seconds", type=float, default=8.0)
    ap.add_argument(
        "--copy-calibration",
        type=int,
        metavar="MIB",
        help="only run a plain device copy of MIB mebibytes (scales rocprofv3 FETCH_SIZE/WRITE_SIZE)",
    )
    return ap.parse_args()


def make_peeler(name, filter_fp32=False, host=None):
    import torch

    from hip_peel import EventPeeler, scenes
    from hip_peel.io import upload

    host = host if host is not None else scenes.CONFIGS[name]()
    arrays = upload(*host, device="cuda:0")
    torch.cuda.synchronize()
    return EventPeeler(*arrays, filter_fp32=filter_fp32), host


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


def run_batch(steps, warmup, count=256, seed=0):
    """Batched uneven instances: ``count`` noisy icospheres (320 to 81,920
    faces each), each under its own pinhole, peeled by one graph launch
    (BatchPeeler); timed beside the members' own graphs launched back to back
    on their own streams. Pools come from the learned pool sizes; members that
    overflow them are rebuilt with the conservative capacities."""
    import torch

    from hip_peel import BatchPeeler, EventPeeler, scenes
    from hip_peel.io import upload

    meshes = scenes.icosphere_instances(count, seed=seed)
    wall = time.perf_counter()
    batch = BatchPeeler.from_meshes(meshes, filter_fp32=True)
    setup_s = time.perf_counter() - wall
    root = batch.stream  # the first member's stream: the batch graph forks and joins there

    def timed(launch):
        for _ in range(warmup):
            launch()
        torch.cuda.synchronize()
        per = []
        for _ in range(steps):
            start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            start.record(root)
            launch()
            end.record(root)
            end.synchronize()
            per.append(start.elapsed_time(end))
        torch.cuda.synchronize()
        return sorted(per)[len(per) // 2]

    batch_ms = timed(batch.run)
    infos = batch.read_status()
    for p in batch.peelers:
        p.capture()

    def separate():
        fork = torch.cuda.Event()
        fork.record(root)
        for p in batch.peelers:
            p.stream.wait_event(fork)
            p.run()
        for p in batch.peelers:
            root.wait_stream(p.stream)

    separate_ms = timed(separate)
    # pool memory of the learned sizes against the conservative capacities, per distinct size
    conservative = {}
    for m in meshes:
        n = len(m[1])
        if n not in conservative:
            probe = EventPeeler(*upload(*m[:4]))
            conservative[n] = probe.device_bytes
            del probe
    faces = sum(len(m[1]) for m in meshes)
    return {
        "instances": count,
        "faces": faces,
        "batch_graph_ms": round(batch_ms, 3),
        "separate_graphs_ms": round(separate_ms, 3),
        "batch_speedup": round(separate_ms / batch_ms, 3),
        "faces_per_s": round(faces / (batch_ms * 1e-3)),
        "layers_max": max(i["layers"] for i in infos),
        "polygons": sum(i["polygons"] for i in infos),
        "triangles": sum(i["triangles"] for i in infos),
        "rebuilt_members": len(batch.rebuilt),
        "pool_gib_learned": round(sum(p.device_bytes for p in batch.peelers) / 2**30, 3),
        "pool_gib_conservative": round(sum(conservative[len(m[1])] for m in meshes) / 2**30, 3),
        "setup_s": round(setup_s, 2),
    }


def run_merged(steps, warmup, count=256, seed=0):
    """The same 256 icospheres peeled as one workspace over an atlas of their
    windows (MergedBatch): every kernel runs once over all 6.1 M faces."""
    import torch

    from hip_peel import MergedBatch, scenes

    meshes = scenes.icosphere_instances(count, seed=seed)
    wall = time.perf_counter()
    merged = MergedBatch(meshes, filter_fp32=True)
    setup_s = time.perf_counter() - wall
    timing = time_replays(merged.peeler, steps, warmup)
    info = merged.read_status()
    faces = sum(len(m[1]) for m in meshes)
    out = {
        "instances": count,
        "faces": faces,
        "merged_ms": round(timing["median_ms"], 3),
        "faces_per_s": round(faces / (timing["median_ms"] * 1e-3)),
        "layers_max": info["layers"],
        "polygons": info["polygons"],
        "triangles": info["triangles"],
        "rebuilt": merged.rebuilt,
        "pool_gib": round(merged.peeler.device_bytes / 2**30, 3),
        "setup_s": round(setup_s, 2),
    }
    del merged
    torch.cuda.synchronize()
    return out


def filter_classes(info):
    """Binary32 decisions and binary64 evaluations per filtered predicate class."""
    from hip_peel.events import FILTER_CLASSES

    out = {}
    for c in FILTER_CLASSES:
        f, b = info[f"filtered_{c}"], info[f"fp64_{c}"]
        out[c] = {"filtered": f, "fp64": b, "fallback_fraction": round(b / (f + b), 6) if f + b else None}
    return out


def run_layered(name, steps):
    """The layered route (DevicePeeler: whole-face ranks, then one clipping
    pass per layer) timed as graph replays. Its per-layer cuts put more
    distinct vertices on single polygon edges than the event route does; where
    a workload exceeds the 128-vertex edge-noding bound (bit 128, the
    reference's own maximum: the torus knot's long thin tube faces are cut
    along the same edges many times) the overflow is reported and an
    81,920-face noisy icosphere under a pinhole (a member of the batched
    workload) is timed instead. The reference publishes 9.767 s for its 64K
    dragon on an L40 (results/gpu/README.md:15)."""
    import torch

    from hip_peel import DevicePeeler, scenes
    from hip_peel.io import upload

    out = {"workload": name, "reference_64k_dragon_l40_ms": 9767.0}
    candidates = [(name, scenes.CONFIGS[name]())]
    candidates.append(("icosphere-81920", scenes.icosphere_instances(1, seed=0, levels=(6, 6))[0]))
    for label, host in candidates:
        p = DevicePeeler(*upload(*host, device="cuda:0"))
        wall = time.perf_counter()
        p.run()
        status, _ = p._read()
        first_s = time.perf_counter() - wall
        if status[9]:
            out[f"{label}_error_bits"] = int(status[9])
            del p
            continue
        info = p.read_status()
        timing = time_replays(p, steps if first_s < 5 else 1, 0)
        out.update({"workload": label, "faces": len(host[1]), "median_ms": round(timing["median_ms"], 3),
                    "layers": info["layers"], "polygons": info["polygons"], "triangles": info["triangles"],
                    "passes": info["peeling_passes"], "peak_residual": info["peak_residual_patches"],
                    "kernel_launches": p.graph_nodes().get("kernel")})
        del p
        break
    torch.cuda.synchronize()
    return out


def learned_pools(name):
    """Pool memory of the learned sizes against the conservative capacities
    for one workload, with the overflow-rerun outcome and equal outputs."""
    import torch

    from hip_peel import EventPeeler, pools, scenes
    from hip_peel.io import upload

    host = scenes.CONFIGS[name]()
    arrays = upload(*host, device="cuda:0")
    base = EventPeeler(*arrays)
    base.run()
    want = base.read_status()
    peeler, rebuilt = pools.admitted_event_peeler(*arrays)
    got = peeler.read_status()
    assert all(got[k] == want[k] for k in ("polygons", "triangles", "layers"))
    used = base.device_bytes if rebuilt else peeler.device_bytes
    out = {"pool_gib_learned": round(used / 2**30, 3), "pool_gib_conservative": round(base.device_bytes / 2**30, 3),
           "rebuilt": bool(rebuilt)}
    del base, peeler
    torch.cuda.synchronize()
    return out


def run_config(name, steps, warmup, filter_fp32=False, host=None, tune=False):
    """Time graph replays of one workload; with ``tune`` the launch tuning is
    autotuned first (outputs do not depend on it) and the report kept on
    ``peeler.tune_report``."""
    peeler, host = make_peeler(name, filter_fp32, host)
    peeler.tune_report = None
    if tune:
        from hip_peel import tune as tuner

        report = tuner.autotune(peeler, replays=5, rounds=2)
        peeler.tune_report = {"best": report["best"], "default_ms": round(report["default_ms"], 4),
                              "tuned_ms": round(report["best_ms"], 4), "trials": len(report["trials"])}
    timing = time_replays(peeler, steps, warmup)
    info = peeler.read_status()
    stages = peeler.profile()
    peeler.read_status()
    total, nverts = algorithmic_bytes(peeler, host, info)
    return peeler, host, info, stages, timing, total, nverts


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (scaled by a known-size copy), or None."""
    path = ROOT / "profiles" / f"traffic_{workload.replace('-', '_')}.md"
    if not path.exists():
        return None
    text = path.read_text()
    block = text.split("```json", 1)[-1].split("```", 1)[0] if "```json" in text else ""
    try:
        table = json.loads(block)["traffic_bytes"]
    except (ValueError, KeyError):
        return None
    return table.get(f"peel::{kernel}")


def cpu_fast_baseline(name, threads_all):
    """The reference's fast CPU method (fast-bvh-peel --resolved-mesh),
    restated in oracle/cpu_fast.cpp, timed by the reference's protocol
    (scripts/benchmark_event_cpu.py:41-62): four extractions per thread count,
    the first discarded, the median of the other three, OMP_WAIT_POLICY=PASSIVE
    (set before the OpenMP runtime starts, at the top of this file). Scope
    "extraction" is the reference's algorithm_seconds (classification, layers,
    meshing; fast_peel.cpp:653-660); scope "full" adds projection and
    topology."""
    import statistics

    import oracle
    from hip_peel import scenes

    host = scenes.CONFIGS[name]()
    per = {}
    result = None
    for threads in sorted({1, 4, threads_all}):
        result = oracle.fast_peel(*host, threads=threads, repeat=4, keep=False)
        kept = result.extract_s[1:]
        med = statistics.median(kept)
        per[threads] = {"extraction_ms": round(1e3 * med, 3), "full_ms": round(1e3 * (med + result.setup_s), 3),
                        "extraction_runs_ms": [round(1e3 * x, 3) for x in kept]}
    cpu_max = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            cpu_max = f.read().strip()
    except OSError:
        pass
    return {
        "value": per[threads_all]["extraction_ms"],
        "unit": "ms",
        "cores": threads_all,
        "kind": "port",
        "method": "fast-bvh-peel --resolved-mesh (src/fast_peel.cpp, peel_common.h), restated in oracle/cpu_fast.cpp",
        "sample": f"4 extractions of {name} ({len(host[1])} faces) per thread count, first discarded, median of 3; "
                  f"scope: extraction and meshing after projection/topology; host nproc={os.cpu_count()}, "
                  f"cgroup cpu.max={cpu_max}, OMP_WAIT_POLICY={os.environ.get('OMP_WAIT_POLICY')}",
        "by_threads": {str(k): v for k, v in per.items()},
        "layers": result.counts["layers"],
        "patches": result.counts["patches"],
        "triangles": result.counts["faces"],
    }


def cpu_port_baseline(name, seconds):
    """Secondary figure: the CPU FP64 restatement of the device's event route
    (oracle/cpu_event.cpp), the same algorithm as the GPU path."""
    import oracle
    from hip_peel import scenes

    host = scenes.CONFIGS[name]()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    out = oracle.cpu_baseline(host, seconds=seconds, threads=threads, workload=name)
    return {k: out[k] for k in ("value", "unit", "cores", "sample", "fewer_threads")}


def progress(message):
    """One line per finished stage on stderr (the JSON result is stdout's only line)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {message}", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(
            json.dumps({"error": "this build runs on one GPU; sharding a scene or a batch over several GPUs is not built",
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
    # plain binary64 predicates first, then the binary32-filtered run (`value`)
    tune = not args.no_tune
    p64, host, info64, stages64, timing64, _, _ = run_config(args.config, args.steps, args.warmup, tune=tune)
    tune64 = p64.tune_report
    del p64
    progress(f"{args.config} binary64: {timing64['median_ms']:.4f} ms")
    peeler, host, info, stages, timing, total, nverts = run_config(args.config, args.steps, args.warmup,
                                                                   filter_fp32=True, host=host, tune=tune)
    progress(f"{args.config} binary32-filtered: {timing['median_ms']:.4f} ms")
    # the patch and triangle sets are deterministic (ray counters depend on
    # which patch represents a component, so they vary between runs)
    for key in ("polygons", "triangles", "layers"):
        assert info[key] == info64[key], (key, info[key], info64[key])
    decided = info["filtered_predicates"] + info["fp64_fallbacks"]
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
            "predicates": "fp32-filtered containment, straddle, front_segment and depth order; fp64 fallback",
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
            "traffic": measured_traffic(args.config, dominant),
            "pipeline_algorithmic_bytes": total,
            "pipeline_GBps": round(total / (timing["median_ms"] * 1e-3) / 1e9, 2),
        },
        "fp32_vs_fp64": {
            args.config: {
                "fp32_filtered_ms": round(timing["median_ms"], 4),
                "fp64_ms": round(timing64["median_ms"], 4),
                "tuning_fp32": peeler.tune_report,
                "tuning_fp64": tune64,
                "speedup": round(timing64["median_ms"] / timing["median_ms"], 3),
                "filtered_predicates": info["filtered_predicates"],
                "fp64_fallbacks": info["fp64_fallbacks"],
                "fallback_fraction": round(info["fp64_fallbacks"] / decided, 6) if decided else None,
                "by_class": filter_classes(info),
                "fp32_stages_ms": {k: round(v, 4) for k, v in stages.items() if k.startswith("k_")},
                "fp64_stages_ms": {k: round(v, 4) for k, v in stages64.items() if k.startswith("k_")},
            }
        },
        "stages_ms": {k: round(v, 4) for k, v in stages.items()},
        "counters": {k: info[k] for k in ("splits", "ray_groups", "ray_hits", "unique_contours", "rounds",
                                          "target_visits", "contour_tests", "ray_candidate_tests")},
        "cpu_baseline": None,
    }
    # one extraction's graph: kernel launches and the node types it holds
    result["graph_nodes"] = peeler.graph_nodes()
    result["launches_per_extraction"] = result["graph_nodes"].get("kernel", 0)
    del peeler
    if not args.no_large:
        name = "nested-shells-717740"
        large = {}
        for flag in (False, True):
            p, h, i, st, t, tot, nv = run_config(name, max(3, args.steps // 2), 1, filter_fp32=flag, tune=tune)
            large[flag] = (i, t, st, p.tune_report)
            del p
            progress(f"{name} filter_fp32={flag}: {t['median_ms']:.3f} ms")
        (i64, t64, st64, tr64), (i32, t32, st32, tr32) = large[False], large[True]
        learned = learned_pools(name)
        progress(f"{name} learned pools: {learned}")
        d32 = i32["filtered_predicates"] + i32["fp64_fallbacks"]
        result["fp32_vs_fp64"][name] = {
            "fp32_filtered_ms": round(t32["median_ms"], 3),
            "fp64_ms": round(t64["median_ms"], 3),
            "speedup": round(t64["median_ms"] / t32["median_ms"], 3),
            "layers": i32["layers"],
            "polygons": i32["polygons"],
            "triangles": i32["triangles"],
            "fallback_fraction": round(i32["fp64_fallbacks"] / d32, 6) if d32 else None,
            "by_class": filter_classes(i32),
            "stages_ms": {k: round(v, 3) for k, v in st32.items() if k.startswith("k_")},
            "fp64_stages_ms": {k: round(v, 3) for k, v in st64.items() if k.startswith("k_")},
            "learned_pools": learned,
            "tuning_fp32": tr32,
            "tuning_fp64": tr64,
        }
    if not args.no_extras:
        extra = {}
        for name in ("stack-65536",):
            p, h, i, st, t, tot, nv = run_config(name, max(3, args.steps // 2), 1)
            extra[name] = {"median_ms": round(t["median_ms"], 3), "layers": i["layers"],
                           "polygons": i["polygons"], "triangles": i["triangles"],
                           "stages_ms": {k: round(v, 3) for k, v in st.items() if k.startswith("k_")}}
            del p
            progress(f"{name}: {extra[name]['median_ms']} ms")
        extra["icospheres-256-merged"] = run_merged(max(3, args.steps // 2), 1)
        progress(f"icospheres merged: {extra['icospheres-256-merged']['merged_ms']} ms")
        if args.batch_graph:
            extra["icospheres-256"] = run_batch(max(3, args.steps // 2), 1)
        extra["layered-" + args.config] = run_layered(args.config, 3)
        progress(f"layered route: {extra['layered-' + args.config]}")
        result["extra"] = extra
    if not args.no_cpu:
        threads_all = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        result["cpu_baseline"] = cpu_fast_baseline(args.config, threads_all)
        progress(f"cpu fast baseline: {result['cpu_baseline']['value']} ms on {threads_all} threads")
        result["cpu_baseline"]["event_route_port"] = cpu_port_baseline(args.config, args.cpu_seconds)
        result["gpu_vs_cpu_fast"] = round(result["cpu_baseline"]["value"] / result["value"], 1)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
