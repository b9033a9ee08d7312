"""Launch autotuning of a workspace.

The extraction's outputs do not depend on how its kernels are launched, only
its time does. ``autotune`` times graph replays of one peeler over candidate
values of its ``PeelTuning`` fields, one field at a time (coordinate descent,
``rounds`` sweeps), checks that every candidate reproduces the default's
polygon, triangle and layer counts, and leaves the peeler on the fastest
setting. ``TuningCache`` keeps the winner per workload signature (route,
switches, face-count octave) in a JSON file, so a workload is tuned once.

Fields (``include/peel355.h`` ``PeelTuning``):

* ``persistent_blocks_per_cu``: blocks of the subdivision work queue per CU;
* ``ray_lanes_target``: lanes in flight allowed for rays of a whole wave each
  (at most 2,048 rays, 64 x rays within it); other rays take 16 lanes each;
* ``heavy_lanes``: lane budget of the subtree jobs of heavy rays and beams;
* ``long_edge_cells``: grid cells beyond which a polygon edge is noded by a
  whole wave instead of one lane;
* ``leaf_bucket_log2``: the ray walks stop descending 2^n (or 2^(n+1)) leaves
  above the bottom and test those leaves' boxes together (the hierarchy's
  leaf size as the walks see it).
"""

import json
import math
import os
from pathlib import Path

CANDIDATES = {
    "persistent_blocks_per_cu": (1, 2, 3, 4),
    "ray_lanes_target": (1024, 131072, 4194304),
    "heavy_lanes": (16384, 32768, 65536, 131072, 262144),
    "long_edge_cells": (16, 32, 48, 96, 192),
    "leaf_bucket_log2": (1, 2, 3, 4),
}
TUNABLE = tuple(CANDIDATES)


def replay_ms(peeler, replays=5, warmup=1):
    """Median ms of ``replays`` graph replays on the peeler's stream."""
    import torch

    for _ in range(warmup):
        peeler.run()
    peeler.synchronize()
    times = []
    for _ in range(replays):
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(peeler.stream)
        peeler.run()
        end.record(peeler.stream)
        end.synchronize()
        times.append(start.elapsed_time(end))
    return sorted(times)[len(times) // 2]


def _outputs(peeler):
    info = peeler.read_status()
    return info["polygons"], info["triangles"], info["layers"]


def signature(peeler):
    """Workload key: route, switches and the octave of the face count."""
    faces = max(1, int(peeler.n))
    return f"route{peeler._route}-flags{peeler.flags}-faces2^{math.ceil(math.log2(faces))}"


def autotune(peeler, replays=5, rounds=2, candidates=None):
    """Tune ``peeler`` in place; returns {"best", "best_ms", "default_ms",
    "trials"}. Every candidate must reproduce the default's output counts."""
    candidates = dict(candidates or CANDIDATES)
    limit = peeler.tuning["max_blocks_per_cu"]
    candidates["persistent_blocks_per_cu"] = tuple(
        v for v in candidates.get("persistent_blocks_per_cu", (1,)) if v <= limit) or (1,)
    best = {k: v for k, v in peeler.tuning.items() if k in TUNABLE}
    peeler.set_tuning(**best)
    best_ms = default_ms = replay_ms(peeler, replays)
    want = _outputs(peeler)
    trials = [(dict(best), default_ms)]
    for _ in range(rounds):
        changed = False
        for field, values in candidates.items():
            for value in values:
                if value == best[field]:
                    continue
                trial = dict(best, **{field: value})
                peeler.set_tuning(**trial)
                ms = replay_ms(peeler, replays)
                if _outputs(peeler) != want:
                    raise AssertionError(f"tuning {trial} changed the extraction")
                trials.append((trial, ms))
                if ms < best_ms:
                    best, best_ms, changed = trial, ms, True
        if not changed:
            break
    peeler.set_tuning(**best)
    return {"best": best, "best_ms": best_ms, "default_ms": default_ms, "trials": trials}


class TuningCache:
    """Best tuning per workload signature, in a JSON file (default
    ``$PEEL355_TUNING_CACHE`` or ``~/.cache/peel355/tuning.json``)."""

    def __init__(self, path=None):
        default = Path(os.environ.get("PEEL355_TUNING_CACHE", Path.home() / ".cache" / "peel355" / "tuning.json"))
        self.path = Path(path) if path is not None else default
        self.entries = json.loads(self.path.read_text()) if self.path.exists() else {}

    def get(self, key):
        return self.entries.get(key)

    def put(self, key, tuning):
        self.entries[key] = {k: int(v) for k, v in tuning.items() if k in TUNABLE}
        self.path.parent.mkdir(parents=True, exist_ok=True)
        self.path.write_text(json.dumps(self.entries, indent=2, sort_keys=True) + "\n")


def tuned(peeler, cache=None, **options):
    """Apply the cached tuning of the peeler's workload, or autotune it and
    cache the winner. Returns (tuning, report or None)."""
    cache = cache if cache is not None else TuningCache()
    key = signature(peeler)
    entry = cache.get(key)
    if entry is not None:
        limit = peeler.tuning["max_blocks_per_cu"]
        entry = dict(entry, persistent_blocks_per_cu=min(entry["persistent_blocks_per_cu"], limit))
        peeler.set_tuning(**entry)
        return entry, None
    report = autotune(peeler, **options)
    cache.put(key, report["best"])
    return report["best"], report
