"""Launch tuning (peel_set_tuning, hip_peel.tune): argument checks, outputs
bit-identical under every setting, and the autotuner with its cache."""

import pytest

from test_gpu_parity import canonical_patches, canonical_triangles, gpu_outputs

pytestmark = pytest.mark.gpu


def knot_peeler(segments=64, sides=32, **options):
    from hip_peel import EventPeeler, scenes
    from hip_peel.io import upload

    return EventPeeler(*upload(*scenes.torus_knot(segments=segments, sides=sides)), **options)


def test_tuning_arguments():
    p = knot_peeler()
    t = p.tuning
    assert t["persistent_blocks_per_cu"] == 1 and t["max_blocks_per_cu"] >= 1
    assert t["ray_lanes_target"] == 4194304 and t["heavy_lanes"] == 65536 and t["long_edge_cells"] == 48
    assert t["leaf_bucket_log2"] == 2
    for bad in ({"persistent_blocks_per_cu": 0}, {"persistent_blocks_per_cu": t["max_blocks_per_cu"] + 1},
                {"ray_lanes_target": 8}, {"heavy_lanes": 10}, {"long_edge_cells": 0}, {"leaf_bucket_log2": 5},
                {"leaf_bucket_log2": -1}):
        with pytest.raises(ValueError):
            p.set_tuning(**bad)
    with pytest.raises(ValueError):
        p.set_tuning(max_blocks_per_cu=3)
    assert p.tuning == t  # rejected settings change nothing


@pytest.mark.parametrize("filter_fp32", [False, True])
def test_outputs_do_not_depend_on_tuning(filter_fp32):
    print(f"[tune] outputs under extreme settings fp32={filter_fp32}", flush=True)
    p = knot_peeler(filter_fp32=filter_fp32)
    p.run()
    _, r0, s0, l0, uv0, w0, ts0, tl0 = gpu_outputs(p)
    want = (canonical_patches(r0, s0, l0), canonical_triangles(uv0, w0, ts0, tl0))
    limit = p.tuning["max_blocks_per_cu"]
    for setting in ({"persistent_blocks_per_cu": limit, "ray_lanes_target": 64, "heavy_lanes": 1024,
                     "long_edge_cells": 1, "leaf_bucket_log2": 0},
                    {"persistent_blocks_per_cu": 1, "ray_lanes_target": 1 << 30, "heavy_lanes": 1 << 24,
                     "long_edge_cells": 1 << 20, "leaf_bucket_log2": 4}):
        p.set_tuning(**setting)
        p.run()
        _, r, s, l, uv, w, ts, tl = gpu_outputs(p)
        assert (canonical_patches(r, s, l), canonical_triangles(uv, w, ts, tl)) == want, setting


@pytest.mark.parametrize("filter_fp32", [False, True])
def test_ray_walks_agree(filter_fp32):
    """Rays in groups of 64 lanes and of 16 (breadth first, k_ray_bfs) give the
    same patches and triangles."""
    print(f"[tune] ray walk modes fp32={filter_fp32}", flush=True)
    p = knot_peeler(filter_fp32=filter_fp32)
    p.run()
    rays = p.read_status()["ray_groups"]
    assert 8 < rays <= 2048
    results = []
    for target in (1 << 30, 64):  # groups of 64, of 16
        p.set_tuning(ray_lanes_target=target)
        p.run()
        info, r, s, l, uv, w, ts, tl = gpu_outputs(p)
        results.append((canonical_patches(r, s, l), canonical_triangles(uv, w, ts, tl)))
    assert results[0] == results[1]


def test_autotune_and_cache(tmp_path):
    from hip_peel import tune

    print("[tune] autotune", flush=True)
    p = knot_peeler(128, 64)
    report = tune.autotune(p, replays=3, rounds=1,
                           candidates={"persistent_blocks_per_cu": (1, 2), "ray_lanes_target": (131072, 524288),
                                       "heavy_lanes": (32768, 65536), "long_edge_cells": (32, 48),
                                       "leaf_bucket_log2": (2, 3)})
    assert report["best_ms"] <= report["default_ms"]
    assert {k: p.tuning[k] for k in tune.TUNABLE} == report["best"]
    cache = tune.TuningCache(tmp_path / "t.json")
    q = knot_peeler(128, 64)
    best, rep = tune.tuned(q, cache=cache, replays=3, rounds=1,
                           candidates={"long_edge_cells": (32, 48)})
    assert rep is not None and cache.get(tune.signature(q)) == best
    r = knot_peeler(128, 64)
    again, rep2 = tune.tuned(r, cache=tune.TuningCache(tmp_path / "t.json"))
    assert rep2 is None and again == best
    assert {k: r.tuning[k] for k in tune.TUNABLE} == best
