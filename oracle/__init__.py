"""ORACLE — test infrastructure only.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package, and only as the checker. The product
(``exact-depth-peeling_amd/``) never imports, links or calls it.

Contents:

* ``liboracle.so`` (``make -C oracle``): exact per-point ranks with Shewchuk
  expansion arithmetic (``exact_rank.cpp``), the region audit
  (``audit.cpp``) and the CPU FP64 restatement of the device route
  (``cpu_*.cpp``), which is also the timed CPU baseline.
* ``load_scene``: an independent reader of the reference's scene JSON
  (warp_peel/io.py:11-48 semantics: rational strings become binary64, faces are
  an unwelded soup, settings = [xmin, ymin, xmax, ymax, perspective, near]).
"""

import ctypes
import json
import subprocess
from fractions import Fraction
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
_lib = None

FLAG_EDGE, FLAG_TIE, FLAG_WINDOW = 1, 2, 4


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = ctypes.CDLL(str(LIB_PATH))
        _lib.oracle_exact_ranks.restype = ctypes.c_long
        _lib.oracle_audit_fragments.restype = None
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def number(x):
    return float(Fraction(x)) if isinstance(x, str) else float(x)


def load_scene(path_or_data):
    """Scene JSON -> (vertices (3N,3) f64, faces (N,3) i32, camera (4,4) f64,
    settings (6,) f64, labels (N,) i64)."""
    data = path_or_data
    if not isinstance(data, dict):
        data = json.loads(Path(path_or_data).read_text())
    tris = data["triangles"]
    vertices = np.array(
        [[[number(x) for x in p] for p in t["vertices"]] for t in tris], dtype=np.float64
    ).reshape((-1, 3))
    faces = np.arange(3 * len(tris), dtype=np.int32).reshape((-1, 3))
    camera = np.eye(4)
    cam = data.get("camera", {})
    if "world_to_camera" in cam:
        camera[:3] = [[number(x) for x in row] for row in cam["world_to_camera"]]
    settings = np.array(
        [number(x) for x in data["window"]]
        + [float(cam.get("type") == "perspective"), number(cam.get("near", 0.01))]
    )
    labels = np.array([t.get("id", i) for i, t in enumerate(tris)], dtype=np.int64)
    return vertices, faces, camera, settings, labels


class Ranks:
    """CSR exact ranks: for point i, entries offsets[i]:offsets[i+1] hold the
    covering sources front to back (rank = position)."""

    def __init__(self, offsets, sources, ranks, flags, point_flags):
        self.offsets, self.sources, self.ranks = offsets, sources, ranks
        self.flags, self.point_flags = flags, point_flags

    def at(self, i):
        a, b = self.offsets[i], self.offsets[i + 1]
        return list(self.sources[a:b])


def exact_ranks(vertices, faces, camera, settings, points):
    vertices = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 3)
    faces = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1, 3)
    camera = np.ascontiguousarray(camera, dtype=np.float64).reshape(4, 4)
    settings = np.ascontiguousarray(settings, dtype=np.float64).reshape(6)
    points = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 2)
    nq = len(points)
    cap = max(16, 8 * nq)
    while True:
        offsets = np.zeros(nq + 1, dtype=np.int64)
        sources = np.zeros(cap, dtype=np.int32)
        ranks = np.zeros(cap, dtype=np.int32)
        flags = np.zeros(cap, dtype=np.int32)
        pflags = np.zeros(max(1, nq), dtype=np.int32)
        total = lib().oracle_exact_ranks(
            _p(vertices), len(vertices), _p(faces), len(faces), _p(camera), _p(settings),
            _p(points), nq, _p(offsets), _p(sources), _p(ranks), _p(flags), _p(pflags),
            ctypes.c_long(cap),
        )
        if total < 0:
            raise ValueError("oracle rejected the scene")
        if total <= cap:
            return Ranks(offsets, sources[:total], ranks[:total], flags[:total], pflags[:nq])
        cap = int(total)


AUDIT_FIELDS = (
    "points_checked",
    "points_skipped",
    "pairs_checked",
    "missing",
    "duplicate",
    "wrong_layer",
    "spurious",
    "pairs_on_fragment_boundary",
)


def audit(polygons, frag_sources, frag_layers, points, ranks, max_examples=16):
    """polygons: list of (n,2) arrays (normalised, counter-clockwise)."""
    sizes = np.array([len(p) for p in polygons], dtype=np.int64)
    offsets = np.zeros(len(polygons) + 1, dtype=np.int64)
    np.cumsum(sizes, out=offsets[1:])
    xy = (
        np.ascontiguousarray(np.concatenate([np.asarray(p, dtype=np.float64).reshape(-1, 2) for p in polygons]))
        if len(polygons)
        else np.zeros((1, 2))
    )
    frag_sources = np.ascontiguousarray(frag_sources, dtype=np.int32)
    frag_layers = np.ascontiguousarray(frag_layers, dtype=np.int32)
    points = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 2)
    counters = np.zeros(8, dtype=np.int64)
    examples = np.zeros((max_examples, 5), dtype=np.int32)
    srcs = np.ascontiguousarray(ranks.sources if len(ranks.sources) else np.zeros(1, np.int32), dtype=np.int32)
    rks = np.ascontiguousarray(ranks.ranks if len(ranks.ranks) else np.zeros(1, np.int32), dtype=np.int32)
    pf = np.ascontiguousarray(ranks.point_flags if len(ranks.point_flags) else np.zeros(1, np.int32), dtype=np.int32)
    lib().oracle_audit_fragments(
        _p(offsets), _p(xy), _p(frag_sources), _p(frag_layers), len(polygons), _p(points),
        len(points), _p(ranks.offsets), _p(srcs), _p(rks), _p(pf), _p(counters), _p(examples),
        max_examples,
    )
    result = dict(zip(AUDIT_FIELDS, map(int, counters)))
    bad = result["missing"] + result["duplicate"] + result["wrong_layer"] + result["spurious"]
    result["mismatches"] = bad
    result["examples"] = examples[: min(bad, max_examples)].tolist()
    return result


CPU_COUNT_FIELDS = (
    "polygons",
    "triangles",
    "layers",
    "splits",
    "ray_groups",
    "ray_hits",
    "error",
    "contour_tests",
    "unique_contours",
    "boundary_vertices",
    "rank_components",
    "rank_connections",
    "beam_attempts",
    "certified_beams",
)


class CpuResult:
    """Output of the CPU FP64 restatement of the event route."""

    def __init__(self, rings, sources, layers, tri_uv, tri_world, tri_src, tri_layer, counts, seconds):
        self.rings, self.sources, self.layers = rings, sources, layers
        self.tri_uv, self.tri_world, self.tri_src, self.tri_layer = tri_uv, tri_world, tri_src, tri_layer
        self.counts, self.seconds = counts, seconds


def cpu_event_peel(vertices, faces, camera, settings, layers=0, threads=0, cap_poly=None, cap_tri=None,
                   propagate_ranks=True, share_regions=True):
    vertices = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 3)
    faces = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1, 3)
    camera = np.ascontiguousarray(camera, dtype=np.float64).reshape(4, 4)
    settings = np.ascontiguousarray(settings, dtype=np.float64).reshape(6)
    n = max(16, len(faces))
    cap_poly = cap_poly or 8 * n
    cap_tri = cap_tri or 32 * n
    poly_xy = np.zeros((cap_poly, 32, 2))
    poly_n = np.zeros(cap_poly, np.int32)
    poly_src = np.zeros(cap_poly, np.int32)
    poly_layer = np.zeros(cap_poly, np.int32)
    tri_uv = np.zeros((cap_tri, 3, 2))
    tri_world = np.zeros((cap_tri, 3, 3))
    tri_src = np.zeros(cap_tri, np.int32)
    tri_layer = np.zeros(cap_tri, np.int32)
    counts = np.zeros(len(CPU_COUNT_FIELDS), np.int64)
    seconds = np.zeros(4)
    lib().oracle_cpu_event_options(int(bool(propagate_ranks)), int(bool(share_regions)))
    f = lib().oracle_cpu_event_peel
    f.restype = ctypes.c_int
    f(
        _p(vertices), len(vertices), _p(faces), len(faces), _p(camera), _p(settings), int(layers),
        int(threads), _p(poly_xy), _p(poly_n), _p(poly_src), _p(poly_layer), int(cap_poly), _p(tri_uv),
        _p(tri_world), _p(tri_src), _p(tri_layer), int(cap_tri), _p(counts), _p(seconds),
    )
    c = dict(zip(CPU_COUNT_FIELDS, map(int, counts)))
    P, T = min(c["polygons"], cap_poly), min(c["triangles"], cap_tri)
    rings = [poly_xy[i, : poly_n[i]].copy() for i in range(P)]
    return CpuResult(
        rings, poly_src[:P].copy(), poly_layer[:P].copy(), tri_uv[:T].copy(), tri_world[:T].copy(),
        tri_src[:T].copy(), tri_layer[:T].copy(), c, seconds.copy(),
    )


def cpu_baseline(host, seconds=20.0, threads=1, workload=""):
    """Time full CPU peels of ``host`` for about ``seconds`` (at least one run);
    report ms per full peel (median)."""
    import os
    import time

    times, runs, start = [], 0, time.perf_counter()
    result = None
    while runs == 0 or (time.perf_counter() - start < seconds and runs < 50):
        t = time.perf_counter()
        result = cpu_event_peel(*host[:4], threads=threads)
        times.append(time.perf_counter() - t)
        runs += 1
    times.sort()
    # fewer threads, a few runs each (the reference's CPU protocol times 1, 4
    # and all cores, scripts/benchmark_event_cpu.py:41-62)
    variants = {}
    for t in sorted({1, 4} - {int(threads)}):
        if t > threads:
            continue
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            cpu_event_peel(*host[:4], threads=t)
            ts.append(time.perf_counter() - t0)
        variants[f"{t}_threads_ms"] = round(1e3 * sorted(ts)[1], 3)
    cpu_max = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            cpu_max = f.read().strip()
    except OSError:
        pass
    return {
        "value": round(1e3 * times[len(times) // 2], 3),
        "unit": "ms",
        "cores": int(threads),
        "kind": "port",
        "sample": f"{runs} full peels of {workload} ({len(host[1])} faces), median; "
        f"host nproc={os.cpu_count()}, cgroup cpu.max={cpu_max}",
        "fewer_threads": variants,
        "stage_seconds": [round(x, 4) for x in result.seconds],
        "polygons": result.counts["polygons"],
        "layers": result.counts["layers"],
    }


LAYERED_COUNT_FIELDS = ("polygons", "layers", "error", "depth_ties", "passes", "peak_residual", "max_hidden",
                        "whole_sources", "residual_sources", "candidate_tests")


class LayeredResult:
    def __init__(self, rings, sources, layers, counts, stats):
        self.rings, self.sources, self.layers = rings, sources, layers
        self.counts, self.stats = counts, stats


def convexify(ring):
    """The layered route's strict-convexity rule applied to one ring (n, 2)."""
    xy = np.ascontiguousarray(ring, dtype=np.float64).reshape(-1, 2).copy()
    n = lib().oracle_convexify(_p(xy), len(xy))
    return xy[:n]


def cpu_layered_peel(vertices, faces, camera, settings, layers=0, layer_capacity=128, pieces=None, residual=None,
                     cap_poly=None, order_seed=0, convexify_rings=True):
    """CPU restatement of the layered (DevicePeeler) route in the device's
    candidate order; ``order_seed`` != 0 shuffles each pass's residual order;
    ``convexify_rings=False`` drops the strict-convexity rule (to show what it
    prevents)."""
    lib().oracle_cpu_layered_options(int(bool(convexify_rings)))
    vertices = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 3)
    faces = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1, 3)
    camera = np.ascontiguousarray(camera, dtype=np.float64).reshape(4, 4)
    settings = np.ascontiguousarray(settings, dtype=np.float64).reshape(6)
    n = max(16, len(faces))
    pieces = pieces or (512 if n <= 8192 else 2048)
    residual = residual or 4 * n
    cap_poly = cap_poly or 16 * n
    poly_xy = np.zeros((cap_poly, 32, 2))
    poly_n = np.zeros(cap_poly, np.int32)
    poly_src = np.zeros(cap_poly, np.int32)
    poly_layer = np.zeros(cap_poly, np.int32)
    counts = np.zeros(len(LAYERED_COUNT_FIELDS), np.int64)
    stats = np.zeros((layer_capacity, 3), np.int64)
    f = lib().oracle_cpu_layered_peel
    f.restype = ctypes.c_int
    f(_p(vertices), len(vertices), _p(faces), len(faces), _p(camera), _p(settings), int(layers),
      int(layer_capacity), int(pieces), int(residual), ctypes.c_uint(order_seed), _p(poly_xy), _p(poly_n),
      _p(poly_src), _p(poly_layer), int(cap_poly), _p(counts), _p(stats), int(layer_capacity))
    c = dict(zip(LAYERED_COUNT_FIELDS, map(int, counts)))
    P = min(c["polygons"], cap_poly)
    rings = [poly_xy[i, : poly_n[i]].copy() for i in range(P)]
    return LayeredResult(rings, poly_src[:P].copy(), poly_layer[:P].copy(), c, stats[: c["passes"]].copy())


FAST_COUNT_FIELDS = ("layers", "patches", "faces", "error", "constant_rank_sources", "residual_sources",
                     "candidate_tests", "polygon_differences", "convexity_repairs", "collapsed_polygons",
                     "depth_ties", "active_sources", "contour_segments", "boundary_edges", "silhouette_edges",
                     "nonmanifold_or_inconsistent_edges", "same_facing_internal_edges_ignored",
                     "camera_clipping_boundary_edges", "point_query_candidates")


class FastResult:
    def __init__(self, rings, sources, layers, tri_uv, tri_src, tri_layer, counts, setup_s, extract_s):
        self.rings, self.sources, self.layers = rings, sources, layers
        self.tri_uv, self.tri_src, self.tri_layer = tri_uv, tri_src, tri_layer
        self.counts, self.setup_s, self.extract_s = counts, setup_s, extract_s


def fast_peel(vertices, faces, camera, settings, layers=0, threads=1, repeat=1, bits=30, keep=True,
              cap_poly=None, cap_tri=None, max_vertices=64):
    """The restatement of the reference's fast CPU method (fast-bvh-peel
    --resolved-mesh). ``extract_s`` lists the seconds of each of ``repeat``
    extractions (the reference's algorithm scope); ``setup_s`` the projection
    and topology before them."""
    vertices = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 3)
    faces = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1, 3)
    camera = np.ascontiguousarray(camera, dtype=np.float64).reshape(4, 4)
    settings = np.ascontiguousarray(settings, dtype=np.float64).reshape(6)
    n = max(16, len(faces))
    cap_poly = cap_poly or 64 * n
    cap_tri = cap_tri or 256 * n
    counts = np.zeros(len(FAST_COUNT_FIELDS), np.int64)
    seconds = np.zeros(1 + repeat)
    if keep:
        poly_xy = np.zeros((cap_poly, max_vertices, 2))
        poly_n = np.zeros(cap_poly, np.int32)
        poly_src = np.zeros(cap_poly, np.int32)
        poly_layer = np.zeros(cap_poly, np.int32)
        tri_uv = np.zeros((cap_tri, 3, 2))
        tri_src = np.zeros(cap_tri, np.int32)
        tri_layer = np.zeros(cap_tri, np.int32)
        ptrs = [_p(poly_xy), _p(poly_n), _p(poly_src), _p(poly_layer), int(cap_poly), int(max_vertices),
                _p(tri_uv), _p(tri_src), _p(tri_layer), int(cap_tri)]
    else:
        ptrs = [None, None, None, None, 0, int(max_vertices), None, None, None, 0]
    f = lib().oracle_fast_peel
    f.restype = ctypes.c_int
    rc = f(_p(vertices), len(vertices), _p(faces), len(faces), _p(camera), _p(settings), int(layers), int(threads),
           int(repeat), int(bits), *ptrs, _p(counts), _p(seconds))
    if rc:
        raise ValueError("fast_peel: bits must be in [12, 30] and repeat positive")
    c = dict(zip(FAST_COUNT_FIELDS, map(int, counts)))
    if not keep:
        return FastResult(None, None, None, None, None, None, c, float(seconds[0]), seconds[1:].tolist())
    P, T = min(c["patches"], cap_poly), min(c["faces"], cap_tri)
    if (poly_n[:P] > max_vertices).any():
        raise ValueError("a fast-route patch exceeds max_vertices")
    rings = [poly_xy[i, : poly_n[i]].copy() for i in range(P)]
    return FastResult(rings, poly_src[:P].copy(), poly_layer[:P].copy(), tri_uv[:T].copy(), tri_src[:T].copy(),
                      tri_layer[:T].copy(), c, float(seconds[0]), seconds[1:].tolist())
