"""Checks the reference's GPU tests make that are restated here through the
C-ABI (tests/gpu/test_peeler.py):

* the captured graph holds only kernel, empty and event nodes: no copies,
  memsets, host callbacks or allocations (test_peeler.py:197-202), for both
  routes and for a batch graph;
* per (layer, source), the fragments cover the exact regions the reference's
  arrangement oracle recorded (results/examples/*/arrangement/exact.json,
  extracted to tests/golden/exact; test_peeler.py:75-83), by area;
* the numerical regression cases of the device's ring functions
  (test_peeler.py:331-375): a round-off backtracking edge does not erase a
  polygon, and a clipping edge of 2e-12 keeps its position.
"""

import ctypes
import json
from collections import defaultdict
from fractions import Fraction

import numpy as np
import pytest

import oracle
from peel_checks import GOLDEN, fixture_path, fragments, polygon_area

pytestmark = pytest.mark.gpu

ALLOWED_NODES = {"kernel", "empty", "event_record", "wait_event"}


def scene(name):
    from hip_peel.io import load_scene

    return load_scene(fixture_path(name))


@pytest.mark.parametrize("route", ["events", "layered"])
@pytest.mark.parametrize("name", ["perspective_torus", "closed_object"])
def test_graph_holds_only_launches(route, name):
    from hip_peel import DevicePeeler, EventPeeler

    arrays = scene(name)[0]
    p = (EventPeeler if route == "events" else DevicePeeler)(*arrays)
    p.run()
    p.read_status()
    nodes = p.graph_nodes()
    print(f"[graph] {route} {name}: {nodes}", flush=True)
    assert set(nodes) <= ALLOWED_NODES, nodes
    assert nodes["kernel"] > 10
    p.run()  # replays stay valid after the census
    p.read_status()


def test_batch_graph_holds_only_launches():
    from hip_peel import BatchPeeler, EventPeeler

    peelers = [EventPeeler(*scene(n)[0]) for n in ("holes", "cover_2", "perspective_torus")]
    batch = BatchPeeler(peelers)
    batch.run()
    batch.read_status()
    nodes = batch.graph_nodes()
    print(f"[graph] batch of 3: {nodes}", flush=True)
    assert set(nodes) <= ALLOWED_NODES, nodes
    single = peelers[0].graph_nodes()["kernel"]
    assert nodes["kernel"] >= 3 * min(single, 1)


def _area(points):
    """Exact doubled-then-halved shoelace area of a rational ring."""
    pts = [(Fraction(x), Fraction(y)) for x, y in points]
    s = sum(a[0] * b[1] - a[1] * b[0] for a, b in zip(pts, pts[1:] + pts[:1]))
    return abs(s) / 2


def golden_regions(name):
    data = json.loads((GOLDEN / "exact" / f"{name}.json").read_text())
    w = [Fraction(x) for x in data["window"]]
    scale = (w[2] - w[0]) * (w[3] - w[1])
    out = defaultdict(Fraction)
    for layer in data["layers"]:
        for r in layer["regions"]:
            for d in r["domain"]:
                a = _area(d["outer"]) - sum(_area(h) for h in d["holes"])
                out[(layer["layer_index"] - 1, r["source_id"])] += a / scale
    return {k: float(v) for k, v in out.items()}


GOLDEN_CASES = ["single", "holes", "closed_object", "components", "cover_2", "cyclic", "perspective_near",
                "perspective_torus", "crossing"]


@pytest.mark.parametrize("route", ["events", "layered"])
@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_regions_match_golden_exact(route, name):
    """Per (layer, source id) the device fragments' area equals the exact
    region's, within 1e-9 of the window (the reference's tolerance on the
    symmetric difference); together with the exact rank at every fragment's
    interior points (test_gpu_fixtures) this pins the regions themselves.
    `crossing` is resolved exactly first and its pieces map back to their
    parent sources."""
    from hip_peel import DevicePeeler, EventPeeler
    from hip_peel.io import upload

    want = golden_regions(name)
    V, F, C, S, labels = oracle.load_scene(fixture_path(name))
    parents = np.arange(len(F))
    if name == "crossing":
        from hip_peel.resolve import resolve_arrays

        V, F, parents = resolve_arrays(V, F)
    p = (EventPeeler if route == "events" else DevicePeeler)(*upload(V, F, C, S))
    p.run()
    rings, srcs, lays = fragments(p)
    got = defaultdict(float)
    for r, s, l in zip(rings, srcs, lays):
        got[(int(l), int(labels[parents[s]]))] += polygon_area(r)
    for key in set(got) | set(want):
        assert abs(got.get(key, 0.0) - want.get(key, 0.0)) <= 1e-9, (key, got.get(key), want.get(key))
    assert set(k for k, v in want.items() if v > 1e-9) <= set(got)


def _ring_op(op, ring, param=None):
    import torch

    from hip_peel import _lib

    lib = _lib.load()
    xy = torch.tensor(np.asarray(ring, dtype=np.float64), device="cuda").contiguous()
    prm = torch.tensor(np.asarray(param if param is not None else [0.0] * 4, dtype=np.float64), device="cuda")
    out = torch.zeros((2, 32, 2), dtype=torch.float64, device="cuda")
    n = torch.zeros(2, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()
    _lib.check(lib.peel_ring_test(op, xy.data_ptr(), len(ring), prm.data_ptr(), out.data_ptr(), n.data_ptr(),
                                  ctypes.c_void_p(stream.cuda_stream)))
    stream.synchronize()
    n = n.cpu().numpy()
    o = out.cpu().numpy()
    return o[0, : n[0]], o[1, : n[1]]


def test_roundoff_edge_does_not_erase_polygon():
    """test_peeler.py:331-348: a vertex 2e-12 back along an edge."""
    ring = [(0.0, 0.0), (1.0, 0.0), (1.0 - 2e-12, 0.0), (1.0, 1.0), (0.0, 1.0)]
    cleaned, _ = _ring_op(0, ring)
    assert len(cleaned) >= 3
    assert abs(abs(polygon_area(cleaned)) - 1.0) < 1e-10
    # the layered route's strict-convexity rule removes the backtracking vertex
    convex, _ = _ring_op(2, cleaned)
    assert len(convex) == 4 and abs(polygon_area(convex) - 1.0) < 1e-10


def test_short_clipping_edge_keeps_its_position():
    """test_peeler.py:351-375: splitting a 2e-6 square through its centre by an
    edge of length ~2.2e-12 evaluated in the edge's own frame halves it."""
    a = np.array([0.3, 0.3])
    h = 1e-6
    square = [a + (-h, -h), a + (h, -h), a + (h, h), a + (-h, h)]
    b = a + (1e-12, 2e-12)
    inside, outside = _ring_op(1, square, [a[0], a[1], b[0], b[1]])
    # areas relative to the first vertex (the shoelace at 0.3 would cancel
    # ~1e-17 of a 4e-12 area; the reference measures with GEOS, which shifts too)
    ai, ao = abs(polygon_area(inside - inside[0])), abs(polygon_area(outside - outside[0]))
    assert abs(ai - ao) < 1e-20
    assert abs(ai + ao - 4e-12) < 1e-20
