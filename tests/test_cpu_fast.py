"""The restatement of the reference's fast CPU method (fast-bvh-peel
--resolved-mesh, oracle/cpu_fast.cpp), the timed CPU baseline, pinned on the
CPU: the mesh topology and rank classification counts the reference recorded
for every fixture (tests/golden/fast_records.json), exact ranks away from
source boundaries (the reference's fast audit skips two grid cells around
them), area conservation to the grid's resolution, and closed forms."""

import json

import numpy as np
import pytest

import oracle
from peel_checks import GOLDEN, fixture_path, interior_points, polygon_area, source_areas

RECORDS = json.loads((GOLDEN / "fast_records.json").read_text())
CELL = 2.0**-30  # one grid cell of the normalised window


def projected_segments(vertices, faces, camera, settings):
    """Window-normalised projected edges of every source (orthographic or
    pinhole, no near clipping: the fixtures that clip keep their vertices in
    front or are checked only away from the clipped edges)."""
    cam = vertices[faces] @ camera[:3, :3].T + camera[:3, 3]
    xy = cam[..., :2] / (cam[..., 2:3] if settings[4] else 1.0)
    uv = (xy - settings[:2]) / (settings[2:4] - settings[:2])
    a = uv.reshape(-1, 2)
    b = np.roll(uv, -1, axis=1).reshape(-1, 2)
    return a, b


def boundary_distance(points, a, b):
    d = b - a
    out = np.empty(len(points))
    for i in range(0, len(points), 512):
        p = points[i : i + 512, None, :]
        t = np.clip(((p - a) * d).sum(-1) / np.maximum((d * d).sum(-1), 1e-300), 0.0, 1.0)
        out[i : i + 512] = np.linalg.norm(p - (a + t[..., None] * d), axis=-1).min(axis=1)
    return out


def check_fast(result, host, cells=2, per_layer=128):
    """The reference's fast audit (src/fast_reference.cpp:7-95): per layer, up
    to `per_layer` fan triangles evenly spaced through the layer's output; at
    each one's centroid, unless it lies within `cells` grid cells of a source
    edge, the exact front-to-back order must hold the triangle's source at the
    triangle's layer. (Interior points of every fragment are not used: the
    fast method keeps ribbons about two cells wide where a hidden part
    collapsed on the grid, and those carry the previous layer.) Returns
    (checked, mismatches)."""
    pts, want = [], []
    for layer in range(result.counts["layers"]):
        idx = np.nonzero(result.tri_layer == layer)[0]
        count = min(per_layer, len(idx))
        for j in range(count):
            k = idx[j * len(idx) // count]
            pts.append(result.tri_uv[k].mean(axis=0))
            want.append((int(result.tri_src[k]), layer))
    pts = np.array(pts).reshape(-1, 2)
    far = boundary_distance(pts, *projected_segments(*host)) > cells * CELL if len(pts) else np.zeros(0, bool)
    ranks = oracle.exact_ranks(*host, pts)
    checked, bad = 0, []
    for k in np.nonzero(far)[0]:
        if ranks.point_flags[k]:
            continue
        a, b = ranks.offsets[k], ranks.offsets[k + 1]
        entries = dict(zip(ranks.sources[a:b].tolist(), ranks.ranks[a:b].tolist()))
        checked += 1
        if entries.get(want[k][0]) != want[k][1]:
            bad.append((int(k), *want[k], entries.get(want[k][0])))
    return checked, bad


@pytest.mark.parametrize("name", sorted(RECORDS))
def test_fast_matches_reference_records(name):
    V, F, C, S, _ = oracle.load_scene(fixture_path(name))
    r = oracle.fast_peel(V, F, C, S)
    want = RECORDS[name]
    assert r.counts["error"] == 0
    assert r.counts["layers"] == want["layers"]
    for k, v in want["topology"].items():
        assert r.counts[k] == v, k
    for k, v in want["classification"].items():
        assert r.counts["residual_sources" if k == "sources_requiring_clipping" else k] == v, k


@pytest.mark.parametrize("name", sorted(RECORDS))
def test_fast_ranks_and_areas(name):
    V, F, C, S, _ = oracle.load_scene(fixture_path(name))
    host = (V, F, C, S)
    r = oracle.fast_peel(*host, threads=2)
    checked, bad = check_fast(r, host)
    assert not bad, bad[:5]
    # per-source area to the grid's resolution: ribbons thinner than two cells
    # are dropped (fast_peel.cpp:92-104)
    got = np.zeros(len(F))
    for ring, s in zip(r.rings, r.sources):
        if len(ring):
            got[s] += polygon_area(ring)
    want = source_areas(*host)
    assert np.abs(got - want).max(initial=0) <= 1e-6 * max(1.0, want.max(initial=0))
    # fan triangles tile the fragments per (layer, source)
    if len(r.tri_uv):
        a, b = r.tri_uv[:, 1] - r.tri_uv[:, 0], r.tri_uv[:, 2] - r.tri_uv[:, 0]
        tri = np.bincount(r.tri_layer * len(F) + r.tri_src, weights=0.5 * (a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]),
                          minlength=(r.counts["layers"] + 1) * len(F))
        frag = np.zeros_like(tri)
        for ring, s, l in zip(r.rings, r.sources, r.layers):
            if len(ring):
                frag[l * len(F) + s] += polygon_area(ring)
        np.testing.assert_allclose(tri, frag, atol=1e-12, rtol=0)


def test_fast_stack_closed_form():
    from test_oracle import stack_scene

    V, F, C, S = stack_scene(200)
    r = oracle.fast_peel(V, F, C, S, threads=2)
    assert r.counts["layers"] == 200 and r.counts["patches"] == 200
    np.testing.assert_array_equal(r.layers, r.sources)


def test_fast_torus_knot_exact_away_from_boundaries():
    """A closed mesh: most sources take their rank from one ray (the
    classification path), the rest are clipped; both are exact."""
    from hip_peel import scenes

    host = scenes.torus_knot(segments=48, sides=24)
    r = oracle.fast_peel(*host, threads=4)
    assert r.counts["error"] == 0
    assert r.counts["constant_rank_sources"] > 0 and r.counts["point_query_candidates"] > 0
    assert r.counts["boundary_edges"] == 0 and r.counts["nonmanifold_or_inconsistent_edges"] == 0
    checked, bad = check_fast(r, host)
    assert checked > 300 and not bad, bad[:5]
    event = oracle.cpu_event_peel(*host, threads=4)
    assert r.counts["layers"] == event.counts["layers"]


def test_fast_threads_agree():
    """The partition does not depend on the thread count (each target's cuts
    depend on its own ring and the fixed hierarchies)."""
    from hip_peel import scenes

    host = scenes.torus_knot(segments=32, sides=16)
    one = oracle.fast_peel(*host, threads=1, repeat=2)
    many = oracle.fast_peel(*host, threads=4)
    assert one.counts == many.counts
    assert len(one.extract_s) == 2

    def rec(r):
        return sorted((int(s), int(l), ring.tobytes()) for ring, s, l in zip(r.rings, r.sources, r.layers))

    assert rec(one) == rec(many)
