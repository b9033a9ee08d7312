"""Exact intersection resolver (hip_peel.resolve; the reference's
src/resolve_intersections.cpp over CGAL autorefine), checked on the CPU:

* golden vector: the reference's resolved `crossing` (results/fast/
  resolved-crossing.json, copied to tests/golden/) is reproduced exactly;
* provenance, as scripts/validate_fast.py:136-186 checks the CGAL output:
  every piece vertex lies on its source's plane and inside the source, and
  the pieces' area vectors sum exactly to the source's;
* resolution: no two pieces of different sources meet except at shared
  vertices and edges, decided with exact orient3d signs (a formulation the
  resolver itself does not use); the checker is pinned by the unresolved
  inputs, which it must reject;
* the reference's rejections (positive-area coincident surfaces, non-manifold
  or inconsistently oriented input) and the CLI's exit codes.
"""

import json
from fractions import Fraction as F
from itertools import combinations

import numpy as np
import pytest

from peel_checks import GOLDEN, fixture_path

CROSSING = ["crossing", "perspective_crossing"] + [f"random_{i:02}" for i in range(12)]


def resolve_module():
    from hip_peel import resolve

    return resolve


def sub(a, b):
    return [x - y for x, y in zip(a, b)]


def dot(a, b):
    return sum(x * y for x, y in zip(a, b))


def cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def orient3d(a, b, c, d):
    return dot(sub(b, a), cross(sub(c, a), sub(d, a)))


def exact(points):
    return [[F(x) for x in p] for p in points]


def provenance(data):
    """scripts/validate_fast.py:156-182, restated."""
    inputs = {t.get("id", i): exact(t["vertices"]) for i, t in enumerate(data["original_triangles"])}
    sums = {i: [F(0)] * 3 for i in inputs}
    for t in data["triangles"]:
        base = inputs[t["original_source_id"]]
        u, v = sub(base[1], base[0]), sub(base[2], base[0])
        uu, uv, vv = dot(u, u), dot(u, v), dot(v, v)
        vertices = exact(t["vertices"])
        for p in vertices:
            w = sub(p, base[0])
            assert dot(w, cross(u, v)) == 0
            b = (dot(w, u) * vv - dot(w, v) * uv) / (uu * vv - uv * uv)
            c = (dot(w, v) * uu - dot(w, u) * uv) / (uu * vv - uv * uv)
            assert b >= 0 and c >= 0 and b + c <= 1
        area = cross(sub(vertices[1], vertices[0]), sub(vertices[2], vertices[0]))
        sums[t["original_source_id"]] = [x + y for x, y in zip(sums[t["original_source_id"]], area)]
    for i, v in inputs.items():
        assert sums[i] == cross(sub(v[1], v[0]), sub(v[2], v[0])), i


def _inside_closed(p, x):
    """x (on p's plane) lies in closed triangle p: the signs of orient3d
    against p's edges and p's normal direction never disagree."""
    n = cross(sub(p[1], p[0]), sub(p[2], p[0]))
    top = [p[0][k] + n[k] for k in range(3)]
    s = [orient3d(p[i], p[(i + 1) % 3], top, x) for i in range(3)]
    return all(v >= 0 for v in s) or all(v <= 0 for v in s)


def improper_contacts(p, q):
    """Ways in which q touches p other than at shared vertices / edges."""
    bad = []
    for i in range(3):
        a, b = q[i], q[(i + 1) % 3]
        sa, sb = orient3d(*p, a), orient3d(*p, b)
        if sa * sb < 0:
            # the edge crosses p's plane strictly between its ends
            s = [orient3d(a, b, p[j], p[(j + 1) % 3]) for j in range(3)]
            if all(v >= 0 for v in s) or all(v <= 0 for v in s):
                bad.append(("edge pierces", i))
        if sa == 0 and a not in p and _inside_closed(p, a):
            bad.append(("vertex on face", i))
        if sa == 0 and sb == 0:
            # edge in p's plane: may not cross any edge of p properly
            n = cross(sub(p[1], p[0]), sub(p[2], p[0]))
            for j in range(3):
                c, d = p[j], p[(j + 1) % 3]
                top = [c[k] + n[k] for k in range(3)]
                o1, o2 = orient3d(c, d, top, a), orient3d(c, d, top, b)
                top2 = [a[k] + n[k] for k in range(3)]
                o3, o4 = orient3d(a, b, top2, c), orient3d(a, b, top2, d)
                if o1 * o2 < 0 and o3 * o4 < 0:
                    bad.append(("edges cross", i, j))
    return bad


def assert_resolved(pieces, sources):
    boxes = [(np.min(np.array(t, dtype=float), axis=0) - 1e-9, np.max(np.array(t, dtype=float), axis=0) + 1e-9)
             for t in pieces]
    for i, j in combinations(range(len(pieces)), 2):
        if sources[i] == sources[j]:
            continue
        if np.any(boxes[i][0] > boxes[j][1]) or np.any(boxes[j][0] > boxes[i][1]):
            continue
        bad = improper_contacts(pieces[i], pieces[j]) + improper_contacts(pieces[j], pieces[i])
        assert not bad, (i, j, bad)


def pieces_of(data):
    return [exact(t["vertices"]) for t in data["triangles"]], [t["original_source_id"] for t in data["triangles"]]


def canonical(triangles):
    return sorted((t["original_source_id"], tuple(sorted(tuple(F(x) for x in p) for p in t["vertices"])))
                  for t in triangles)


def test_golden_crossing_reproduced():
    R = resolve_module()
    data = json.loads(fixture_path("crossing").read_text())
    golden = json.loads((GOLDEN / "resolved-crossing.json").read_text())
    out = R.resolve_scene(data)
    assert canonical(out["triangles"]) == canonical(golden["triangles"])
    for key in ("input_triangles", "output_triangles", "marked_intersection_edges",
                "positive_area_coincident_surfaces_supported"):
        assert out["intersection_preprocessing"][key] == golden["intersection_preprocessing"][key], key
    assert out["original_triangles"] == data["triangles"]


@pytest.mark.parametrize("name", CROSSING)
def test_fixture_resolution(name):
    R = resolve_module()
    data = json.loads(fixture_path(name).read_text())
    out = R.resolve_scene(data)
    provenance(out)
    assert_resolved(*pieces_of(out))
    # each piece keeps the winding of its source
    inputs = {t["id"]: exact(t["vertices"]) for t in data["triangles"]}
    for t in out["triangles"]:
        src = inputs[t["original_source_id"]]
        v = exact(t["vertices"])
        assert dot(cross(sub(v[1], v[0]), sub(v[2], v[0])), cross(sub(src[1], src[0]), sub(src[2], src[0]))) > 0


def test_checker_rejects_unresolved_inputs():
    # the crossing fixtures' own triangles cross: the checker must say so
    for name in CROSSING[:4]:
        data = json.loads(fixture_path(name).read_text())
        tris = [exact(t["vertices"]) for t in data["triangles"]]
        with pytest.raises(AssertionError):
            assert_resolved(tris, list(range(len(tris))))


def test_binary64_soup():
    """A random binary64 soup (the random-triangles generator at 12
    triangles): exact pieces are resolved and conserve every source's area;
    the binary64 output keeps one parent per piece."""
    R = resolve_module()
    from hip_peel import scenes

    v, f, _, _ = scenes.random_triangles(12, seed=3)
    tris = [[tuple(F(float(x)) for x in v[i]) for i in face] for face in f]
    pieces, parents, marked = R.resolve_triangles(tris)
    assert marked > 0 and len(pieces) > len(tris)
    assert_resolved([list(map(list, t)) for t in pieces], parents)
    for s, t in enumerate(tris):
        total = [F(0)] * 3
        for piece, parent in zip(pieces, parents):
            if parent == s:
                total = [x + y for x, y in zip(total, cross(sub(piece[1], piece[0]), sub(piece[2], piece[0])))]
        assert total == cross(sub(t[1], t[0]), sub(t[2], t[0]))
    rv, rf, rp = R.resolve_arrays(v, f)
    assert rf.shape == (len(pieces), 3) and rv.shape == (3 * len(pieces), 3)
    np.testing.assert_array_equal(rp, parents)
    np.testing.assert_allclose(rv.reshape(-1, 3, 3), np.array(pieces, dtype=float), rtol=0, atol=0)


def test_parallel_refinement_matches_serial():
    R = resolve_module()
    from hip_peel import scenes

    v, f, _, _ = scenes.random_triangles(10, seed=1)
    tris = [[tuple(F(float(x)) for x in v[i]) for i in face] for face in f]
    one = R.resolve_triangles(tris, form="text")
    two = R.resolve_triangles(tris, workers=2, form="text")
    assert one == two


def test_coplanar_touching_is_split():
    """Two coplanar triangles sharing half of an edge: the T-junction
    vertices are inserted, and nothing else changes."""
    R = resolve_module()
    tris = [
        [(F(0), F(0), F(0)), (F(4), F(0), F(0)), (F(0), F(4), F(0))],
        [(F(2), F(0), F(0)), (F(6), F(0), F(0)), (F(4), F(-3), F(0))][::-1],
    ]
    pieces, parents, marked = R.resolve_triangles(tris)
    assert marked == 0
    assert len(pieces) == 4  # each source split once at the other's vertex on its edge
    assert_resolved([list(map(list, t)) for t in pieces], parents)


def test_rejections():
    R = resolve_module()
    overlap = {"window": [-5, -5, 5, 5], "triangles": [
        {"id": 0, "vertices": [[0, 0, 1], [2, 0, 1], [0, 2, 1]]},
        {"id": 1, "vertices": [[1, 0, 1], [3, 0, 1], [1, 2, 1]]},
    ]}
    with pytest.raises(ValueError, match="coincident"):
        R.resolve_scene(overlap)
    repeated = {"window": [-5, -5, 5, 5], "triangles": [
        {"id": 0, "vertices": [[0, 0, 1], [2, 0, 1], [0, 2, 1]]},
        {"id": 1, "vertices": [[0, 0, 1], [2, 0, 1], [0, -2, 3]]},
    ]}
    with pytest.raises(ValueError, match="non-manifold or inconsistently oriented"):
        R.resolve_scene(repeated)
    with pytest.raises(ValueError, match="degenerate"):
        R.resolve_triangles([[(F(0), F(0), F(0)), (F(1), F(1), F(1)), (F(2), F(2), F(2))]])


def test_cli(tmp_path):
    R = resolve_module()
    out = tmp_path / "resolved.json"
    assert R.main([str(fixture_path("crossing")), str(out)]) == 0
    data = json.loads(out.read_text())
    assert data["intersection_preprocessing"]["output_triangles"] == 4
    bad = tmp_path / "bad.json"
    bad.write_text(json.dumps({"window": [0, 0, 1, 1], "triangles": [{"vertices": [[0, 0, 0], [1, 0, 0]]}]}))
    assert R.main([str(bad), str(tmp_path / "x.json")]) == 1
    with pytest.raises(SystemExit) as e:
        R.main([str(bad)])
    assert e.value.code == 2
