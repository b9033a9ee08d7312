"""Pins the exact oracle before it is trusted (CPU only).

* Golden vectors: the reference's arrangement oracle records, for every cell
  of the exact arrangement, an interior point and the exact front-to-back
  source order (results/examples/*/arrangement/exact.json, written by
  src/arrangement_oracle.cpp:55-69). The oracle must reproduce every order.
* Independent restatement: a pure-Python Fraction evaluation of the
  reference's exact scene semantics (src/camera.cpp:74-141: near clip in camera
  space, projection, affine score fit `fit`, smaller score is nearer) on small
  random scenes, orthographic and perspective. The oracle evaluates depth by
  barycentric / ray-plane formulas instead, so a shared formula mistake cannot
  pass both.
* Closed forms: the identical-footprint stack (rank k for the k-th triangle).
* Planted mistakes: flipping the depth order or dropping a covering source in
  the audit input must be reported.
"""

import json
import random
from fractions import Fraction as Fr

import numpy as np
import pytest

import oracle
from peel_checks import GOLDEN

EXAMPLES = [
    "single",
    "holes",
    "closed_object",
    "components",
    "cover_2",
    "crossing",
    "cyclic",
    "perspective_near",
    "perspective_torus",
]


@pytest.mark.parametrize("case", EXAMPLES)
def test_golden_cells(case):
    V, F, C, S, labels = oracle.load_scene(GOLDEN / "fixtures" / f"{case}.json")
    g = json.loads((GOLDEN / "exact" / f"{case}.json").read_text())
    w = [Fr(x) for x in g["window"]]
    pts, want = [], []
    for c in g["oracle_cells"]:
        x, y = (Fr(v) for v in c["interior_point"])
        pts.append([float((x - w[0]) / (w[2] - w[0])), float((y - w[1]) / (w[3] - w[1]))])
        want.append(c["depth_order"])
    r = oracle.exact_ranks(V, F, C, S, np.array(pts))
    for i, order in enumerate(want):
        assert [int(labels[s]) for s in r.at(i)] == order, (case, i)
    assert len(want) > 0


def test_golden_layer_counts_bound_cells():
    # the deepest cell of each example has exactly the recorded layer count
    layers = json.loads((GOLDEN / "layers.json").read_text())
    for case in EXAMPLES:
        g = json.loads((GOLDEN / "exact" / f"{case}.json").read_text())
        deepest = max(len(c["depth_order"]) for c in g["oracle_cells"])
        assert deepest == layers[case] == len(g["layers"]), case


# --- independent Fraction restatement of src/camera.cpp:74-141 ---------------
def frac_scene(V, F, C, S):
    persp, near = S[4] != 0, Fr(S[5])
    M = [[Fr(C[r][c]) for c in range(4)] for r in range(3)]
    sources = []
    for f in F:
        cam = []
        for vi in f:
            p = [Fr(x) for x in V[vi]]
            cam.append([M[r][0] * p[0] + M[r][1] * p[1] + M[r][2] * p[2] + M[r][3] for r in range(3)])
        ring = cam
        if persp:
            ring = []
            for i in range(3):
                a, b = cam[i], cam[(i + 1) % 3]
                da, db = a[2] - near, b[2] - near
                if da >= 0:
                    ring.append(a)
                if (da < 0 < db) or (db < 0 < da):
                    t = da / (da - db)
                    ring.append([a[k] + (b[k] - a[k]) * t for k in range(3)])
        if len(ring) < 3:
            sources.append(None)
            continue
        proj = [(p[0] / p[2], p[1] / p[2]) if persp else (p[0], p[1]) for p in ring]
        score = [-1 / p[2] if persp else p[2] for p in ring]
        k = 2
        while k < len(proj) and orient(proj[0], proj[1], proj[k]) == 0:
            k += 1
        if k == len(proj):
            sources.append(None)
            continue
        (x0, y0), (x1, y1), (x2, y2) = proj[0], proj[1], proj[k]
        z0, z1, z2 = score[0], score[1], score[k]
        X1, Y1, X2, Y2 = x1 - x0, y1 - y0, x2 - x0, y2 - y0
        det = X1 * Y2 - X2 * Y1
        a = ((z1 - z0) * Y2 - (z2 - z0) * Y1) / det
        b = (X1 * (z2 - z0) - X2 * (z1 - z0)) / det
        c = z0 - a * x0 - b * y0
        if orient(*proj[:3]) < 0:
            proj = proj[::-1]
        sources.append((proj, (a, b, c)))
    return sources


def orient(a, b, c):
    return (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])


def frac_ranks(sources, S, q):
    w = [Fr(x) for x in S[:4]]
    x = w[0] + Fr(q[0]) * (w[2] - w[0])
    y = w[1] + Fr(q[1]) * (w[3] - w[1])
    hits, edge = [], False
    for i, src in enumerate(sources):
        if src is None:
            continue
        proj, (a, b, c) = src
        signs = [orient(proj[j], proj[(j + 1) % len(proj)], (x, y)) for j in range(len(proj))]
        if min(signs) < 0:
            continue
        if min(signs) == 0:
            edge = True
            continue
        hits.append((a * x + b * y + c, i))
    hits.sort()
    return [i for _, i in hits], edge


def random_scene(rng, n, perspective):
    V = np.array([rng.uniform(-1, 1) for _ in range(9 * n)]).reshape(-1, 3)
    if perspective:
        V[:, 2] = np.array([rng.uniform(0.2, 3.0) for _ in range(3 * n)])
    F = np.arange(3 * n, dtype=np.int32).reshape(-1, 3)
    C = np.eye(4)
    if perspective:
        C[0, 3], C[1, 3] = 0.125, -0.25
    S = np.array([-1.5, -1.5, 1.5, 1.5, float(perspective), 0.5])
    return V, F, C, S


@pytest.mark.parametrize("perspective", [False, True])
@pytest.mark.parametrize("seed", range(4))
def test_independent_fraction_restatement(seed, perspective):
    rng = random.Random(seed)
    V, F, C, S = random_scene(rng, 10, perspective)
    pts = np.array([[rng.random(), rng.random()] for _ in range(120)])
    r = oracle.exact_ranks(V, F, C, S, pts)
    sources = frac_scene(V, F, C, S)
    compared = 0
    for i, q in enumerate(pts):
        order, edge = frac_ranks(sources, S, q)
        if edge or r.point_flags[i] & oracle.FLAG_EDGE:
            continue
        assert r.at(i) == order, (i, r.at(i), order)
        compared += 1
    assert compared > 100


def stack_scene(n):
    tri = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
    V = np.zeros((n, 3, 3))
    V[:, :, :2] = tri
    V[:, :, 2] = (np.arange(n) * 2.0**-20)[:, None]
    F = np.arange(3 * n, dtype=np.int32).reshape(-1, 3)
    S = np.array([-1 / 16, -1 / 16, 17 / 16, 17 / 16, 0.0, 0.01])
    return V.reshape(-1, 3), F, np.eye(4), S


def test_stack_closed_form():
    n = 257
    V, F, C, S = stack_scene(n)
    pts = np.array([[0.3, 0.3], [0.1, 0.6], [0.55, 0.2]])
    r = oracle.exact_ranks(V, F, C, S, pts)
    for i in range(len(pts)):
        assert r.at(i) == list(range(n))
        assert list(r.ranks[r.offsets[i] : r.offsets[i + 1]]) == list(range(n))
    # outside the footprint (x + y > 17/16 * ... in window units) nothing covers
    r = oracle.exact_ranks(V, F, C, S, np.array([[0.9, 0.9]]))
    assert r.at(0) == []


def test_edge_and_tie_flags():
    V, F, C, S = stack_scene(2)
    # put the point exactly on the hypotenuse x + y = 1 (window units)
    q = (0.5 + 1 / 16) / (18 / 16)
    r = oracle.exact_ranks(V, F, C, S, np.array([[q, q]]))
    assert r.point_flags[0] & oracle.FLAG_EDGE
    V2 = V.copy()
    V2[3:, 2] = 0.0  # two coincident triangles: equal depth everywhere
    r = oracle.exact_ranks(V2, F, C, S, np.array([[0.3, 0.3]]))
    assert r.point_flags[0] & oracle.FLAG_TIE
    assert r.at(0) == [0, 1]


def test_audit_detects_planted_mistakes():
    V, F, C, S = stack_scene(3)
    w = S
    ring = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
    ring = (ring - w[:2]) / (w[2:4] - w[:2])
    rings = [ring.copy() for _ in range(3)]
    pts = np.array([[0.2, 0.2], [0.3, 0.5], [0.6, 0.1]])
    ranks = oracle.exact_ranks(V, F, C, S, pts)
    ok = oracle.audit(rings, [0, 1, 2], [0, 1, 2], pts, ranks)
    assert ok["mismatches"] == 0 and ok["pairs_checked"] == 9
    swapped = oracle.audit(rings, [0, 1, 2], [1, 0, 2], pts, ranks)
    assert swapped["wrong_layer"] == 6
    missing = oracle.audit(rings[:2], [0, 1], [0, 1], pts, ranks)
    assert missing["missing"] == 3
    dup = oracle.audit(rings + [ring], [0, 1, 2, 2], [0, 1, 2, 2], pts, ranks)
    assert dup["duplicate"] == 3
    spurious = oracle.audit(rings + [ring], [0, 1, 2, 7], [0, 1, 2, 3], pts, ranks)
    assert spurious["spurious"] == 3


def test_audit_boundary_distance():
    """A fragment whose edge passes within 1e-11 of a sample point is
    ambiguous there (audit_gpu.py:78-80 skips such points), while a fragment
    short of the point by 1e-9 is a reported gap. The rational fixture
    random_04 places a grid point on a source edge in real arithmetic; its
    binary64 coordinates sit ~1e-17 inside the exact source and outside the
    fragment's rounded edge."""
    V, F, C, S = stack_scene(1)
    ring = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
    ring = (ring - S[:2]) / (S[2:4] - S[:2])
    pts = np.array([[0.3, 0.2]])
    ranks = oracle.exact_ranks(V, F, C, S, pts)
    assert ranks.at(0) == [0]
    # move the hypotenuse (outward normal) past the point
    normal = np.array([1.0, 1.0]) / np.sqrt(2.0)
    edge = ring[1:]
    d_point = float(np.dot(pts[0] - edge[0], normal))
    # the point ends up outside the moved edge by 1e-12, then by 1e-9
    for offset, missing in ((d_point - 1e-12, 0), (d_point - 1e-9, 1)):
        moved = ring.copy()
        moved[1:] += offset * normal
        report = oracle.audit([moved], [0], [0], pts, ranks)
        assert report["missing"] == missing, (offset, report)
    from peel_checks import fixture_path, grid_points

    host = oracle.load_scene(fixture_path("random_04"))[:4]
    cpu = oracle.cpu_event_peel(*host, threads=1)
    gp = grid_points(48)
    report = oracle.audit(cpu.rings, cpu.sources, cpu.layers, gp, oracle.exact_ranks(*host, gp))
    assert report["mismatches"] == 0 and report["pairs_on_fragment_boundary"] > 0


def test_source_edge_distance():
    """The cleanup-band helper of the merged-batch check: distances to the
    nearest projected edge, in normalised window units, against hand values."""
    from peel_checks import source_edge_distance

    V = np.array([[0.0, 0.0, 1.0], [2.0, 0.0, 1.0], [0.0, 2.0, 1.0], [3.0, 3.0, 2.0]])
    F = np.array([[0, 1, 2], [1, 3, 2]])
    S = np.array([0.0, 0.0, 4.0, 4.0, 0.0, 0.1])
    pts = np.array([[0.25, 0.1], [0.1, 0.1], [0.25, 0.25], [0.9, 0.9]])
    d = source_edge_distance(V, F, np.eye(4), S, pts, pad=0.2)
    # (0.25, 0.1): 0.1 above the edge y = 0; (0.1, 0.1): 0.1 from two edges;
    # (0.25, 0.25): on the shared edge x + y = 0.5; (0.9, 0.9): 0.21 from the
    # nearest vertex, beyond the pad
    np.testing.assert_allclose(d, [0.1, 0.1, 0.0, 0.2], atol=1e-15)
