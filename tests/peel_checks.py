"""Parity checks of a device extraction against the exact oracle.

The checks follow the reference's GPU test strategy (tests/gpu/test_peeler.py:
93-136 complete regions, triangle unions, orientation, reprojection, plane
distance; scripts/audit_gpu.py:22-137 exact ray audit), with exact per-point
ranks from ``oracle`` instead of shapely unions of CGAL regions:

* every fragment: at interior points (centroid and centroid-vertex midpoints),
  the exact rank of its source equals its layer;
* grid audit: at sample points, each covering source owns exactly one
  fragment containing the point, on the exact layer, and no other fragment
  contains it;
* per source, fragment areas sum to the area of the clipped projected source
  (no gap, no overlap);
* triangles: positive orientation, per (layer, source) area equal to the
  fragments' area, reprojection within 2e-9, distance to the source plane
  within 2e-8 (test_peeler.py:112-133).
"""

import json
from collections import defaultdict
from pathlib import Path

import numpy as np

import oracle

GOLDEN = Path(__file__).resolve().parent / "golden"
PACKAGE_DIR = Path(__file__).resolve().parents[1] / "exact-depth-peeling_amd"


def golden_layers():
    return json.loads((GOLDEN / "layers.json").read_text())


def fixture_path(name):
    return GOLDEN / "fixtures" / f"{name}.json"


def accepted_fixtures():
    manifest = json.loads((GOLDEN / "fixtures" / "manifest.json").read_text())
    return [f["name"] for f in manifest["fixtures"] if f["expected"] == "accept"]


def polygon_area(p):
    x, y = p[:, 0], p[:, 1]
    return 0.5 * float(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))


def clip_halfplane(poly, a, b, c):
    """Keep a*x + b*y + c >= 0 of a convex polygon (list of (x, y))."""
    out = []
    n = len(poly)
    for i in range(n):
        p, q = poly[i], poly[(i + 1) % n]
        dp, dq = a * p[0] + b * p[1] + c, a * q[0] + b * q[1] + c
        if dp >= 0:
            out.append(p)
        if (dp > 0 and dq < 0) or (dp < 0 and dq > 0):
            t = dp / (dp - dq)
            out.append((p[0] + t * (q[0] - p[0]), p[1] + t * (q[1] - p[1])))
    return out


def source_areas(vertices, faces, camera, settings):
    """Area (normalised window units) of each projected, near- and
    window-clipped source: the area its fragments must tile."""
    cam = vertices[faces] @ camera[:3, :3].T + camera[:3, 3]
    w = settings
    if not w[4] and len(faces):
        uv = (cam[..., :2] - w[:2]) / (w[2:4] - w[:2])
        if uv.min() >= 0.0 and uv.max() <= 1.0:  # nothing clipped: plain triangle areas
            a, b = uv[:, 1] - uv[:, 0], uv[:, 2] - uv[:, 0]
            return 0.5 * np.abs(a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0])
    areas = np.zeros(len(faces))
    for f in range(len(faces)):
        ring = [tuple(v) for v in cam[f]]
        if w[4]:
            out = []
            for i in range(3):
                p, q = ring[i], ring[(i + 1) % 3]
                dp, dq = p[2] - w[5], q[2] - w[5]
                if dp >= 0:
                    out.append(p)
                if (dp > 0 and dq < 0) or (dp < 0 and dq > 0):
                    t = dp / (dp - dq)
                    out.append(tuple(p[k] + t * (q[k] - p[k]) for k in range(3)))
            ring2 = [(p[0] / p[2], p[1] / p[2]) for p in out]
        else:
            ring2 = [(p[0], p[1]) for p in ring]
        ring2 = [((x - w[0]) / (w[2] - w[0]), (y - w[1]) / (w[3] - w[1])) for x, y in ring2]
        if len(ring2) < 3:
            continue
        a = polygon_area(np.array(ring2))
        if a < 0:
            ring2 = ring2[::-1]
        for plane in ((1, 0, 0), (-1, 0, 1), (0, 1, 0), (0, -1, 1)):
            ring2 = clip_halfplane(ring2, *plane)
            if len(ring2) < 3:
                break
        if len(ring2) >= 3:
            areas[f] = abs(polygon_area(np.array(ring2)))
    return areas


CLEANUP_WIDTH = 1e-10  # warp_peel/README.md:138-141: remnants narrower than ~1e-10 are discarded


def source_edge_distance(vertices, faces, camera, settings, points, pad=1e-6):
    """Distance (normalised window units) from each point to the nearest
    projected edge of any source of an orthographic scene without near
    clipping; only sources whose projected box, padded by `pad`, holds the
    point are examined (a nearer edge of a farther source would need a
    distance above `pad`, where this returns `pad`)."""
    assert not settings[4]
    cam = vertices[faces] @ camera[:3, :3].T + camera[:3, 3]
    uv = (cam[..., :2] - settings[:2]) / (settings[2:4] - settings[:2])
    lo, hi = uv.min(axis=1) - pad, uv.max(axis=1) + pad
    out = np.full(len(points), pad)
    for k, p in enumerate(np.asarray(points, dtype=np.float64).reshape(-1, 2)):
        near = np.nonzero(np.all(lo <= p, axis=1) & np.all(hi >= p, axis=1))[0]
        for f in near:
            for j in range(3):
                a, b = uv[f, j], uv[f, (j + 1) % 3]
                d = b - a
                t = np.clip(np.dot(p - a, d) / max(np.dot(d, d), 1e-300), 0.0, 1.0)
                out[k] = min(out[k], float(np.hypot(*(p - (a + t * d)))))
    return out


def fragments(peeler, info=None):
    info = info or peeler.read_status()
    count = info["polygons"]
    polys = peeler.output.numpy(count)
    sources = peeler.output_sources[:count].cpu().numpy()
    layers = peeler.output_layers[:count].cpu().numpy()
    keep = polys["n"] > 0
    rings = [p["v"][: p["n"]].copy() for p in polys[keep]]
    return rings, sources[keep], layers[keep]


def interior_points(rings):
    points, owner = [], []
    for i, r in enumerate(rings):
        c = r.mean(axis=0)
        points.append(c)
        owner.append(i)
        for v in r:
            points.append(0.5 * (c + v))
            owner.append(i)
    return np.array(points).reshape(-1, 2), np.array(owner, dtype=np.int64)


def grid_points(n=48, seed=7):
    rng = np.random.default_rng(seed)
    g = (np.arange(n) + 0.5) / n
    x, y = np.meshgrid(g, g)
    pts = np.stack([x.ravel(), y.ravel()], axis=1)
    return np.concatenate([pts, rng.random((n * n, 2))])


def check_extraction(peeler, host, *, expected_layers=None, grid=48, triangles=True, area_tol=1e-9,
                     orientation_tol=0.0):
    """Full parity check; returns a report dict and raises AssertionError on failure."""
    vertices, faces, camera, settings = host
    info = peeler.read_status()
    rings, srcs, lays = fragments(peeler, info)
    report = {"polygons": len(rings), "layers": info["layers"]}
    if expected_layers is not None:
        assert info["layers"] == expected_layers, (info["layers"], expected_layers)
    # 1) per-fragment exact ranks at interior points
    pts, owner = interior_points(rings) if rings else (np.zeros((0, 2)), np.zeros(0, np.int64))
    ranks = oracle.exact_ranks(vertices, faces, camera, settings, pts)
    bad, checked = [], 0
    for k in range(len(pts)):
        f = owner[k]
        a, b = ranks.offsets[k], ranks.offsets[k + 1]
        entries = dict(zip(ranks.sources[a:b].tolist(), ranks.ranks[a:b].tolist()))
        if ranks.point_flags[k]:
            continue
        checked += 1
        if entries.get(int(srcs[f])) != int(lays[f]):
            bad.append((k, int(srcs[f]), int(lays[f]), entries.get(int(srcs[f]))))
    report["fragment_points_checked"] = checked
    report["fragment_rank_mismatches"] = len(bad)
    assert not bad, bad[:10]
    # layer count: largest exact rank reached + 1
    if len(lays):
        assert info["layers"] == int(lays.max()) + 1
    # 2) grid audit (gaps, overlaps, wrong layers)
    if grid:
        gp = grid_points(grid)
        granks = oracle.exact_ranks(vertices, faces, camera, settings, gp)
        audit = oracle.audit(rings, srcs, lays, gp, granks)
        report["audit"] = {k: v for k, v in audit.items() if k != "examples"}
        assert audit["mismatches"] == 0, audit
    # 3) area conservation per source
    areas = source_areas(vertices, faces, camera, settings)
    got = np.zeros(len(faces))
    for r, s in zip(rings, srcs):
        a = polygon_area(r)
        assert a > 0, "fragment must be counter-clockwise"
        got[s] += a
    err = np.abs(got - areas)
    report["max_area_error"] = float(err.max()) if len(err) else 0.0
    assert report["max_area_error"] <= area_tol * max(1.0, float(areas.max(initial=0))), report
    # 4) triangles
    if triangles:
        report.update(check_triangles(peeler, info, rings, srcs, lays, host, orientation_tol))
    return report


def check_triangles(peeler, info, rings, srcs, lays, host, orientation_tol=0.0):
    """``orientation_tol`` admits triangles whose doubled signed area is within
    that noise level of zero (the fan's own cut-off is |cross| > EPS^2 =
    1e-20, kernels.py:1101-1104, below binary64 resolution for sub-segments
    of ~1e-15 next to near-collinear vertices)."""
    vertices, faces, camera, settings = host
    T = info["triangles"]
    tris = peeler.triangles_uv[:T].cpu().numpy()
    world = peeler.triangles_world[:T].cpu().numpy()
    tsrc = peeler.triangle_sources[:T].cpu().numpy()
    tlay = peeler.triangle_layers[:T].cpu().numpy()
    a, b = tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0]
    cross = a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]
    assert (cross > -orientation_tol).all(), ("triangles must be counter-clockwise", cross.min())
    tri_area = defaultdict(float)
    for k in range(T):
        tri_area[(int(tlay[k]), int(tsrc[k]))] += 0.5 * cross[k]
    frag_area = defaultdict(float)
    for r, s, l in zip(rings, srcs, lays):
        frag_area[(int(l), int(s))] += polygon_area(r)
    for key in frag_area.keys() | tri_area.keys():
        assert abs(frag_area.get(key, 0.0) - tri_area.get(key, 0.0)) < 1e-9, key
    if T:
        transformed = world @ camera[:3, :3].T + camera[:3, 3]
        uv = transformed[:, :, :2].copy()
        if settings[4]:
            uv /= transformed[:, :, 2, None]
        uv = (uv - settings[:2]) / (settings[2:4] - settings[:2])
        np.testing.assert_allclose(uv, tris, atol=2e-9, rtol=0)
        inputs = vertices[faces[tsrc]]
        normals = np.cross(inputs[:, 1] - inputs[:, 0], inputs[:, 2] - inputs[:, 0])
        normals /= np.linalg.norm(normals, axis=1)[:, None]
        distances = np.sum((world - inputs[:, :1]) * normals[:, None], axis=2)
        np.testing.assert_allclose(distances, 0, atol=2e-8, rtol=0)
    return {"triangles": T}
