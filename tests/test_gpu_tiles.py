"""Tiled peeling (hip_peel.tiles) against the untiled extraction and the exact
oracle.

A tile peels the faces whose projected box meets it, with the tile as the
window (the reference clips to the window and makes clipped edges contours,
warp_peel/kernels.py:424-430, 532-533). The union over a partition must give
the untiled extraction's layers and, per (layer, source), its area; every
fragment must carry the exact rank of its source at its interior points, and
the grid audit must find each covering source once per point. The full-size
256-triangle random scene, whose arrangement outgrows any workspace, is
peeled this way inside a window at its densest and checked fragment by
fragment.
"""

import time

import numpy as np
import pytest

import oracle
from peel_checks import grid_points, interior_points, source_areas

pytestmark = pytest.mark.gpu


def ring_areas(vertices, offsets):
    """Signed shoelace areas of CSR rings (every ring non-empty), taken about
    each ring's first vertex so that the rounding scales with the ring's size,
    not with its distance from the origin."""
    if len(offsets) < 2:
        return np.zeros(0)
    local = vertices - np.repeat(vertices[offsets[:-1]], np.diff(offsets), axis=0)
    nxt = np.arange(1, len(vertices) + 1)
    nxt[offsets[1:] - 1] = offsets[:-1]
    cross = local[:, 0] * local[nxt, 1] - local[:, 1] * local[nxt, 0]
    return 0.5 * np.add.reduceat(cross, offsets[:-1])


def ring_perimeters(vertices, offsets):
    nxt = np.arange(1, len(vertices) + 1)
    nxt[offsets[1:] - 1] = offsets[:-1]
    return np.add.reduceat(np.hypot(*(vertices[nxt] - vertices).T), offsets[:-1])


def assert_counter_clockwise(vertices, offsets):
    """Every fragment has positive area, up to the rounding of the map from
    its tile's frame to the whole window (2^-53 per coordinate, so at most
    about 2^-52 times the perimeter)."""
    a = ring_areas(vertices, offsets)
    tol = 2.0**-50 * ring_perimeters(vertices, offsets)
    bad = np.nonzero(a <= -tol)[0]
    if len(bad):
        i = bad[0]
        print("[tiles] clockwise fragment", i, a[i], tol[i], vertices[offsets[i]:offsets[i + 1]].tolist(), flush=True)
    assert not len(bad), (len(bad), float(a[bad].min()))
    return a


def rank_mismatches(host, rings, sources, layers, parents=None):
    """Fragments whose interior points do not carry their source's exact rank
    (``parents`` maps resolved pieces to the sources of an unresolved scene
    given as ``host``)."""
    pts, owner = interior_points(rings)
    r = oracle.exact_ranks(*host, pts)
    bad, checked = [], 0
    for k in range(len(pts)):
        if r.point_flags[k]:
            continue
        a, b = r.offsets[k], r.offsets[k + 1]
        entries = dict(zip(r.sources[a:b].tolist(), r.ranks[a:b].tolist()))
        f = owner[k]
        src = int(sources[f]) if parents is None else int(parents[sources[f]])
        checked += 1
        if entries.get(src) != int(layers[f]):
            bad.append((int(f), src, int(layers[f]), entries.get(src)))
    return bad, checked


def tile_audit(host, tile, points):
    """Grid audit of one tile at the points strictly inside it; returns how
    many were audited. A tile without fragments must have no covering source
    at any of its points."""
    u0, v0, u1, v1 = tile.rect
    inside = (points[:, 0] > u0) & (points[:, 0] < u1) & (points[:, 1] > v0) & (points[:, 1] < v1)
    pts = points[inside]
    if not len(pts):
        return 0
    ranks = oracle.exact_ranks(*host, pts)
    if not len(tile.sources):
        assert ranks.offsets[-1] == 0, tile.rect
        return len(pts)
    audit = oracle.audit(tile.rings(), tile.sources, tile.layers, pts, ranks)
    assert audit["mismatches"] == 0, (tile.rect, audit)
    return len(pts)


def tile_triangles(host, tile, layer_count):
    """Triangles of one tile: counter-clockwise, tiling the fragments' area
    per (layer, source), and reprojecting onto their uv within 2e-9.

    The route emits triangles with doubled area above 1e-20 in the tile's
    frame (warp_peel/kernels.py:1101-1104); mapping to the whole window rounds
    each coordinate by up to 2^-53, which can flip a sliver whose doubled area
    is below about 8 * 2^-53 times its longest edge, so that much is admitted."""
    vertices, faces, camera, settings = host
    uv, world, tsrc, tlay = tile.triangles_uv, tile.triangles_world, tile.triangle_sources, tile.triangle_layers
    a, b = uv[:, 1] - uv[:, 0], uv[:, 2] - uv[:, 0]
    cross = a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]
    edge = np.abs(np.concatenate([a, b], axis=1)).max(axis=1, initial=0.0)
    assert (cross > -8 * 2.0**-53 * edge).all(), cross.min(initial=0.0)
    F = len(faces)
    tri = np.bincount(tlay * F + tsrc, weights=0.5 * cross, minlength=layer_count * F)
    frag = np.bincount(tile.layers * F + tile.sources, weights=ring_areas(tile.ring_vertices, tile.ring_offsets),
                       minlength=layer_count * F)
    assert np.abs(tri - frag).max(initial=0.0) < 1e-9
    if len(uv):
        pc = world @ camera[:3, :3].T + camera[:3, 3]
        xy = pc[..., :2] / pc[..., 2:3] if settings[4] else pc[..., :2]
        xy = (xy - settings[:2]) / (settings[2:4] - settings[:2])
        assert np.abs(xy - uv).max() < 2e-9
    return len(uv)


def pinhole_sphere():
    from hip_peel import scenes

    v, f = scenes.geodesic_sphere(1.0, 8, amplitude=0.3, seed=5)
    return v, f.astype(np.int32), scenes.pinhole_look_at((2.2, 1.1, 0.9)), np.array([-1.0, -1.0, 1.0, 1.0, 1.0, 0.1])


def random_resolved():
    from hip_peel import scenes

    return scenes.random_triangles_resolved(48, 0)[:4]


@pytest.mark.parametrize("route", ["events", "layered"])
@pytest.mark.parametrize("make", [random_resolved, pinhole_sphere], ids=["random-48-resolved", "pinhole-sphere"])
def test_tiled_matches_untiled(make, route):
    from hip_peel.tiles import TiledPeeler

    host = make()
    print(f"[tiles] {make.__name__} {route} faces {len(host[1])}", flush=True)
    whole = TiledPeeler(*host, route=route)
    whole.run()
    assert len(whole.results) == 1 and whole.splits == 0
    ref_vert, ref_off, ref_src, ref_lay = whole.fragments()
    L = whole.layers

    tiled = TiledPeeler(*host, tiles=(2, 2), max_faces=len(host[1]) // 6, route=route)
    tiled.run()
    assert tiled.splits > 0 and len(tiled.results) > 4
    assert tiled.layers == L
    rv, off, src, lay = tiled.fragments()
    F = len(host[1])
    got = np.bincount(lay * F + src, weights=ring_areas(rv, off), minlength=L * F)
    want = np.bincount(ref_lay * F + ref_src, weights=ring_areas(ref_vert, ref_off), minlength=L * F)
    # the layered route clips pieces repeatedly through constructed binary64
    # points, so its areas carry the 1e-9 tolerance of its own parity checks
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 if route == "events" else 1e-9)
    assert_counter_clockwise(rv, off)
    bad, checked = rank_mismatches(host, np.split(rv, off[1:-1]), src, lay)
    assert checked > 0 and not bad, bad[:5]
    # the untiled run carries the exact ranks too
    bad, checked = rank_mismatches(host, np.split(ref_vert, ref_off[1:-1]), ref_src, ref_lay)
    assert checked > 0 and not bad, bad[:5]
    points = grid_points(48)
    audited = sum(tile_audit(host, t, points) for t in tiled.results)
    inside = sum(int(((points[:, 0] > r[0]) & (points[:, 0] < r[2]) & (points[:, 1] > r[1])
                      & (points[:, 1] < r[3])).sum()) for r in (t.rect for t in tiled.results))
    assert audited == inside and inside > 0.2 * len(points)
    assert sum(tile_triangles(host, t, L) for t in tiled.results) > 0


def test_random_256_resolved_window_full_density():
    """The 256 random unit-cube triangles, resolved exactly (272,806 pieces),
    peeled by the event route inside a window of 1/128 x 1/128 of the scene's
    window at its densest, split into 2 x 2 tiles: every fragment's exact rank
    on the resolved and on the unresolved scene, the grid audit of every tile,
    per-source area conservation, and the triangles.

    The whole window is out of reach of the event route: tiles of 1/256 of
    the window near its centre already hold 25-90 thousand fragments under
    45-51 layers, about 10^9 fragments over the whole window."""
    from hip_peel import scenes
    from hip_peel.tiles import TiledPeeler, projected_boxes

    t0 = time.perf_counter()
    v, f, c, s, parents = scenes.random_triangles_resolved(256, 0, workers=16)
    print(f"[tiles] random 256 resolved: {len(f)} pieces in {time.perf_counter() - t0:.1f} s", flush=True)
    assert len(f) == 272806
    # 1/128 of the window per side, around the point (0.41, 0.285) of the window
    u0, v0 = 52 / 128, 36 / 128
    win = s.copy()
    win[0], win[2] = s[0] + u0 * (s[2] - s[0]), s[0] + (u0 + 1 / 128) * (s[2] - s[0])
    win[1], win[3] = s[1] + v0 * (s[3] - s[1]), s[1] + (v0 + 1 / 128) * (s[3] - s[1])
    host = (v, f, c, win)
    b = projected_boxes(*host)
    near = np.nonzero((b[:, 0] <= 1) & (b[:, 2] >= 0) & (b[:, 1] <= 1) & (b[:, 3] >= 0))[0]
    tiled = TiledPeeler(*host, tiles=(2, 2), route="events", scales=(4, 16, 64))
    tiled.run()
    rv, off, src, lay = tiled.fragments()
    print(f"[tiles] window: {len(near)} faces, {len(src)} fragments, {tiled.layers} layers, "
          f"{len(tiled.results)} tiles, {time.perf_counter() - t0:.1f} s", flush=True)
    assert len(src) > 50000 and tiled.layers > 30
    areas = assert_counter_clockwise(rv, off)
    got = np.bincount(src, weights=areas, minlength=len(f))
    sub = (v, f[near], c, win)
    want = np.zeros(len(f))
    want[near] = source_areas(*sub)
    assert np.abs(got - want).max() < 1e-9
    rings = np.split(rv, off[1:-1])
    bad, checked = rank_mismatches(host, rings, src, lay)
    assert checked > 0 and not bad, bad[:5]
    original = scenes.random_triangles(256, 0)
    bad, checked = rank_mismatches((*original[:3], win), rings, src, lay, parents=parents)
    assert checked > 0 and not bad, bad[:5]
    points = grid_points(48)
    assert sum(tile_audit(host, t, points) for t in tiled.results) == len(points)
    assert sum(tile_triangles(host, t, tiled.layers) for t in tiled.results) > 0
    print(f"[tiles] checked {checked} fragment points ({time.perf_counter() - t0:.1f} s)", flush=True)
