"""Host side of tiled peeling (hip_peel.tiles): projected face boxes, the
quadtree partition and face culling, without a GPU."""

import numpy as np
import pytest

from hip_peel import scenes
from hip_peel.tiles import TiledPeeler, projected_boxes


def test_projected_boxes_orthographic_and_pinhole():
    v, f, c, s = scenes.random_triangles(32, 3)
    b = projected_boxes(v, f, c, s)
    uv = ((v @ c[:3, :3].T + c[:3, 3])[:, :2] - s[:2]) / (s[2:4] - s[:2])
    np.testing.assert_array_equal(b[:, :2], uv[f].min(axis=1))
    np.testing.assert_array_equal(b[:, 2:], uv[f].max(axis=1))
    assert b.min() >= 0 and b.max() <= 1
    # pinhole: division by depth; a face crossing the near plane is unbounded
    cam = scenes.pinhole_look_at((0.0, -3.0, 0.0))
    st = np.array([-1.0, -1.0, 1.0, 1.0, 1.0, 0.1])
    verts = np.array([[0.0, 0.0, 0.0], [0.5, 0.0, 0.0], [0.0, 0.0, 0.5], [0.0, -3.0, 0.2], [0.1, -2.0, 0.0],
                      [0.0, 1.0, 0.0]])
    faces = np.array([[0, 1, 2], [3, 4, 5]])
    b = projected_boxes(verts, faces, cam, st)
    pc = verts[:3] @ cam[:3, :3].T + cam[:3, 3]
    uv = (pc[:, :2] / pc[:, 2:3] + 1.0) / 2.0
    np.testing.assert_allclose(b[0], np.r_[uv.min(axis=0), uv.max(axis=0)], rtol=0, atol=1e-15)
    assert np.isinf(b[1]).all()


def test_quadtree_partition_and_culling():
    v, f, c, s = scenes.random_triangles(64, 1)
    t = TiledPeeler(v, f, c, s, tiles=(2, 4))
    rects = [(i / 2, j / 4, (i + 1) / 2, (j + 1) / 4) for i in range(2) for j in range(4)]
    quarters = [q for r in rects for q in t._quarters(r)]
    assert np.isclose(sum((q[2] - q[0]) * (q[3] - q[1]) for q in quarters), 1.0)
    # shared seams are identical numbers in both neighbours
    for r in rects:
        a, b, cc, d = t._quarters(r)
        assert a[2] == b[0] and a[3] == cc[1] and b[3] == d[1] and cc[2] == d[0]
    # a face is kept by every tile its (exactly computed) footprint meets
    rng = np.random.default_rng(0)
    pc = (v @ c[:3, :3].T + c[:3, 3])[:, :2]
    uv = (pc - s[:2]) / (s[2:4] - s[:2])
    for q in quarters:
        keep = set(t._faces_in(q).tolist())
        for k in range(len(f)):
            w = rng.dirichlet(np.ones(3), 64) @ uv[f[k]]
            inside = ((w[:, 0] > q[0]) & (w[:, 0] < q[2]) & (w[:, 1] > q[1]) & (w[:, 1] < q[3])).any()
            if inside:
                assert k in keep


def test_tile_grid_must_be_powers_of_two():
    v, f, c, s = scenes.random_triangles(4, 0)
    with pytest.raises(ValueError):
        TiledPeeler(v, f, c, s, tiles=(3, 2))


def test_tile_outputs_concatenate_in_window_coordinates():
    """fragments()/triangles() join the kept tiles' CSR rings and triangles,
    with ring offsets shifted by the vertices before them; a tile's rings are
    already in whole-window coordinates (the map is applied per tile)."""
    from hip_peel.tiles import Tile, _to_window

    v, f, c, s = scenes.random_triangles(4, 0)
    t = TiledPeeler(v, f, c, s, tiles=(2, 2))
    # a ring in tile (0.5, 0, 1, 0.5) at tile-local (0, 0), (1, 0), (0, 1)
    local = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
    a = _to_window((0.5, 0.0, 1.0, 0.5), local)
    np.testing.assert_array_equal(a, [[0.5, 0.0], [1.0, 0.0], [0.5, 0.5]])

    def tile(rect, rings, sources, layers):
        counts = [len(r) for r in rings]
        return Tile(rect=rect, faces=np.arange(4), scale=1, info={"layers": max(layers) + 1},
                    ring_vertices=np.concatenate(rings), ring_offsets=np.r_[0, np.cumsum(counts)],
                    sources=np.array(sources), layers=np.array(layers, dtype=np.int32),
                    triangles_uv=np.zeros((len(rings), 3, 2)), triangles_world=np.zeros((len(rings), 3, 3)),
                    triangle_sources=np.array(sources), triangle_layers=np.array(layers, dtype=np.int32),
                    device_bytes=0)

    r1 = [np.array([[0.0, 0.0], [0.1, 0.0], [0.0, 0.1]]), np.array([[0.2, 0.2], [0.3, 0.2], [0.3, 0.3], [0.2, 0.3]])]
    r2 = [np.array([[0.6, 0.6], [0.7, 0.6], [0.6, 0.7]])]
    t.results = [tile((0, 0, 0.5, 0.5), r1, [0, 1], [0, 2]), tile((0.5, 0.5, 1, 1), r2, [3], [1])]
    rv, off, src, lay = t.fragments()
    np.testing.assert_array_equal(off, [0, 3, 7, 10])
    rings = np.split(rv, off[1:-1])
    for got, want in zip(rings, r1 + r2):
        np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(src, [0, 1, 3])
    np.testing.assert_array_equal(lay, [0, 2, 1])
    assert t.layers == 3
    uv, world, tsrc, tlay = t.triangles()
    assert uv.shape == (3, 3, 2) and world.shape == (3, 3, 3)
    np.testing.assert_array_equal(tsrc, [0, 1, 3])
    t.results = []
    assert len(t.fragments()[2]) == 0 and t.layers == 0
