"""Host-side logic without a GPU: scene parsing, capacities, error decoding,
procedural scenes."""

import json

import numpy as np
import pytest

import oracle
from peel_checks import GOLDEN, fixture_path


def test_scene_arrays_match_independent_reader():
    from hip_peel.io import scene_arrays

    for name in ("perspective_torus", "camera_transform", "binary_float", "holes"):
        data = json.loads(fixture_path(name).read_text())
        ours = scene_arrays(data)
        theirs = oracle.load_scene(data)
        for a, b in zip(ours, theirs):
            np.testing.assert_array_equal(np.asarray(a), np.asarray(b))


def test_scene_validation_errors():
    from hip_peel.io import scene_arrays

    data = json.loads(fixture_path("single").read_text())
    bad = dict(data, window=["1", "0", "0", "1"])
    with pytest.raises(ValueError, match="window"):
        scene_arrays(bad)
    dup = dict(data, triangles=data["triangles"] * 2)
    with pytest.raises(ValueError, match="unique"):
        scene_arrays(dup)
    cam = dict(data, camera={"type": "fisheye"})
    with pytest.raises(ValueError, match="camera"):
        scene_arrays(cam)


def test_capacities_match_reference_defaults():
    from hip_peel import Capacities, EventCapacities

    c = Capacities.for_mesh(10)
    assert (c.residual, c.polygons, c.vertices, c.triangles, c.workers, c.pieces_per_worker) == (
        64, 256, 1536, 1024, 16, 512)
    assert Capacities.for_mesh(10000).pieces_per_worker == 2048
    e = EventCapacities.for_mesh(1000)
    assert (e.regions, e.hits, e.vertices, e.triangles, e.rounds, e.layers, e.edge_vertices) == (
        8000, 32000, 32000, 32000, 4096, 1000, 128)


def test_procedural_scenes_are_closed_meshes():
    from hip_peel import scenes

    v, f, c, s = scenes.torus_knot()
    assert v.shape == (32768, 3) and f.shape == (65536, 3)
    e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    directed = set(map(tuple, e.tolist()))
    assert len(directed) == len(e)
    assert all((b, a) in directed for a, b in e.tolist())
    v, f, c, s = scenes.stack(64)
    assert len(f) == 64 and np.all(v[3::3, 2] > v[0::3, 2][:-1] - 1)
    assert np.isclose(np.linalg.det(c[:3, :3]), 1.0)


def _direct_geodesic(radius, m, amplitude, seed):
    """Point-by-point construction of the class-I geodesic sphere: lattice
    points keyed by their (corner, weight) pairs in a dict, numbered in order
    of first use."""
    from hip_peel.scenes import _value_noise, icosahedron

    corners, ico = icosahedron()
    index, points, faces = {}, [], []

    def vertex(*weights):
        key = tuple(sorted((c, w) for c, w in weights if w > 0))
        if key not in index:
            p = sum(w * corners[c] for c, w in key) / m
            index[key] = len(points)
            points.append(p / np.linalg.norm(p))
        return index[key]

    for A, B, C in ico:
        for a in range(m):
            for b in range(m - a):
                faces.append((vertex((A, m - a - b), (B, a), (C, b)), vertex((A, m - a - 1 - b), (B, a + 1), (C, b)),
                              vertex((A, m - a - b - 1), (B, a), (C, b + 1))))
                if a + b < m - 1:
                    faces.append((vertex((A, m - a - 1 - b), (B, a + 1), (C, b)),
                                  vertex((A, m - a - b - 2), (B, a + 1), (C, b + 1)),
                                  vertex((A, m - a - b - 1), (B, a), (C, b + 1))))
    dirs = np.array(points)
    vertices = dirs * (radius + amplitude * _value_noise(dirs * radius, seed))[:, None]
    faces = np.array(faces, dtype=np.int32)
    a, b, c = vertices[faces[:, 0]], vertices[faces[:, 1]], vertices[faces[:, 2]]
    inward = np.einsum("ij,ij->i", np.cross(b - a, c - a), a + b + c) < 0
    faces[inward] = faces[inward][:, [0, 2, 1]]
    return vertices, faces


@pytest.mark.parametrize("m", [1, 2, 3, 5, 8, 13, 53])
def test_geodesic_sphere_matches_direct_construction(m):
    from hip_peel import scenes

    v, f = scenes.geodesic_sphere(1.7, m, amplitude=0.2, seed=m)
    dv, df = _direct_geodesic(1.7, m, 0.2, m)
    np.testing.assert_array_equal(v, dv)
    np.testing.assert_array_equal(f, df)
    assert len(f) == 20 * m * m and len(v) == 10 * m * m + 2
