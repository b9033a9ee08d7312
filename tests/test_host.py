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
