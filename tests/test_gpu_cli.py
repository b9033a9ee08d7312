"""The command line and its on-disk output (warp_peel/__main__.py:13-77,
warp_peel/io.py:51-72): `python -m hip_peel` on fixtures through both
methods, the NPZ reloaded and checked against the exact oracle."""

import json
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from peel_checks import PACKAGE_DIR, fixture_path, golden_layers, source_areas

pytestmark = pytest.mark.gpu


def _run_cli(scene, out, *extra):
    env = dict(os.environ)
    env["PYTHONPATH"] = os.pathsep.join([str(PACKAGE_DIR), env.get("PYTHONPATH", "")])
    cmd = [sys.executable, "-m", "hip_peel", str(scene), "--out", str(out), *extra]
    print("cli:", " ".join(cmd[2:]), flush=True)
    r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def _check_mesh(mesh, data, name):
    vertices, faces, camera, settings, labels = oracle.load_scene(data)[:5]
    assert set(mesh.files) == {"uv", "world", "sources", "layers"}
    uv, world, sources, layers = mesh["uv"], mesh["world"], mesh["sources"], mesh["layers"]
    assert uv.dtype == np.float64 and uv.shape[1:] == (3, 2)
    assert world.dtype == np.float64 and world.shape[1:] == (3, 3)
    assert sources.dtype == np.int64 and layers.dtype == np.int32
    T = len(uv)
    assert world.shape[0] == sources.shape[0] == layers.shape[0] == T
    # external labels back to face indices
    index = {int(l): i for i, l in enumerate(labels)}
    faces_of = np.array([index[int(s)] for s in sources], dtype=np.int64)
    # zero-based layers, all present
    if T:
        assert layers.min() == 0
        assert set(layers.tolist()) == set(range(int(layers.max()) + 1))
        assert int(layers.max()) + 1 == golden_layers()[f"{name}-mesh"]
    # counter-clockwise, area conservation per source
    a, b = uv[:, 1] - uv[:, 0], uv[:, 2] - uv[:, 0]
    cross = a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]
    assert (cross > 0).all()
    got = np.bincount(faces_of, weights=0.5 * cross, minlength=len(faces))
    np.testing.assert_allclose(got, source_areas(vertices, faces, camera, settings), atol=1e-9, rtol=0)
    # exact rank of the source at every triangle centroid equals its layer
    centroids = uv.mean(axis=1)
    ranks = oracle.exact_ranks(vertices, faces, camera, settings, centroids)
    checked = 0
    for k in range(T):
        if ranks.point_flags[k]:
            continue
        lo, hi = ranks.offsets[k], ranks.offsets[k + 1]
        entry = dict(zip(ranks.sources[lo:hi].tolist(), ranks.ranks[lo:hi].tolist()))
        assert entry.get(int(faces_of[k])) == int(layers[k]), (k, int(faces_of[k]), int(layers[k]))
        checked += 1
    assert checked > 0 or T == 0
    # world corners reproject onto the uv corners and lie on the source plane
    if T:
        cam = world @ camera[:3, :3].T + camera[:3, 3]
        xy = cam[:, :, :2] / (cam[:, :, 2:3] if settings[4] else 1.0)
        np.testing.assert_allclose((xy - settings[:2]) / (settings[2:4] - settings[:2]), uv, atol=2e-9, rtol=0)
        src = vertices[faces[faces_of]]
        n = np.cross(src[:, 1] - src[:, 0], src[:, 2] - src[:, 0])
        n /= np.linalg.norm(n, axis=1)[:, None]
        np.testing.assert_allclose(np.sum((world - src[:, :1]) * n[:, None], axis=2), 0, atol=2e-8, rtol=0)
    return checked


@pytest.mark.parametrize("method", ["events", "peel"])
@pytest.mark.parametrize("name", ["perspective_torus", "cover_2", "holes"])
def test_cli_writes_npz_and_sidecar(tmp_path, method, name):
    scene = fixture_path(name)
    out = tmp_path / f"{name}_{method}.npz"
    status = _run_cli(scene, out, "--method", method)
    sidecar = json.loads(out.with_suffix(".json").read_text())
    assert sidecar["format"] == "warp-peel-approximate-v1"
    assert sidecar["zero_based_layers"] is True and sidecar["exact_geometry"] is False
    assert sidecar["method"] == method and sidecar["scene"] == str(scene)
    assert sidecar["output"] == status
    with np.load(out) as mesh:
        assert len(mesh["uv"]) == status["triangles"]
        _check_mesh(mesh, json.loads(scene.read_text()), name)


def test_cli_pool_scale_and_pieces(tmp_path):
    out = tmp_path / "t.npz"
    status = _run_cli(fixture_path("cover_2"), out, "--method", "peel", "--pieces", "64", "--pool-scale", "2",
                      "--layers", "1")
    assert status["layers"] == 1
    with np.load(out) as mesh:
        assert set(mesh["layers"].tolist()) <= {0}


def test_cli_rejects_pieces_for_events(tmp_path):
    env = dict(os.environ)
    env["PYTHONPATH"] = os.pathsep.join([str(PACKAGE_DIR), env.get("PYTHONPATH", "")])
    r = subprocess.run([sys.executable, "-m", "hip_peel", str(fixture_path("single")), "--out",
                        str(tmp_path / "x.npz"), "--method", "events", "--pieces", "4"],
                       env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 2 and "--pieces" in r.stderr
