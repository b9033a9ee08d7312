"""Learned pool sizes, CPU side: the binary64 network oracle pinned by
central finite differences and by closed forms; the mesh statistics against
meshes whose contour counts are known; margins, initialisation and the model
file."""

import numpy as np
import pytest

from oracle import pool_mlp


def random_problem(seed, inputs=12, outputs=5, n=37):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, pool_mlp.param_count(inputs, outputs))
    x = rng.normal(size=(n, inputs))
    t = rng.normal(size=(n, outputs))
    return p, x, t


@pytest.mark.parametrize("under", [1.0, 4.0])
def test_oracle_gradient_matches_central_differences(under):
    p, x, t = random_problem(0)
    value, grad = pool_mlp.gradient(p, x, t, under)
    assert value == pytest.approx(pool_mlp.loss(p, x, t, under), rel=1e-12)
    rng = np.random.default_rng(1)
    for k in rng.choice(len(p), 60, replace=False):
        e = np.zeros_like(p)
        e[k] = 1e-6
        fd = (pool_mlp.loss(p + e, x, t, under) - pool_mlp.loss(p - e, x, t, under)) / 2e-6
        assert grad[k] == pytest.approx(fd, rel=1e-5, abs=1e-9), k


def test_oracle_closed_forms():
    inputs, outputs = 3, 2
    p = np.zeros(pool_mlp.param_count(inputs, outputs))
    # zero weights: the output is b3, the loss a weighted mean of squares
    p[-outputs:] = [0.5, -1.0]
    x = np.ones((4, inputs))
    t = np.zeros((4, outputs))
    assert np.allclose(pool_mlp.forward(p, x, outputs), [[0.5, -1.0]] * 4)
    assert pool_mlp.loss(p, x, t, under=3.0) == pytest.approx((0.25 + 3.0 * 1.0) / 2)
    value, grad = pool_mlp.gradient(p, x, t, under=3.0)
    # only b3 receives gradient: 2 w r / O averaged over samples
    assert np.allclose(grad[-outputs:], [2 * 0.5 / 2, 2 * 3.0 * -1.0 / 2])
    assert np.count_nonzero(grad[:-outputs]) == 0
    # the first Adam step moves each parameter by lr * g / (|g| + eps)
    g = np.array([0.3, -2.0, 0.0])
    q, m, v = pool_mlp.adam(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, lr=0.1)
    assert np.allclose(q, -0.1 * g / (np.abs(g) + 1e-8))


def test_margin_and_init():
    from hip_peel import pools

    target = np.log1p(np.array([[10.0, 100.0], [20.0, 50.0], [30.0, 80.0]]))
    pred = target - np.array([[0.1, -0.2], [0.3, 0.0], [-0.1, 0.1]])
    m = pools.calibrate_margin(target, pred, quantile=1.0, slack=0.0)
    assert np.allclose(m, [0.3, 0.1])
    assert np.all(pools.calibrate_margin(target, target + 1.0) >= 0)
    p = pools.init_params(12, 5, seed=3)
    assert p.dtype == np.float32 and len(p) == pool_mlp.param_count(12, 5)


def test_model_file_round_trip(tmp_path):
    from hip_peel import pools

    p = pools.init_params(12, 5)
    m = pools.PoolModel(p, np.zeros(12), np.ones(12), np.zeros(5), np.ones(5), np.full(5, 0.2),
                        {"runs": 10.0})
    path = tmp_path / "m.npz"
    m.save(path)
    q = pools.PoolModel.load(path)
    assert np.array_equal(q.params, p) and np.allclose(q.margin, 0.2) and q.meta["runs"] == 10.0


def single(z=1.0):
    v = np.array([[0, 0, z], [1, 0, z], [0, 1, z]], dtype=float)
    return v, np.array([[0, 1, 2]])


def test_features_contours():
    from hip_peel import pools

    cam = np.eye(4)
    win = np.array([-1.0, -1.0, 2.0, 2.0, 0.0, 0.01])
    names = pools.FEATURES
    v, f = single()
    x = pools.mesh_features(v, f, cam, win)
    assert x[names.index("log_contours")] == pytest.approx(np.log1p(3))
    assert x[names.index("log_faces")] == 0.0
    assert x[names.index("log_total_area")] == pytest.approx(np.log1p(0.5 / 9))
    # two same-facing triangles sharing an edge: 4 contour slots, 2 internal
    v = np.array([[0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=float)
    f = np.array([[0, 1, 2], [0, 2, 3]])
    x = pools.mesh_features(v, f, cam, win)
    assert x[names.index("log_contours")] == pytest.approx(np.log1p(4))
    # the same pair folded so that one faces away: every slot is a contour
    f2 = np.array([[0, 1, 2], [0, 3, 2]])
    x = pools.mesh_features(v, f2, cam, win)
    assert x[names.index("log_contours")] == pytest.approx(np.log1p(6))
    assert x[names.index("back_fraction")] == pytest.approx(0.5)
    # a closed octahedron seen along y: the four faces around (0, 1, 0) face
    # the camera, the four around (0, -1, 0) face away, and the four rim
    # edges between them are contours (8 slots); the other edges are internal
    v = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], dtype=float)
    f = np.array([[0, 2, 4], [2, 1, 4], [1, 3, 4], [3, 0, 4], [2, 0, 5], [1, 2, 5], [3, 1, 5], [0, 3, 5]])
    rot = np.eye(4)
    rot[:3, :3] = np.array([[1, 0, 0], [0, 0, -1], [0, 1, 0]])  # look along y: equator x-z edges are rims
    rot[2, 3] = 5.0
    x = pools.mesh_features(v, f, rot, np.array([-2.0, -2.0, 2.0, 2.0, 0.0, 0.01]))
    assert x[names.index("log_contours")] == pytest.approx(np.log1p(8))
    # stacked soup: no shared vertices, every slot is a contour
    from hip_peel import scenes

    v, f, c, s = scenes.stack(16)
    x = pools.mesh_features(v, f, c, s)
    assert x[names.index("log_contours")] == pytest.approx(np.log1p(48))


def test_generated_meshes_cover_every_kind():
    from hip_peel import pools

    kinds = [k for k, _ in pools.generated_meshes(40, seed=2)]
    assert set(kinds) == {"icosphere", "knot", "stack", "soup"}


def test_depth_complexity_closed_form():
    """Two stacked squares exactly filling the window: every cell covered at
    depth 2; one triangle over half the window's cells touches all cells of
    its bounding box at depth 1/2."""
    from hip_peel import pools

    names = pools.FEATURES
    v = np.array([[0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=float)
    f = np.array([[0, 1, 2], [0, 2, 3]])
    v2 = np.concatenate([v, v + [0, 0, 1]])
    f2 = np.concatenate([f, f + 4])
    x = pools.mesh_features(v2, f2, np.eye(4), np.array([0.0, 0.0, 1.0, 1.0, 0.0, 0.01]))
    assert x[names.index("covered_fraction")] == pytest.approx(1.0)
    assert x[names.index("log_depth_mean")] == pytest.approx(np.log1p(2.0))
    assert x[names.index("log_depth_max")] == pytest.approx(np.log1p(2.0))
    x = pools.mesh_features(v, f[:1], np.eye(4), np.array([0.0, 0.0, 1.0, 1.0, 0.0, 0.01]))
    assert x[names.index("log_depth_mean")] == pytest.approx(np.log1p(0.5))


def test_closed_mesh_detection():
    from hip_peel import pools, scenes

    v, f = scenes.geodesic_sphere(1.0, 4)
    assert pools.closed_mesh(v, f)
    # two separate closed surfaces are still closed
    assert pools.closed_mesh(np.concatenate([v, v + 5.0]), np.concatenate([f, f + len(v)]))
    assert not pools.closed_mesh(v, f[1:])  # a hole
    flipped = f.copy()
    flipped[0] = flipped[0][::-1]
    assert not pools.closed_mesh(v, flipped)  # inconsistent orientation
    assert not pools.closed_mesh(*scenes.stack(8)[:2])  # unwelded soup
    # an unwelded copy of a closed mesh welds back to it
    assert pools.closed_mesh(v[f].reshape(-1, 3), np.arange(3 * len(f)).reshape(-1, 3))


def test_closed_mesh_floor_and_capacities():
    """The floor is the per-face use quantile of the large closed-surface
    runs, and only closed meshes of at least FLOOR_MIN_FACES faces get it."""
    from hip_peel import pools

    usage = np.array([[10, 20, 30, 40, 50], [100, 100, 100, 100, 100], [1, 1, 1, 1, 1], [9e5] * 5], float)
    faces = np.array([5000, 10000, 100, 6000])
    kinds = np.array(["icosphere", "knot", "icosphere", "soup"])
    floor = pools.closed_mesh_floor(usage, faces, kinds, quantile=1.0)
    # per-face use [0.002 .. 0.01] and 0.01: soups and meshes below 4,096 faces are excluded
    np.testing.assert_allclose(floor, [0.01] * 5)
    m = pools.load_default()
    assert (m.floor > 0).all()


def test_nested_shells_floor_covers_recorded_use():
    """The committed model's floor admits the 717,740-face nested shells, a
    closed mesh none of the training generators produce: their recorded use
    per face (frontier 228,226 slots, 827,321 polygons, 1,353,120 ray hits,
    2,734,657 triangles) lies below it."""
    from hip_peel import pools

    m = pools.load_default()
    used = np.array([228226, 827321, 1353120, 2734657]) / 717740
    np.testing.assert_array_less(used, m.floor[[0, 1, 2, 4]])
