"""Device API behaviour, following the reference's GPU tests
(tests/gpu/test_peeler.py, test_events.py, test_components.py) with exact-rank
audits in place of shapely region unions."""

import json
from dataclasses import replace

import numpy as np
import pytest

import oracle
from peel_checks import check_extraction, fixture_path, fragments

pytestmark = pytest.mark.gpu


def scene(name):
    from hip_peel.io import load_scene

    arrays, labels, data = load_scene(fixture_path(name))
    return arrays, labels, data, oracle.load_scene(data)[:4]


def host_of(arrays):
    v, f, c, s = (a.cpu().numpy() for a in arrays)
    return v.reshape(-1, 3), f.reshape(-1, 3), c.reshape(4, 4), s.reshape(6)


def stack(n):
    """test_events.py:63-71: rotated identical footprints at z = 1..n."""
    import torch

    vertices = np.array([[-0.75, -0.75, 0], [0.75, -0.75, 0], [0, 0.75, 0]], dtype=np.float64)
    mesh = np.stack([np.roll(vertices, i % 3, axis=0) for i in range(n)])
    mesh[:, :, 2] = np.arange(1, n + 1)[:, None]
    return [
        torch.tensor(mesh.reshape(-1, 3), device="cuda"),
        torch.tensor(np.arange(3 * n, dtype=np.int32).reshape(-1, 3), device="cuda"),
        torch.tensor(np.eye(4)[None], device="cuda"),
        torch.tensor([-1.0, -1.0, 1.0, 1.0, 0.0, 0.01], dtype=torch.float64, device="cuda"),
    ]


def test_external_capture_and_changed_input():
    """test_peeler.py:155-202: enqueue inside an external capture on the
    workspace stream; replay after changing inputs in place."""
    import torch

    from hip_peel import EventPeeler

    arrays, labels, data, host = scene("partial")
    p = EventPeeler(*arrays)
    p.warmup()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=p.stream):
        p.enqueue()
    g.replay()
    before = fragments(p)
    check_extraction(p, host, expected_layers=2)
    with torch.cuda.stream(p.stream):
        arrays[0][:, 0] += 0.5
        arrays[0][:, 1] += 0.25
    g.replay()
    changed = host_of(arrays)
    after = fragments(p)
    assert not all(np.array_equal(a, b) for a, b in zip(before[0], after[0]))
    check_extraction(p, changed)
    # another workspace does not disturb this graph
    other = EventPeeler(*scene("perspective_torus")[0])
    other.run()
    other.read_status()
    g.replay()
    check_extraction(p, changed)


@pytest.mark.parametrize("limit", [1, 2])
def test_prefix(limit):
    """test_peeler.py:205-213: prefixes keep the first `limit` layers."""
    from hip_peel import EventPeeler

    arrays, labels, data, host = scene("holes")
    p = EventPeeler(*arrays, layers=limit)
    p.run()
    info = p.read_status()
    rings, srcs, lays = fragments(p, info)
    assert info["layers"] == min(limit, 2)
    assert len(lays) and lays.max() < limit
    full = EventPeeler(*arrays)
    full.run()
    frings, fsrcs, flays = fragments(full)
    keep = flays < limit
    area = lambda rs: sum(0.5 * abs(np.dot(r[:, 0], np.roll(r[:, 1], -1)) - np.dot(r[:, 1], np.roll(r[:, 0], -1))) for r in rs)
    assert abs(area(rings) - area([r for r, k in zip(frings, keep) if k])) < 1e-12


@pytest.mark.parametrize(
    "field,match",
    [("hits", "shared ray hit"), ("rounds", "subdivision round"), ("frontier", "residual pool"),
     ("regions", "output polygon"), ("vertices", "boundary vertex"), ("triangles", "triangle pool")],
)
def test_event_capacity(field, match):
    """test_events.py:91-106 (plus the vertex/triangle pools of test_peeler.py:216-238)."""
    from hip_peel import EventCapacities, EventPeeler

    arrays = scene("holes")[0]
    cap = replace(EventCapacities.for_mesh(arrays[1].shape[0]), **{field: 1})
    p = EventPeeler(*arrays, capacities=cap)
    for _ in range(2):  # replays keep failing the same way
        p.run()
        with pytest.raises(RuntimeError, match=match):
            p.read_status()


@pytest.mark.parametrize("field,bit", [("residual", 4), ("polygons", 8), ("vertices", 16), ("triangles", 32)])
def test_device_peeler_overflow(field, bit):
    """test_peeler.py:216-238 for the pools this route has."""
    from hip_peel import Capacities, DevicePeeler

    arrays = scene("holes")[0]
    cap = replace(Capacities.for_mesh(arrays[1].shape[0]), **{field: 1})
    p = DevicePeeler(*arrays, capacities=cap)
    p.run()
    with pytest.raises(RuntimeError, match="GPU extraction failed"):
        p.read_status()
    assert int(p.status[9].item()) & bit


@pytest.mark.parametrize("invalid", ["index", "camera"])
def test_invalid_device_input(invalid):
    """test_peeler.py:241-252."""
    import torch

    from hip_peel import EventPeeler

    arrays = scene("single")[0]
    p = EventPeeler(*arrays)
    p.run()
    p.read_status()
    with torch.cuda.stream(p.stream):
        if invalid == "index":
            p.faces[0] = torch.tensor([-1, 0, 1000000], dtype=torch.int32, device="cuda")
        else:
            p.camera.zero_()
    p.run()
    with pytest.raises(RuntimeError, match="invalid input"):
        p.read_status()


@pytest.mark.parametrize("name", ["partial", "holes", "closed_object"])
def test_mesh_has_no_t_junctions(name):
    """test_peeler.py:255-280: no triangle edge passes through another
    triangle vertex of the same layer."""
    from hip_peel import EventPeeler

    p = EventPeeler(*scene(name)[0])
    p.run()
    info = p.read_status()
    T = info["triangles"]
    tris = p.triangles_uv[:T].cpu().numpy()
    layers = p.triangle_layers[:T].cpu().numpy()
    for layer in np.unique(layers):
        ts = tris[layers == layer]
        points = np.unique(ts.reshape(-1, 2), axis=0)
        for t in ts:
            for a, b in zip(t, np.roll(t, -1, axis=0)):
                d = b - a
                tval = (points - a) @ d / (d @ d)
                dist = np.abs(d[0] * (points[:, 1] - a[1]) - d[1] * (points[:, 0] - a[0])) / np.linalg.norm(d)
                assert not np.any((tval > 1e-8) & (tval < 1 - 1e-8) & (dist < 1e-11))


def test_device_camera_update():
    """test_peeler.py:290-302: a camera changed on the device is honoured."""
    import torch

    from hip_peel import EventPeeler

    arrays, labels, data, host = scene("partial")
    p = EventPeeler(*arrays)
    p.run()
    before = fragments(p)
    with torch.cuda.stream(p.stream):
        p.camera[0, 0, 3] += 0.5
    p.run()
    after = fragments(p)
    assert not all(np.array_equal(a, b) for a, b in zip(before[0], after[0]))
    check_extraction(p, host_of(arrays))


def test_device_welding_signed_zero():
    """test_peeler.py:305-319: sorting agrees with == on signed zeros."""
    import torch

    from hip_peel import EventPeeler

    vertices = torch.tensor([[0, 2, 1], [-0.0, 1, 1], [0, 1, 1], [-0.0, 2, 1]], dtype=torch.float64, device="cuda")
    p = EventPeeler(
        vertices,
        torch.empty((0, 3), dtype=torch.int32, device="cuda"),
        torch.tensor(np.eye(4)[None], device="cuda"),
        torch.tensor([-5.0, -5.0, 5.0, 5.0, 0.0, 1.0], dtype=torch.float64, device="cuda"),
    )
    p.run()
    p.read_status()
    ids = p.canonical.cpu().numpy()
    assert ids[0] == ids[3] and ids[1] == ids[2] and ids[0] != ids[1]


def test_empty_download():
    """test_peeler.py:322-328."""
    from hip_peel import EventPeeler
    from hip_peel.io import download_mesh

    p = EventPeeler(*scene("empty")[0])
    p.run()
    data = download_mesh(p)
    assert data["world"].shape == (0, 3, 3)


@pytest.mark.parametrize("n", [16, 33, 65, 128, 257, 1024])
def test_shared_stack(n):
    """test_events.py:74-88: closed-form work counts."""
    from hip_peel import EventPeeler

    p = EventPeeler(*stack(n))
    p.run()
    info = p.read_status()
    assert info["polygons"] == n
    assert info["splits"] == 0
    assert info["unique_contours"] == 3
    assert info["ray_groups"] == 1
    assert info["ray_candidate_tests"] == n
    assert info["ray_hits"] == n
    assert info["contour_tests"] <= 3 * n
    np.testing.assert_array_equal(p.output_layers[:n].cpu().numpy(), p.output_sources[:n].cpu().numpy())


def test_foreground_unsplit():
    """test_events.py:126-133."""
    from hip_peel import EventPeeler

    p = EventPeeler(*scene("cover_2")[0])
    p.run()
    info = p.read_status()
    ids = p.output_sources[: info["polygons"]].cpu().numpy()
    ranks = p.output_layers[: info["polygons"]].cpu().numpy()
    assert np.count_nonzero((ids == 0) & (ranks == 0)) == 1


def test_nonmanifold_edge():
    """test_events.py:136-151: a non-manifold junction keeps its contours."""
    from hip_peel import EventPeeler
    from hip_peel.io import scene_arrays, upload

    a, b = [-1, 0, 2], [1, 0, 2]
    data = {
        "window": [-2, -2, 2, 2],
        "camera": {"type": "orthographic"},
        "triangles": [{"id": i, "vertices": [a, b, c]} for i, c in enumerate([[0, 1, 1], [0, -1, 1], [0, 1, 3]])],
    }
    host = scene_arrays(data)
    p = EventPeeler(*upload(*host[:4]))
    p.run()
    check_extraction(p, oracle.load_scene(data)[:4], expected_layers=2)


@pytest.mark.parametrize("slanted", [False, True])
def test_certified_distinct_samples(slanted):
    """test_events.py:155-190: nested footprints have distinct samples but one
    certified common order, so one ray serves all of them."""
    import torch

    from hip_peel import EventCapacities, EventPeeler

    n = 128
    arrays = stack(n)
    xyz = arrays[0].cpu().numpy().reshape(n, 3, 3)
    t = np.arange(n) / n
    xyz[:, :, :2] *= (1 - 0.1 * t)[:, None, None]
    xyz[:, :, 0] += 0.01 * t[:, None]
    if slanted:
        xyz[:, :, 2] = 10 + 0.001 * np.arange(n)[:, None] + 10 * xyz[:, :, 0]
    arrays[0] = torch.tensor(xyz.reshape(-1, 3), device="cuda")
    p = EventPeeler(*arrays)
    p.run()
    info = p.read_status()
    assert info["splits"] == 0
    assert info["polygons"] == n
    assert info["certified_beams"] == 1
    assert info["ray_groups"] == 1
    assert info["ray_candidate_tests"] == n
    assert info["beam_candidate_tests"] == n
    assert info["heavy_beams"] == 1  # 128 candidates pass the 32 handled by one lane
    np.testing.assert_array_equal(p.output_sources[:n].cpu().numpy(), p.output_layers[:n].cpu().numpy())
    baseline = EventPeeler(*arrays, share_regions=False, capacities=replace(EventCapacities.for_mesh(n), hits=n * n))
    baseline.run()
    control = baseline.read_status()
    assert control["ray_groups"] == n
    assert control["ray_candidate_tests"] == n * n
    assert control["certified_beams"] == 0 and control["beam_attempts"] == 0
    np.testing.assert_array_equal(
        baseline.output_sources[:n].cpu().numpy(), baseline.output_layers[:n].cpu().numpy())


def test_beam_rejected_when_a_candidate_straddles():
    """A cell whose rectangle is crossed by a source edge keeps its exact
    groups (beams.py:88-97: every candidate must strictly cover the corners)."""
    import torch

    from hip_peel import EventPeeler

    n = 8
    arrays = stack(n)
    xyz = arrays[0].cpu().numpy().reshape(n, 3, 3)
    t = np.arange(n) / n
    xyz[:, :, :2] *= (1 - 0.1 * t)[:, None, None]
    xyz[:, :, 0] += 0.01 * t[:, None]
    # a small far triangle whose edge passes through the cell of the samples
    extra = np.array([[[0.004, -0.26, 50.0], [0.3, -0.26, 50.0], [0.004, 0.1, 50.0]]])
    xyz = np.concatenate([xyz, extra])
    arrays[0] = torch.tensor(xyz.reshape(-1, 3), device="cuda")
    arrays[1] = torch.tensor(np.arange(3 * (n + 1), dtype=np.int32).reshape(-1, 3), device="cuda")
    p = EventPeeler(*arrays)
    p.run()
    info = p.read_status()
    assert info["beam_attempts"] >= 1
    assert info["certified_beams"] == 0
    check_extraction(p, host_of(arrays))


def test_heavy_disjoint_groups():
    """test_events.py:193-206."""
    import torch

    from hip_peel import EventPeeler

    arrays = stack(130)
    xyz = arrays[0].cpu().numpy().reshape(130, 3, 3)
    xyz[:, :, 0] *= 0.25
    xyz[:, :, 0] += np.where(np.arange(130) % 2, 0.5, -0.5)[:, None]
    arrays[0] = torch.tensor(xyz.reshape(-1, 3), device="cuda")
    p = EventPeeler(*arrays)
    p.run()
    info = p.read_status()
    assert info["ray_groups"] == info["heavy_ray_groups"] == 2
    assert info["ray_hits"] == 130
    np.testing.assert_array_equal(p.output_sources[:130].cpu().numpy() // 2, p.output_layers[:130].cpu().numpy())


def test_replay_changes_sort_bucket():
    """test_events.py:209-238: the same graph serves n and n^2 hits."""
    import torch

    from hip_peel import EventCapacities, EventPeeler

    n = 64
    arrays = stack(n)
    cap = replace(EventCapacities.for_mesh(n), hits=n * n)
    p = EventPeeler(*arrays, capacities=cap, share_regions=False)
    p.run()
    assert p.read_status()["ray_hits"] == n
    xyz = arrays[0].cpu().numpy().reshape(n, 3, 3)
    xyz[:, :, :2] *= (1 - 0.1 * np.arange(n) / n)[:, None, None]
    xyz[:, :, 0] += (0.01 * np.arange(n) / n)[:, None]
    with torch.cuda.stream(p.stream):
        p.vertices.copy_(torch.tensor(xyz.reshape(-1, 3), device="cuda"))
    p.run()
    assert p.read_status()["ray_hits"] == n * n
    np.testing.assert_array_equal(p.output_sources[:n].cpu().numpy(), p.output_layers[:n].cpu().numpy())
    with torch.cuda.stream(p.stream):
        p.vertices.copy_(stack(n)[0])
    p.run()
    assert p.read_status()["ray_hits"] == n


def grid_scene(half_cover=False):
    """test_components.py:13-32: two 4x4 quad grids at z = 1, 2."""
    triangles = []
    for z in (1, 2):
        for y in range(4):
            for x in range(4):
                if half_cover and z == 1 and x >= 2:
                    continue
                a, b, c, d = [x, y, z], [x + 1, y, z], [x + 1, y + 1, z], [x, y + 1, z]
                for vertices in ([a, b, c], [a, c, d]):
                    triangles.append({"id": len(triangles), "vertices": vertices})
    return {"window": [-1, -1, 5, 5], "camera": {"type": "orthographic"}, "triangles": triangles}


def welded(data):
    """The reference uploads an unwelded soup; welding happens on the device.
    Return the same soup (faces index 3k..3k+2)."""
    from hip_peel.io import scene_arrays, upload

    host = scene_arrays(data)
    return upload(*host[:4]), host


@pytest.mark.parametrize("half_cover", [False, True])
def test_mesh_rank_components(half_cover):
    """test_components.py:35-61."""
    import torch

    from hip_peel import EventPeeler

    data = grid_scene(half_cover)
    arrays, host = welded(data)
    p = EventPeeler(*arrays, share_regions=False)
    p.run()
    info = p.read_status()
    assert info["rank_connections"] > 0
    assert info["rank_components"] <= 3
    assert info["ray_groups"] <= 3
    assert info["merged_polygons"] == 0
    assert info["polygons"] == len(host[4])
    check_extraction(p, oracle.load_scene(data)[:4])
    baseline = EventPeeler(*arrays, share_regions=False, propagate_ranks=False, merge_regions=False)
    baseline.run()
    assert baseline.read_status()["ray_groups"] >= 16
    # deform both surfaces and replay the same graphs
    with torch.cuda.stream(p.stream):
        xyz = arrays[0]
        xyz[:, 2] += 0.05 * xyz[:, 0] * xyz[:, 1]
    torch.cuda.synchronize()
    p.run()
    baseline.run()
    deformed = host_of(arrays)
    check_extraction(p, deformed)
    check_extraction(baseline, deformed)


def test_stack_hash_is_bounded():
    """test_components.py:81-93."""
    from hip_peel import EventPeeler

    p = EventPeeler(*stack(4096))
    p.run()
    info = p.read_status()
    assert info["edge_hash_misses"] == 0
    assert info["matched_patch_edges"] == 0
    assert info["rank_components"] == 4096
    assert info["ray_groups"] == 1
    np.testing.assert_array_equal(p.output_layers[:4096].cpu().numpy(), p.output_sources[:4096].cpu().numpy())


@pytest.mark.parametrize("n", [2049, 5000])
def test_huge_group_order(n):
    """One ray group larger than a workgroup's LDS sort (2,048 hits): the
    stack's depths in shuffled order, so the merge passes must order the hits
    by depth for every layer to equal the source's depth rank."""
    import torch

    from hip_peel import EventPeeler

    arrays = stack(n)
    perm = np.random.default_rng(n).permutation(n)
    xyz = arrays[0].cpu().numpy().reshape(n, 3, 3)
    xyz[:, :, 2] = (perm + 1)[:, None]
    arrays[0] = torch.tensor(xyz.reshape(-1, 3), device="cuda")
    p = EventPeeler(*arrays)
    p.run()
    info = p.read_status()
    assert info["polygons"] == n and info["ray_groups"] == 1 and info["layers"] == n
    src = p.output_sources[:n].cpu().numpy()
    np.testing.assert_array_equal(p.output_layers[:n].cpu().numpy(), perm[src])


def test_query_sort_bucket_replay():
    """test_components.py:96-114: one captured graph, changing polygon counts."""
    import torch

    from hip_peel import EventPeeler

    arrays = stack(1025)
    xyz = arrays[0].cpu().numpy()
    p = EventPeeler(*arrays)
    p.capture()
    for count in (1025, 127, 0, 257, 256, 1025):
        changed = xyz.copy().reshape(1025, 3, 3)
        changed[count:, :, 0] += 10.0
        with torch.cuda.stream(p.stream):
            p.vertices.copy_(torch.tensor(changed.reshape(-1, 3), device="cuda"))
        p.run()
        info = p.read_status()
        assert info["polygons"] == count
        assert info["layers"] == count
        assert info["ray_groups"] == int(count > 0)
        np.testing.assert_array_equal(p.output_sources[:count].cpu().numpy(), p.output_layers[:count].cpu().numpy())


@pytest.mark.parametrize("name", ["holes", "partial", "closed_object", "coplanar_adjacent"])
@pytest.mark.parametrize("propagate,merge", [(True, False), (False, True), (True, True), (False, False)])
def test_component_switches(name, propagate, merge):
    """test_components.py:64-78: rank propagation and convex coalescing leave
    the labelled rank regions exact."""
    from hip_peel import EventPeeler

    arrays, labels, data, host = scene(name)
    p = EventPeeler(*arrays, propagate_ranks=propagate, merge_regions=merge)
    p.run()
    check_extraction(p, host)
    assert p.read_status()["edge_hash_misses"] == 0


@pytest.mark.parametrize("propagate", [True, False])
def test_merge_regions_coalesces(propagate):
    """Convex coalescing (components.py:274-395) joins same-source, same-layer
    patches across matched edges: fewer polygons, the same exact regions and
    triangles, and the merged count reported."""
    from hip_peel import EventPeeler, scenes
    from hip_peel.io import upload

    host = scenes.torus_knot(segments=32, sides=16)
    plain = EventPeeler(*upload(*host[:4]), propagate_ranks=propagate)
    plain.run()
    merged = EventPeeler(*upload(*host[:4]), propagate_ranks=propagate, merge_regions=True)
    merged.run()
    a, b = plain.read_status(), merged.read_status()
    assert b["merged_polygons"] > 0
    assert b["polygons"] == a["polygons"] - b["merged_polygons"]
    assert b["layers"] == a["layers"]
    check_extraction(merged, host[:4], grid=32)


def test_event_edge_capacity():
    """Edge noding capacity (kernels.py:1051-1057, status bit 128): a scene
    whose busiest edge holds more distinct vertices than edge_vertices fails
    with the reference's message, on every replay."""
    from hip_peel import EventCapacities, EventPeeler

    arrays = scene("perspective_torus")[0]
    p = EventPeeler(*arrays)
    p.run()
    busiest = p.read_status()["max_edge_vertices"]
    assert busiest >= 2
    cap = replace(EventCapacities.for_mesh(arrays[1].shape[0]), edge_vertices=busiest - 1)
    q = EventPeeler(*arrays, capacities=cap)
    for _ in range(2):
        q.run()
        with pytest.raises(RuntimeError, match="edge noding capacity"):
            q.read_status()
    ok = EventPeeler(*arrays, capacities=replace(cap, edge_vertices=busiest))
    ok.run()
    ok.read_status()
