"""DevicePeeler (whole-face ranks, then per-layer clipping of the residual
pieces) against the exact oracle, following the reference's DevicePeeler tests
(tests/gpu/test_peeler.py) with exact-rank audits in place of shapely unions."""

from collections import defaultdict
from dataclasses import replace

import numpy as np
import pytest

import oracle
from peel_checks import check_extraction, fixture_path, fragments, polygon_area
from test_gpu_fixtures import CASES, expected_layers

pytestmark = pytest.mark.gpu


def scene(name):
    from hip_peel.io import load_scene

    arrays, labels, data = load_scene(fixture_path(name))
    return arrays, labels, data, oracle.load_scene(data)[:4]


def host_of(arrays):
    v, f, c, s = (a.cpu().numpy() for a in arrays)
    return v.reshape(-1, 3), f.reshape(-1, 3), c.reshape(4, 4), s.reshape(6)


def areas_by_key(peeler):
    rings, srcs, lays = fragments(peeler)
    out = defaultdict(float)
    for r, s, l in zip(rings, srcs, lays):
        out[(int(l), int(s))] += polygon_area(r)
    return out


@pytest.mark.parametrize("name", CASES)
def test_device_fixture_matches_oracle(name):
    """test_peeler.py:93-136: regions, triangles, reprojection, replay."""
    from hip_peel import DevicePeeler

    print(f"[device] {name}", flush=True)
    arrays, labels, data, host = scene(name)
    p = DevicePeeler(*arrays)
    p.run()
    check_extraction(p, host, expected_layers=expected_layers(name))
    info = p.read_status()
    stats = p.stats.cpu().numpy()
    assert stats[info["peeling_passes"]:].sum() == 0
    assert info["peak_residual_patches"] >= stats[0, 0]
    p.run()
    check_extraction(p, host, expected_layers=expected_layers(name))


@pytest.mark.parametrize("limit", [1, 2])
def test_device_prefix(limit):
    """test_peeler.py:205-213."""
    from hip_peel import DevicePeeler

    arrays = scene("holes")[0]
    p = DevicePeeler(*arrays, layers=limit)
    p.run()
    info = p.read_status()
    assert info["layers"] == min(limit, 2)
    full = DevicePeeler(*arrays)
    full.run()
    want = {k: v for k, v in areas_by_key(full).items() if k[0] < limit}
    got = areas_by_key(p)
    assert set(got) == set(want)
    for k in want:
        assert abs(got[k] - want[k]) < 1e-12, k


@pytest.mark.parametrize(
    "field,bit,name",
    [("residual", 4, "holes"), ("polygons", 8, "holes"), ("vertices", 16, "holes"), ("triangles", 32, "holes"),
     ("layers", 64, "holes"), ("pieces_per_worker", 2, "holes"), ("edge_vertices", 128, "perspective_torus")],
)
def test_device_overflow(field, bit, name):
    """test_peeler.py:216-238: each pool reports its own bit, on every replay.
    edge_vertices bounds distinct vertices per edge here (the reference counts
    every copy), so its case needs a scene with two distinct vertices on one
    edge; `holes` has at most one."""
    from hip_peel import Capacities, DevicePeeler

    arrays = scene(name)[0]
    cap = replace(Capacities.for_mesh(arrays[1].shape[0]), **{field: 1})
    p = DevicePeeler(*arrays, capacities=cap)
    for _ in range(2):
        p.run()
        with pytest.raises(RuntimeError, match="GPU extraction failed"):
            p.read_status()
        assert int(p.status[9].item()) & bit


@pytest.mark.parametrize("invalid", ["index", "camera"])
def test_device_invalid_input(invalid):
    """test_peeler.py:241-252."""
    import torch

    from hip_peel import DevicePeeler

    p = DevicePeeler(*scene("single")[0])
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
def test_device_mesh_has_no_t_junctions(name):
    """test_peeler.py:255-280."""
    from hip_peel import DevicePeeler

    p = DevicePeeler(*scene(name)[0])
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


def test_device_camera_update_and_external_capture():
    """test_peeler.py:155-202 and 290-302: a device-side camera change is
    honoured by replays of an externally captured graph."""
    import torch

    from hip_peel import DevicePeeler

    arrays, labels, data, host = scene("partial")
    p = DevicePeeler(*arrays)
    p.warmup()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=p.stream):
        p.enqueue()
    g.replay()
    check_extraction(p, host, expected_layers=2)
    before = areas_by_key(p)
    with torch.cuda.stream(p.stream):
        p.camera[0, 0, 3] += 0.5
    g.replay()
    after = areas_by_key(p)
    assert before != after
    check_extraction(p, host_of(arrays))


@pytest.mark.parametrize("grid", [(32, 16), (64, 32)])
def test_device_route_matches_event_route(grid):
    """Both routes partition the same labelled rank regions: per (layer,
    source) areas agree, and both pass the exact audit."""
    from hip_peel import DevicePeeler, EventPeeler, scenes
    from hip_peel.io import upload

    host = scenes.torus_knot(segments=grid[0], sides=grid[1])
    d = DevicePeeler(*upload(*host[:4]))
    d.run()
    e = EventPeeler(*upload(*host[:4]))
    e.run()
    di, ei = d.read_status(), e.read_status()
    assert di["layers"] == ei["layers"]
    da, ea = areas_by_key(d), areas_by_key(e)
    assert set(da) == set(ea)
    for k in ea:
        assert abs(da[k] - ea[k]) <= 1e-10, (k, da[k], ea[k])
    # the per-layer cuts leave near-collinear vertex runs on this mesh; one
    # fan triangle of ~1e-15 x 1e-3 has a noise-level negative area (-3.8e-20)
    check_extraction(d, host[:4], grid=32, triangles=grid[0] <= 32, orientation_tol=1e-18)


def layered_records(rings, srcs, lays):
    """Multiset of (source, layer, ring bytes) of the non-empty fragments."""
    return sorted((int(s), int(l), np.ascontiguousarray(r).tobytes()) for r, s, l in zip(rings, srcs, lays) if len(r))


def _device_vs_cpu(host, cap=None):
    from hip_peel import DevicePeeler
    from hip_peel.io import upload

    p = DevicePeeler(*upload(*host[:4]), capacities=cap)
    p.run()
    info = p.read_status()
    c = p.cap
    cpu = oracle.cpu_layered_peel(*host[:4], layer_capacity=c.layers, pieces=c.pieces_per_worker,
                                  residual=c.residual, cap_poly=c.polygons)
    assert cpu.counts["error"] == 0
    assert (info["layers"], info["peeling_passes"], info["depth_ties"]) == (
        cpu.counts["layers"], cpu.counts["passes"], cpu.counts["depth_ties"])
    assert info["peak_residual_patches"] == cpu.counts["peak_residual"]
    stats = p.stats.cpu().numpy()[: info["peeling_passes"]]
    np.testing.assert_array_equal(stats, cpu.stats)
    rings, srcs, lays = fragments(p, info)
    assert layered_records(rings, srcs, lays) == layered_records(cpu.rings, cpu.sources, cpu.layers)
    return p, info


@pytest.mark.parametrize("name", CASES)
def test_device_matches_cpu_restatement(name):
    """Bit parity with the CPU restatement of the layered route: every piece
    is cut in the same candidate order with the same binary64 operations, so
    the fragment multisets are identical, as are the per-layer statistics."""
    print(f"[device-cpu] {name}", flush=True)
    _device_vs_cpu(oracle.load_scene(fixture_path(name))[:4])


def test_device_matches_cpu_restatement_torus_knot():
    from hip_peel import scenes

    _device_vs_cpu(scenes.torus_knot(segments=32, sides=16))


def test_device_resolved_soup_exact_ranks():
    """The resolved 48-triangle random soup (3,024 pieces, 16 layers): bit
    parity with the CPU restatement, and every fragment on its exact rank with
    no allowance (the strict-convexity rule of the layered route's clipping;
    without it some candidate orders left ~3e-10-wide strips one layer off)."""
    from hip_peel import Capacities, scenes

    host = scenes.random_triangles_resolved(48, 0, workers=8)[:4]
    F = len(host[1])
    cap = replace(Capacities.for_mesh(F), residual=64 * F, polygons=64 * F, vertices=512 * F, triangles=512 * F,
                  pieces_per_worker=4096)
    p, info = _device_vs_cpu(host, cap)
    assert info["layers"] == 16
    # fan triangles over near-collinear vertex runs carry noise-level doubled
    # areas (~-2e-20), as on the torus knot above
    report = check_extraction(p, host, grid=48, orientation_tol=1e-18)
    print(f"[device-cpu] soup 48: {report}", flush=True)
