"""Crossing scenes through the exact resolver, then the device routes
(tests/gpu/test_events.py:29-42 `test_resolved_oracle`): the reference's
crossing fixtures and the 256 random triangles are split exactly
(hip_peel.resolve), peeled on the device, and checked against the exact
oracle on the resolved scene and on the original crossing scene, whose exact
pointwise depth order resolution does not change."""

import json

import numpy as np
import pytest

import oracle
from peel_checks import check_extraction, fixture_path, golden_layers, interior_points
from test_gpu_parity import compare

pytestmark = pytest.mark.gpu

CROSSING = ["crossing", "perspective_crossing"] + [f"random_{i:02}" for i in range(12)]


def original_rank_mismatches(rings, srcs, lays, parents, original, sample=None, seed=0):
    """Exact ranks on the unresolved scene at fragment interior points: the
    fragment's layer must be the rank of its parent source there."""
    idx = np.arange(len(rings))
    if sample is not None and sample < len(rings):
        idx = np.random.default_rng(seed).choice(len(rings), sample, replace=False)
    pts, owner = interior_points([rings[i] for i in idx])
    owner = idx[owner]
    r = oracle.exact_ranks(*original, pts)
    bad, checked = [], 0
    for k in range(len(pts)):
        if r.point_flags[k]:
            continue
        a, b = r.offsets[k], r.offsets[k + 1]
        entries = dict(zip(r.sources[a:b].tolist(), r.ranks[a:b].tolist()))
        checked += 1
        f = owner[k]
        if entries.get(int(parents[srcs[f]])) != int(lays[f]):
            bad.append((int(f), int(parents[srcs[f]]), int(lays[f]), entries.get(int(parents[srcs[f]]))))
    return bad, checked


@pytest.mark.parametrize("name", CROSSING)
def test_resolved_crossing_matches_oracle(name):
    from hip_peel import DevicePeeler
    from hip_peel.io import scene_arrays, upload
    from hip_peel.resolve import resolve_scene

    print(f"[resolved] {name}", flush=True)
    source = json.loads(fixture_path(name).read_text())
    data = resolve_scene(source)
    v, f, c, s, labels = scene_arrays(data)
    host = (v, f, c, s)
    golden = golden_layers()
    layers = golden.get(f"{name}-preprocessed", golden[name])
    # event route: bit parity with the CPU restatement, then the exact oracle
    p, info = compare(host)
    check_extraction(p, host, expected_layers=layers)
    # the unresolved scene's exact ranks at every fragment
    parents = np.array([t["original_source_id"] for t in data["triangles"]])
    ids = {t.get("id", i): i for i, t in enumerate(source["triangles"])}
    parents = np.array([ids[x] for x in parents])
    original = oracle.load_scene(source)[:4]
    polys = p.output.numpy(info["polygons"])
    rings = [x["v"][: x["n"]] for x in polys]
    srcs = p.output_sources[: info["polygons"]].cpu().numpy()
    lays = p.output_layers[: info["polygons"]].cpu().numpy()
    bad, checked = original_rank_mismatches(rings, srcs, lays, parents, original)
    assert checked > 0 and not bad, bad[:5]
    # per-layer clipping route on the same resolved scene
    d = DevicePeeler(*upload(*host))
    d.run()
    check_extraction(d, host, expected_layers=layers)


def peel_with_rerun(host, scales=(1, 4, 16)):
    """EventPeeler with the default capacities, rerun with scaled pools while
    the status reports a pool overflow (the reference's rerun protocol)."""
    from dataclasses import replace

    from hip_peel import EventCapacities, EventPeeler
    from hip_peel.io import upload

    arrays = upload(*host[:4])
    for scale in scales:
        cap = EventCapacities.for_mesh(host[1].shape[0])
        cap = replace(cap, **{k: getattr(cap, k) * scale for k in ("frontier", "regions", "hits", "vertices", "triangles")})
        p = EventPeeler(*arrays, capacities=cap)
        p.run()
        status, _ = p._read()
        if status[9] == 0:
            return p, scale
        del p
    raise AssertionError("pools still overflow at the largest scale")


@pytest.mark.parametrize("n", [16, 48])
def test_random_soup_resolved(n):
    """The random-triangles generator (vertices i.i.d. uniform in [0, 1]^3,
    orthographic along -z) at n triangles, resolved exactly, peeled on the
    device: bit parity with the CPU restatement, the exact oracle on the
    resolved mesh, and the exact ranks of the unresolved soup at every
    fragment."""
    from hip_peel import scenes

    print(f"[resolved] random soup {n}", flush=True)
    v, f, c, s, parents = scenes.random_triangles_resolved(n, 0)
    host = (v, f, c, s)
    p, scale = peel_with_rerun(host)
    info = p.read_status()
    P, T = info["polygons"], info["triangles"]
    cpu = oracle.cpu_event_peel(*host, threads=16, cap_poly=8 * scale * len(f), cap_tri=32 * scale * len(f))
    assert cpu.counts["error"] == 0
    for key in ("polygons", "triangles", "layers", "splits", "unique_contours"):
        assert info[key] == cpu.counts[key], key
    polys = p.output.numpy(P)
    rings = [x["v"][: x["n"]] for x in polys]
    srcs = p.output_sources[:P].cpu().numpy()
    lays = p.output_layers[:P].cpu().numpy()
    from test_gpu_parity import canonical_patches

    assert canonical_patches(rings, srcs, lays) == canonical_patches(cpu.rings, cpu.sources, cpu.layers)
    report = check_extraction(p, host, grid=48)
    assert report["fragment_rank_mismatches"] == 0
    original = scenes.random_triangles(n, 0)
    bad, checked = original_rank_mismatches(rings, srcs, lays, parents, original)
    assert checked > 0 and not bad, bad[:5]
    assert T == cpu.counts["triangles"]


def test_random_256_resolution_full_size():
    """The 256-triangle random scene resolved at full size in worker
    processes on the host: 272,806 pieces, each on its parent's plane and
    inside it, and the pieces' areas summing to their parent's (binary64
    rounding of the exact pieces only). Its whole window holds about 10^9
    event fragments; test_gpu_tiles peels it inside a window at its densest."""
    from hip_peel import scenes

    print("[resolved] random 256 resolution", flush=True)
    v0, f0, _, _ = scenes.random_triangles(256, 0)
    v, f, _, _, parents = scenes.random_triangles_resolved(256, 0, workers=16)
    assert f.shape == (272806, 3)
    tri0 = v0[f0]
    tri = v[f]
    par = tri0[parents]
    n = np.cross(par[:, 1] - par[:, 0], par[:, 2] - par[:, 0])
    scale = np.linalg.norm(n, axis=1)
    # vertices on the parent's plane (binary64 rounding of exact points)
    dist = np.abs(np.einsum("pk,pvk->pv", n / scale[:, None], tri - par[:, :1]))
    assert dist.max() < 1e-14
    # pieces keep the parent's winding and tile its area
    a = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    assert (np.einsum("pk,pk->p", a, n) > -1e-24).all()
    sums = np.zeros((256, 3))
    np.add.at(sums, parents, a)
    full = np.cross(tri0[:, 1] - tri0[:, 0], tri0[:, 2] - tri0[:, 0])
    np.testing.assert_allclose(sums, full, rtol=0, atol=1e-12)
