"""The CPU FP64 restatement of the event route (oracle) against the exact
per-point ranks, on the reference's fixtures (CPU only). This pins the
restatement that the device route is later compared to bit for bit."""

import numpy as np
import pytest

import oracle
from peel_checks import (accepted_fixtures, check_extraction, fixture_path, golden_layers,
                         grid_points, interior_points, source_areas, polygon_area)
from test_gpu_fixtures import CASES, expected_layers


class CpuView:
    """Adapter so check_extraction-like checks read a CpuResult."""


def check_cpu(result, host, expected):
    vertices, faces, camera, settings = host
    rings, srcs, lays = result.rings, result.sources, result.layers
    keep = [i for i, r in enumerate(rings) if len(r)]
    rings = [rings[i] for i in keep]
    srcs, lays = srcs[keep], lays[keep]
    assert result.counts["error"] == 0
    assert result.counts["layers"] == expected
    pts, owner = interior_points(rings)
    ranks = oracle.exact_ranks(vertices, faces, camera, settings, pts)
    for k in range(len(pts)):
        if ranks.point_flags[k]:
            continue
        a, b = ranks.offsets[k], ranks.offsets[k + 1]
        entries = dict(zip(ranks.sources[a:b].tolist(), ranks.ranks[a:b].tolist()))
        assert entries.get(int(srcs[owner[k]])) == int(lays[owner[k]]), k
    gp = grid_points(32)
    audit = oracle.audit(rings, srcs, lays, gp, oracle.exact_ranks(vertices, faces, camera, settings, gp))
    assert audit["mismatches"] == 0, audit
    areas = source_areas(vertices, faces, camera, settings)
    got = np.zeros(len(faces))
    for r, s in zip(rings, srcs):
        got[s] += polygon_area(r)
    assert np.abs(got - areas).max(initial=0) <= 1e-9


@pytest.mark.parametrize("name", CASES)
def test_cpu_route_matches_exact_oracle(name):
    V, F, C, S, _ = oracle.load_scene(fixture_path(name))
    result = oracle.cpu_event_peel(V, F, C, S, threads=2)
    check_cpu(result, (V, F, C, S), expected_layers(name))


def test_cpu_route_stack_closed_form():
    from test_oracle import stack_scene

    V, F, C, S = stack_scene(300)
    r = oracle.cpu_event_peel(V, F, C, S, threads=2)
    assert r.counts["polygons"] == 300 and r.counts["splits"] == 0
    assert r.counts["unique_contours"] == 3 and r.counts["ray_groups"] == 1
    assert r.counts["ray_hits"] == 300 and r.counts["layers"] == 300
    np.testing.assert_array_equal(np.sort(r.layers), np.arange(300))
    np.testing.assert_array_equal(r.layers, r.sources)


def nested_stack(n, slanted=False, extra=None):
    """test_events.py:155-170: shrinking, shifted footprints of a stack."""
    base = np.array([[-0.75, -0.75, 0], [0.75, -0.75, 0], [0, 0.75, 0]], dtype=np.float64)
    xyz = np.stack([np.roll(base, i % 3, axis=0) for i in range(n)])
    xyz[:, :, 2] = np.arange(1, n + 1)[:, None]
    t = np.arange(n) / n
    xyz[:, :, :2] *= (1 - 0.1 * t)[:, None, None]
    xyz[:, :, 0] += 0.01 * t[:, None]
    if slanted:
        xyz[:, :, 2] = 10 + 0.001 * np.arange(n)[:, None] + 10 * xyz[:, :, 0]
    if extra is not None:
        xyz = np.concatenate([xyz, extra])
    m = len(xyz)
    return (xyz.reshape(-1, 3), np.arange(3 * m, dtype=np.int32).reshape(-1, 3), np.eye(4),
            np.array([-1.0, -1.0, 1.0, 1.0, 0.0, 0.01]))


@pytest.mark.parametrize("slanted", [False, True])
def test_cpu_route_certified_beam(slanted):
    """The reference's expectations for nested footprints (test_events.py:155-190):
    one certified cell, one ray, ranks = source order; without sharing, one ray
    per patch and the same ranks."""
    n = 128
    host = nested_stack(n, slanted)
    r = oracle.cpu_event_peel(*host, threads=2)
    assert r.counts["splits"] == 0 and r.counts["polygons"] == n
    assert r.counts["certified_beams"] == 1 and r.counts["beam_attempts"] == 1
    assert r.counts["ray_groups"] == 1 and r.counts["ray_hits"] == n
    np.testing.assert_array_equal(r.layers, r.sources)
    control = oracle.cpu_event_peel(*host, threads=2, share_regions=False)
    assert control.counts["ray_groups"] == n and control.counts["ray_hits"] == n * n
    assert control.counts["certified_beams"] == 0
    np.testing.assert_array_equal(control.layers, control.sources)


def test_cpu_route_beam_rejected():
    """A far source whose edge crosses the samples' rectangle blocks sharing
    (beams.py:88-97), and ranks stay exact."""
    extra = np.array([[[0.004, -0.26, 50.0], [0.3, -0.26, 50.0], [0.004, 0.1, 50.0]]])
    host = nested_stack(8, extra=extra)
    r = oracle.cpu_event_peel(*host, threads=2)
    assert r.counts["beam_attempts"] >= 1 and r.counts["certified_beams"] == 0
    check_cpu(r, host, r.counts["layers"])
