"""The CPU FP64 restatement of the layered (DevicePeeler) route against the
exact per-point ranks (CPU only). It pins the restatement that the device's
layered route is compared to bit for bit, and the strict-convexity rule that
keeps its clipping exact."""

from fractions import Fraction

import numpy as np
import pytest

import oracle
from peel_checks import fixture_path, interior_points
from test_cpu_route import check_cpu
from test_gpu_fixtures import CASES, expected_layers


@pytest.mark.parametrize("name", CASES)
def test_cpu_layered_matches_exact_oracle(name):
    V, F, C, S, _ = oracle.load_scene(fixture_path(name))
    result = oracle.cpu_layered_peel(V, F, C, S)
    check_cpu(result, (V, F, C, S), expected_layers(name))


def rank_mismatches(host, result):
    """Fragments whose interior points do not carry their source's exact rank."""
    pts, owner = interior_points(result.rings)
    r = oracle.exact_ranks(*host, pts)
    bad = set()
    for k in range(len(pts)):
        if r.point_flags[k]:
            continue
        a, b = r.offsets[k], r.offsets[k + 1]
        entries = dict(zip(r.sources[a:b].tolist(), r.ranks[a:b].tolist()))
        if entries.get(int(result.sources[owner[k]])) != int(result.layers[owner[k]]):
            bad.add(int(owner[k]))
    return sorted(bad)


@pytest.fixture(scope="module")
def soup32():
    from hip_peel import scenes

    return scenes.random_triangles_resolved(32, 0, workers=8)[:4]


def _peel(host, **kw):
    F = len(host[1])
    return oracle.cpu_layered_peel(*host, pieces=8192, residual=64 * F, cap_poly=64 * F, **kw)


@pytest.mark.parametrize("order_seed", [0, 6, 11])
def test_cpu_layered_resolved_soup_exact(soup32, order_seed):
    """The resolved 32-triangle soup (1,292 pieces, ~41 thousand fragments):
    every fragment on its exact rank, in the deterministic candidate order and
    in shuffled orders."""
    result = _peel(soup32, order_seed=order_seed)
    assert result.counts["error"] == 0
    assert not rank_mismatches(soup32, result)


def test_cpu_layered_without_convexity_misranks(soup32):
    """Planted in the oracle: without the strict-convexity rule, the shuffled
    order 6 clips by the rotated line of a ~1e-12 constructed edge and leaves a
    fragment one layer off; the exact check catches it."""
    result = _peel(soup32, order_seed=6, convexify_rings=False)
    assert result.counts["error"] == 0
    assert rank_mismatches(soup32, result)


def _turns_exact(ring):
    n = len(ring)
    q = [(Fraction(x), Fraction(y)) for x, y in ring]
    out = []
    for j in range(n):
        a, b, c = q[j - 1], q[j], q[(j + 1) % n]
        out.append((b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]))
    return out


def _boundary_distance(p, ring):
    best = np.inf
    for j in range(len(ring)):
        a, b = ring[j], ring[(j + 1) % len(ring)]
        d = b - a
        t = np.clip(np.dot(p - a, d) / np.dot(d, d), 0.0, 1.0)
        best = min(best, float(np.linalg.norm(p - (a + t * d))))
    return best


def test_convexify_rule():
    rng = np.random.default_rng(4)
    for trial in range(200):
        # a random convex polygon
        k = int(rng.integers(3, 12))
        ang = np.sort(rng.random(k)) * 2 * np.pi
        base = np.stack([0.5 + 0.3 * np.cos(ang), 0.5 + 0.2 * np.sin(ang)], axis=1)
        if min(_turns_exact(base)) <= 0:
            continue
        kept = oracle.convexify(base)
        if min(_turns_exact(base)) > 1e-9:
            np.testing.assert_array_equal(kept, base)  # strictly convex rings are untouched
        # insert a vertex within 1e-14 of an edge's interior, on either side
        j = int(rng.integers(k))
        a, b = base[j], base[(j + 1) % k]
        t = rng.uniform(0.2, 0.8)
        d = b - a
        normal = np.array([-d[1], d[0]]) / np.linalg.norm(d)
        extra = a + t * d + normal * rng.uniform(-1e-14, 1e-14)
        ring = np.insert(base, j + 1, extra, axis=0)
        out = oracle.convexify(ring)
        turns = _turns_exact(out)
        assert len(out) >= 3 and min(turns) > 0, trial
        # only vertices within ~1e-12 of the kept boundary are dropped
        assert len(out) >= k
        for p in ring:
            if not any((p == q).all() for q in out):
                assert _boundary_distance(p, out) < 1e-11, (trial, p)
    # a ring that collapses to a needle is emptied
    needle = np.array([[0.1, 0.1], [0.9, 0.9], [0.5, 0.5 + 1e-15]])
    assert len(oracle.convexify(needle)) == 0
