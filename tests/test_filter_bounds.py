"""The static error bounds of the binary32 filters (csrc/peel_common.hpp
`orientation_tol32`, `candidate_tol32`, `value32` with DEPTH_TOL32), checked on
the CPU by emulating the device's binary32 arithmetic in numpy (round to
nearest, one rounding per operation, no contraction) on adversarial,
near-degenerate inputs: wherever the binary32 value exceeds its bound, its
sign equals the exact sign (rational arithmetic on the binary64 inputs) and
the binary64 evaluation's, which is then nonzero. The same checks with a
bound 64 times smaller fail, so the inputs reach the bound's scale."""

from fractions import Fraction as Fr

import numpy as np

U32 = 2.0**-24
DEPTH_TOL32 = np.float32(16.0) * np.float32(U32)


def round_up32(t):
    """The binary32 value nearest t, rounded up (__double2float_ru)."""
    f = np.float32(t)
    return np.nextafter(f, np.float32(np.inf)) if float(f) < t else f


def orientation_tol32(D, M):
    return round_up32(16.0 * U32 * D * (M + D) + 1.0e-37)


def cross32(a, b, c):
    a, b, c = (np.asarray(v, dtype=np.float32) for v in (a, b, c))
    return (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])


def cross64(a, b, c):
    return (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])


def cross_exact(a, b, c):
    a, b, c = ([Fr(float(x)) for x in v] for v in (a, b, c))
    return (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])


def sign(x):
    return int(x > 0) - int(x < 0)


def near_collinear(rng, count, scale=1.0):
    """Segment a-b and point p within ~1e-12..1e-3 of its line (times scale)."""
    for _ in range(count):
        a = rng.uniform(-1, 1, 2) * scale
        b = a + rng.uniform(-1, 1, 2) * scale * 10.0 ** rng.uniform(-4, 0)
        t = rng.uniform(-0.5, 1.5)
        d = b - a
        normal = np.array([-d[1], d[0]]) / max(np.linalg.norm(d), 1e-300)
        p = a + t * d + normal * scale * 10.0 ** rng.uniform(-12, -3) * rng.choice([-1.0, 1.0])
        yield a, b, p


def orientation_violations(shrink, count=4000, seed=0):
    rng = np.random.default_rng(seed)
    decided = bad = 0
    for scale in (1.0, 1.0 / 16.0, 16.0):
        for a, b, p in near_collinear(rng, count, scale):
            pts = np.stack([a, b, p])
            D = float(max(np.ptp(pts[:, 0]), np.ptp(pts[:, 1])))
            M = float(np.abs(pts).max())
            tol = orientation_tol32(D, M) / np.float32(shrink)
            s32 = cross32(a, b, p)
            if abs(s32) <= tol:
                continue
            decided += 1
            s64 = cross64(a, b, p)
            if not (sign(s32) == sign(cross_exact(a, b, p)) == sign(s64) != 0):
                bad += 1
    return decided, bad


def test_orientation_bound_holds():
    decided, bad = orientation_violations(1)
    assert decided > 2000 and bad == 0, (decided, bad)


def test_orientation_bound_is_not_slack():
    _, bad = orientation_violations(64)
    assert bad > 0


def candidate_tol32(box, a, b):
    x0, x1 = min(box[0], a[0], b[0]), max(box[2], a[0], b[0])
    y0, y1 = min(box[1], a[1], b[1]), max(box[3], a[1], b[1])
    return orientation_tol32(max(x1 - x0, y1 - y0), max(abs(x0), abs(x1), abs(y0), abs(y1)))


def test_candidate_bound_covers_every_ring_vertex():
    """Straddle and front_segment signs of every vertex and edge of a ring
    against a segment, under the bound built from the ring's box and the
    segment (the subdivision's per-candidate bound)."""
    rng = np.random.default_rng(3)
    decided = 0
    for _ in range(1500):
        c = rng.uniform(0, 1, 2)
        ang = np.sort(rng.uniform(0, 2 * np.pi, int(rng.integers(3, 9))))
        r = 10.0 ** rng.uniform(-6, -1)
        ring = c + r * np.stack([np.cos(ang), np.sin(ang)], axis=1)
        # a segment through the ring's centre, nudged to pass ~1e-12..1e-8
        # from one vertex
        v = ring[int(rng.integers(len(ring)))]
        d = rng.normal(size=2)
        d /= np.linalg.norm(d)
        n = np.array([-d[1], d[0]])
        off = 10.0 ** rng.uniform(-12, -8) * rng.choice([-1.0, 1.0])
        a = v + n * off - d * r * rng.uniform(0.5, 3)
        b = v + n * off + d * r * rng.uniform(0.5, 3)
        box = (ring[:, 0].min(), ring[:, 1].min(), ring[:, 0].max(), ring[:, 1].max())
        tol = candidate_tol32(box, a, b)
        for j in range(len(ring)):
            u, w = ring[j], ring[(j + 1) % len(ring)]
            for s32, s64, exact in ((cross32(a, b, u), cross64(a, b, u), cross_exact(a, b, u)),
                                    (cross32(u, w, a), cross64(u, w, a), cross_exact(u, w, a)),
                                    (cross32(u, w, b), cross64(u, w, b), cross_exact(u, w, b))):
                if abs(s32) > tol:
                    decided += 1
                    assert sign(s32) == sign(exact) == sign(s64) != 0
    assert decided > 10000


def value32(s, x):
    s, x = np.asarray(s, dtype=np.float32), np.asarray(x, dtype=np.float32)
    A = np.abs(s[0]) * np.abs(x[0]) + np.abs(s[1]) * np.abs(x[1]) + np.abs(s[2])
    return s[0] * x[0] + s[1] * x[1] + s[2], A


def value64(s, x):
    return s[0] * x[0] + s[1] * x[1] + s[2]


def value_exact(s, x):
    return Fr(float(s[0])) * Fr(float(x[0])) + Fr(float(s[1])) * Fr(float(x[1])) + Fr(float(s[2]))


def depth_violations(shrink, count=6000, seed=1):
    """Two depth planes that nearly agree at a point (their difference 1e-12
    to 1e-3 of the depths), as for surfaces meeting near a sample."""
    rng = np.random.default_rng(seed)
    decided = bad = 0
    for _ in range(count):
        x = rng.uniform(0, 1, 2)
        s1 = rng.normal(size=3) * 10.0 ** rng.uniform(-2, 2)
        z1 = value64(s1, x)
        s2 = s1 + rng.normal(size=3) * abs(s1).max() * 10.0 ** rng.uniform(-3, 0)
        # move s2's constant so its depth at x is z1 plus a tiny offset
        s2[2] += z1 - value64(s2, x) + abs(z1) * 10.0 ** rng.uniform(-12, -3) * rng.choice([-1.0, 1.0])
        h32, Ah = value32(s2, x)
        z32, Az = value32(s1, x)
        tol = (DEPTH_TOL32 * (Ah + Az) + np.float32(1e-37)) / np.float32(shrink)
        diff = h32 - z32
        if abs(diff) <= tol:
            continue
        decided += 1
        e = value_exact(s2, x) - value_exact(s1, x)
        h64, z64 = value64(s2, x), value64(s1, x)
        if not (sign(diff) == sign(e) and (h64 < z64) == (diff < 0) and h64 != z64):
            bad += 1
    return decided, bad


def test_depth_bound_holds():
    decided, bad = depth_violations(1)
    assert decided > 1000 and bad == 0, (decided, bad)


def test_depth_bound_is_not_slack():
    _, bad = depth_violations(64)
    assert bad > 0
