"""Procedural benchmark scenes (float64, seeded), host-side numpy.

Each generator returns (vertices (V, 3) f64, faces (F, 3) i32, camera (4, 4) f64,
settings (6,) f64) in the peelers' input layout:

* ``random_triangles``  256 triangles, vertices i.i.d. uniform in [0, 1]^3,
  orthographic camera looking along -z.
* ``torus_knot``        closed (2,3) torus-knot tube, 256 x 128 grid = 65,536
  faces, orthographic view along a seeded random unit direction.
* ``stack``             k-th congruent triangle (0,0),(1,0),(0,1) at z = k 2^-20.
* ``nested_shells``     13 closed noisy geodesic spheres of radius 1.0 .. 2.2,
  ~55K faces each (717,740 total), orthographic along z.
"""

import functools

import numpy as np


def orthographic_view(vertices, direction, margin=0.05):
    """World-to-camera rotation whose camera z axis is ``direction`` (det +1)
    and a window around the projected bounding box."""
    d = np.asarray(direction, dtype=np.float64)
    d = d / np.linalg.norm(d)
    helper = np.array([1.0, 0.0, 0.0]) if abs(d[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    e1 = np.cross(helper, d)
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(d, e1)
    camera = np.eye(4)
    camera[0, :3], camera[1, :3], camera[2, :3] = e1, e2, d
    xy = vertices @ camera[:2, :3].T
    lo, hi = xy.min(axis=0), xy.max(axis=0)
    pad = margin * (hi - lo).max()
    settings = np.array([lo[0] - pad, lo[1] - pad, hi[0] + pad, hi[1] + pad, 0.0, 0.01])
    return camera, settings


def grid_faces(rows, cols, wrap_rows=True, wrap_cols=True):
    """Two triangles per quad of a (rows x cols) vertex grid, counter-clockwise."""
    faces = []
    r_end = rows if wrap_rows else rows - 1
    c_end = cols if wrap_cols else cols - 1
    for i in range(r_end):
        for j in range(c_end):
            a = i * cols + j
            b = i * cols + (j + 1) % cols
            c = ((i + 1) % rows) * cols + (j + 1) % cols
            d = ((i + 1) % rows) * cols + j
            faces.append((a, b, c))
            faces.append((a, c, d))
    return np.array(faces, dtype=np.int32)


def torus_knot(seed=0, p=2, q=3, segments=256, sides=128, radius=0.08, scale=1.0):
    t = np.linspace(0.0, 2.0 * np.pi, segments, endpoint=False)

    def curve(t):
        r = 2.0 + np.cos(q * t)
        return scale * np.stack([r * np.cos(p * t), r * np.sin(p * t), -np.sin(q * t)], axis=-1)

    c = curve(t)
    h = 1e-5
    tangent = curve(t + h) - curve(t - h)
    tangent /= np.linalg.norm(tangent, axis=1)[:, None]
    # parallel transport of a normal, then remove the holonomy linearly
    normals = np.zeros_like(c)
    n0 = np.cross(tangent[0], [0.0, 0.0, 1.0])
    normals[0] = n0 / np.linalg.norm(n0)
    for i in range(1, segments):
        n = normals[i - 1] - np.dot(normals[i - 1], tangent[i]) * tangent[i]
        normals[i] = n / np.linalg.norm(n)
    end = normals[-1] - np.dot(normals[-1], tangent[0]) * tangent[0]
    end /= np.linalg.norm(end)
    binormal0 = np.cross(tangent[0], normals[0])
    twist = np.arctan2(np.dot(end, binormal0), np.dot(end, normals[0]))
    verts = np.zeros((segments, sides, 3))
    for i in range(segments):
        b = np.cross(tangent[i], normals[i])
        angle = -twist * i / segments
        n = np.cos(angle) * normals[i] + np.sin(angle) * b
        b = np.cross(tangent[i], n)
        phi = 2.0 * np.pi * np.arange(sides) / sides
        verts[i] = c[i] + radius * scale * (np.cos(phi)[:, None] * n + np.sin(phi)[:, None] * b)
    vertices = verts.reshape(-1, 3)
    faces = grid_faces(segments, sides)
    rng = np.random.default_rng(seed)
    direction = rng.normal(size=3)
    camera, settings = orthographic_view(vertices, direction)
    return vertices, faces, camera, settings


def stack(n=65536):
    tri = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
    v = np.zeros((n, 3, 3))
    v[:, :, :2] = tri
    v[:, :, 2] = (np.arange(n) * 2.0**-20)[:, None]
    faces = np.arange(3 * n, dtype=np.int32).reshape(-1, 3)
    settings = np.array([-1 / 16, -1 / 16, 17 / 16, 17 / 16, 0.0, 0.01])
    return v.reshape(-1, 3), faces, np.eye(4), settings


def _value_noise(points, seed, octaves=3, frequency=2.0):
    """Smooth 3D value noise in [-1, 1] (trilinear interpolation of hashed
    lattice values, smoothstep weights), summed over octaves."""
    rng = np.random.default_rng(seed)
    table = rng.uniform(-1.0, 1.0, size=4096)
    perm = rng.permutation(4096)
    total = np.zeros(len(points))
    amplitude, norm = 1.0, 0.0
    for o in range(octaves):
        x = points * frequency * (2.0**o)
        i = np.floor(x).astype(np.int64)
        f = x - i
        f = f * f * (3.0 - 2.0 * f)
        acc = np.zeros(len(points))
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    h = perm[(i[:, 0] + dx) % 4096]
                    h = perm[(h + i[:, 1] + dy) % 4096]
                    h = perm[(h + i[:, 2] + dz) % 4096]
                    wgt = (
                        (f[:, 0] if dx else 1 - f[:, 0])
                        * (f[:, 1] if dy else 1 - f[:, 1])
                        * (f[:, 2] if dz else 1 - f[:, 2])
                    )
                    acc += wgt * table[h]
        total += amplitude * acc
        norm += amplitude
        amplitude *= 0.5
    return total / norm


def uv_sphere(radius, n_lat=126, n_lon=222, amplitude=0.0, seed=0):
    """Closed UV sphere: 2 n_lon (n_lat - 1) faces; radial value-noise offset."""
    theta = np.pi * np.arange(1, n_lat) / n_lat
    phi = 2.0 * np.pi * np.arange(n_lon) / n_lon
    st, ct = np.sin(theta), np.cos(theta)
    ring = np.stack(
        [st[:, None] * np.cos(phi)[None], st[:, None] * np.sin(phi)[None], np.repeat(ct[:, None], n_lon, 1)],
        axis=-1,
    ).reshape(-1, 3)
    dirs = np.concatenate([np.array([[0.0, 0.0, 1.0]]), ring, np.array([[0.0, 0.0, -1.0]])])
    r = radius + amplitude * _value_noise(dirs * radius, seed)
    vertices = dirs * r[:, None]
    rows = n_lat - 1
    faces = []
    for j in range(n_lon):  # north cap
        faces.append((0, 1 + j, 1 + (j + 1) % n_lon))
    body = grid_faces(rows, n_lon, wrap_rows=False, wrap_cols=True) + 1
    faces = np.concatenate([np.array(faces, dtype=np.int32), body])
    south = len(dirs) - 1
    base = 1 + (rows - 1) * n_lon
    cap = [(base + j, south, base + (j + 1) % n_lon) for j in range(n_lon)]
    faces = np.concatenate([faces, np.array(cap, dtype=np.int32)])
    # poles on the x axis (in the image plane of a view along z), so no polar
    # fan faces the camera; a rotation about y keeps the mesh closed
    vertices = np.stack([vertices[:, 2], vertices[:, 1], -vertices[:, 0]], axis=1)
    # orient every face outward
    a, b, c = vertices[faces[:, 0]], vertices[faces[:, 1]], vertices[faces[:, 2]]
    outward = np.einsum("ij,ij->i", np.cross(b - a, c - a), a + b + c) < 0
    faces[outward] = faces[outward][:, [0, 2, 1]]
    return vertices, faces.astype(np.int32)


def icosahedron():
    t = (1.0 + 5.0**0.5) / 2.0
    v = np.array(
        [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
         [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]],
        dtype=np.float64,
    )
    f = np.array(
        [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
         [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8],
         [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]],
        dtype=np.int32,
    )
    return v / np.linalg.norm(v, axis=1)[:, None], f


def _lattice_triangles(m):
    """Lattice (a, b) of the three corners of every small triangle of one
    icosahedron face split m x m, in row order: for each a, the upright
    triangle at b followed by the inverted one at b (none after the last b)."""
    rows = []
    for a in range(m):
        b = np.arange(m - a)
        up = np.stack([np.stack([np.full_like(b, a), b], 1), np.stack([np.full_like(b, a + 1), b], 1),
                       np.stack([np.full_like(b, a), b + 1], 1)], 1)
        c = b[:-1]
        down = np.stack([np.stack([np.full_like(c, a + 1), c], 1), np.stack([np.full_like(c, a + 1), c + 1], 1),
                         np.stack([np.full_like(c, a), c + 1], 1)], 1)
        both = np.empty((len(b) + len(c), 3, 2), dtype=np.int64)
        both[0::2], both[1::2] = up, down
        rows.append(both)
    return np.concatenate(rows)


@functools.lru_cache(maxsize=16)
def _geodesic_lattice(m):
    """Unit directions and faces of the class-I geodesic sphere of frequency m
    (before orientation): a lattice point is identified by its integer
    barycentric weights on the icosahedron corners, so points shared by
    neighbouring faces are one vertex (closed mesh), numbered in order of first
    use; its position is the weighted sum of its corners in ascending corner
    order, divided by m, then normalised."""
    if not 1 <= m < 128:
        raise ValueError("geodesic_sphere: 1 <= m < 128 (weights are 7-bit fields of the point keys)")
    corners, ico = icosahedron()
    tri = _lattice_triangles(m)
    a, b = tri[..., 0].ravel(), tri[..., 1].ravel()
    n = len(a)
    ids = np.repeat(ico.astype(np.int64), n, axis=0).reshape(len(ico), n, 3)
    w = np.stack([np.broadcast_to(m - a - b, (len(ico), n)), np.broadcast_to(a, (len(ico), n)),
                  np.broadcast_to(b, (len(ico), n))], axis=-1)
    # key: the (corner << 7 | weight) pairs with positive weight, ascending;
    # absent pairs are 15 << 7 so that they sort last
    pair = np.where(w > 0, (ids << 7) | w, 15 << 7)
    pair.sort(axis=-1)
    key = (pair[..., 0] << 22) | (pair[..., 1] << 11) | pair[..., 2]
    unique, first, inverse = np.unique(key.ravel(), return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    number = np.empty(len(order), dtype=np.int64)
    number[order] = np.arange(len(order))
    faces = number[inverse].reshape(-1, 3).astype(np.int32)
    k = unique[order]
    p = np.zeros((len(k), 3))
    for shift in (22, 11, 0):
        q = (k >> shift) & 0x7FF
        corner, weight = q >> 7, (q & 0x7F).astype(np.float64)
        p = np.where((corner < 12)[:, None], p + weight[:, None] * corners[np.minimum(corner, 11)], p)
    p = p / m
    dirs = p / np.array([np.linalg.norm(x) for x in p])[:, None]
    dirs.flags.writeable = False
    faces.flags.writeable = False
    return dirs, faces


def geodesic_sphere(radius, m, amplitude=0.0, seed=0):
    """Class-I geodesic sphere: every icosahedron face split into m^2 triangles
    (20 m^2 faces), lattice points projected to the sphere, every face
    oriented outward. Radial value-noise offset of the given amplitude."""
    dirs, faces = _geodesic_lattice(m)
    r = radius + amplitude * _value_noise(dirs * radius, seed)
    vertices = dirs * r[:, None]
    faces = faces.copy()
    a, b, c = vertices[faces[:, 0]], vertices[faces[:, 1]], vertices[faces[:, 2]]
    inward = np.einsum("ij,ij->i", np.cross(b - a, c - a), a + b + c) < 0
    faces[inward] = faces[inward][:, [0, 2, 1]]
    return vertices, faces


def nested_shells(count=13, r0=1.0, step=0.1, amplitude=0.02, seed=0):
    """13 concentric noisy geodesic spheres, radius 1.0 .. 2.2: the 6 inner
    shells with 20 * 52^2 = 54,080 faces, the 7 outer with 20 * 53^2 = 56,180
    (717,740 faces in total), orthographic view along z."""
    vs, fs, offset = [], [], 0
    for k in range(count):
        v, f = geodesic_sphere(r0 + step * k, 52 if k < 6 else 53, amplitude, seed + k)
        vs.append(v)
        fs.append(f + offset)
        offset += len(v)
    vertices = np.concatenate(vs)
    faces = np.concatenate(fs).astype(np.int32)
    settings = np.array([-2.3, -2.3, 2.3, 2.3, 0.0, 0.01])
    return vertices, faces, np.eye(4), settings


def pinhole_look_at(eye, target=(0.0, 0.0, 0.0)):
    """Affine world-to-camera matrix (4 x 4) of a pinhole at ``eye`` looking at
    ``target``: camera z forward (depth > 0 in front), det +1."""
    eye = np.asarray(eye, dtype=np.float64)
    z = np.asarray(target, dtype=np.float64) - eye
    z /= np.linalg.norm(z)
    up = np.array([0.0, 0.0, 1.0]) if abs(z[2]) < 0.99 else np.array([0.0, 1.0, 0.0])
    x = np.cross(up, z)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    m = np.eye(4)
    m[:3, :3] = np.stack([x, y, z])
    m[:3, 3] = -m[:3, :3] @ eye
    return m


def icosphere_instances(count=256, seed=0, levels=(2, 6), amplitude=0.3, distance=3.0):
    """Batched uneven instances: ``count`` icospheres subdivided ``levels``
    times (20 * 4^L faces, L drawn uniformly by seed), radially displaced by
    3-octave value noise of the given amplitude, each with its own normalized
    pinhole (window [-1, 1]^2, near 0.1) at ``distance`` from the origin,
    aimed at it from a random azimuth and elevation. Returns a list of
    (vertices, faces, camera, settings)."""
    rng = np.random.default_rng(seed)
    settings = np.array([-1.0, -1.0, 1.0, 1.0, 1.0, 0.1])
    out = []
    for k in range(count):
        level = int(rng.integers(levels[0], levels[1] + 1))
        azimuth = rng.uniform(0.0, 2.0 * np.pi)
        elevation = rng.uniform(-0.5 * np.pi, 0.5 * np.pi)
        eye = distance * np.array(
            [np.cos(elevation) * np.cos(azimuth), np.cos(elevation) * np.sin(azimuth), np.sin(elevation)])
        vertices, faces = geodesic_sphere(1.0, 2**level, amplitude=amplitude, seed=seed * 100003 + k)
        out.append((vertices, faces, pinhole_look_at(eye), settings.copy()))
    return out


def random_triangles(n=256, seed=0):
    rng = np.random.default_rng(seed)
    vertices = rng.random((3 * n, 3))
    faces = np.arange(3 * n, dtype=np.int32).reshape(-1, 3)
    camera = np.diag([1.0, -1.0, -1.0, 1.0])  # look along -z, det +1
    xy = vertices @ camera[:2, :3].T
    lo, hi = xy.min(axis=0), xy.max(axis=0)
    pad = 0.05 * (hi - lo).max()
    settings = np.array([lo[0] - pad, lo[1] - pad, hi[0] + pad, hi[1] + pad, 0.0, 0.01])
    return vertices, faces, camera, settings


def random_triangles_resolved(n=256, seed=0, workers=1):
    """``random_triangles`` with its surface intersections split exactly
    (``resolve.resolve_arrays``), the input contract of the contour routes.
    Returns (vertices, faces, camera, settings, parents); the window and camera
    are those of the unresolved scene."""
    from .resolve import resolve_arrays

    vertices, faces, camera, settings = random_triangles(n, seed)
    rv, rf, parents = resolve_arrays(vertices, faces, workers=workers)
    return rv, rf, camera, settings, parents


CONFIGS = {
    "torus-knot-65536": torus_knot,
    "stack-65536": stack,
    "nested-shells-717740": nested_shells,
    "random-256": random_triangles,
}
