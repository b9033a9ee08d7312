"""Learned pool sizes for the event route.

The reference sizes every pool as a fixed multiple of the face count
(``EventCapacities.for_mesh``, warp_peel/events.py:19-33) and reruns with a
larger ``--pool-scale`` when a pool overflows (warp_peel/__main__.py:24-47).
Those multiples are far above what closed meshes use and below what crossing
soups need. Here a small network predicts each pool's use from cheap mesh
statistics, a per-pool margin calibrated on held-out runs turns the prediction
into a capacity, and the status bits catch the rare overflow, after which the
extraction is rerun with the conservative capacities.

* ``mesh_features``: projected-area and edge-length statistics, contour-edge
  count and length, clipping and facing fractions (host numpy, before upload);
* ``PoolModel``: standardisation, network parameters and margins; inference
  runs the device kernel ``pool_mlp_forward``;
* ``PoolTrainer``: full-batch training on the device with the library's own
  gradient and Adam kernels (``pool_mlp_gradient``, ``pool_mlp_adam``);
* ``collect_runs``: peels generated meshes and logs what each pool used.

The network, its loss and the kernels are described in ``csrc/pool_mlp.hip``.
"""

import ctypes
from dataclasses import replace
from pathlib import Path

import numpy as np

from . import _lib
from .events import EventCapacities

MODEL_PATH = Path(__file__).resolve().parent / "pool_model.npz"

FEATURES = (
    "log_faces",
    "log_vertices",
    "log_total_area",
    "log_mean_area",
    "log_median_area",
    "log_mean_edge",
    "edge_spread",
    "log_contours",
    "log_contour_length",
    "clipped_fraction",
    "perspective",
    "back_fraction",
    "covered_fraction",
    "log_depth_mean",
    "log_depth_max",
)
DEPTH_GRID = 64
# pool -> status key of what the pool holds after an extraction
TARGETS = (
    ("frontier", "frontier_slots"),
    ("regions", "polygons"),
    ("hits", "ray_hits"),
    ("vertices", "boundary_vertices"),
    ("triangles", "triangles"),
)
# approximate device bytes per element of each pool (for reports only; the
# workspace reports its exact size as ``device_bytes``)
ELEMENT_BYTES = {"frontier": 532, "regions": 650, "hits": 40, "vertices": 126, "triangles": 128}
MIN_CAPACITY = 64
# Closed meshes of at least this many faces get a per-face floor on every
# predicted pool (PoolModel.floor).
FLOOR_MIN_FACES = 4096
CLOSED_KINDS = ("icosphere", "knot")


def mesh_features(vertices, faces, camera, settings):
    """Feature vector (len(FEATURES),) of one mesh under its camera."""
    V = np.asarray(vertices, dtype=np.float64).reshape(-1, 3)
    F = np.asarray(faces, dtype=np.int64).reshape(-1, 3)
    C = np.asarray(camera, dtype=np.float64).reshape(4, 4)
    S = np.asarray(settings, dtype=np.float64).reshape(6)
    nf = len(F)
    out = np.zeros(len(FEATURES))
    out[0] = np.log(max(nf, 1))
    out[10] = float(S[4] != 0)
    if nf == 0:
        return out
    cam = V @ C[:3, :3].T + C[:3, 3]
    if S[4]:
        front = cam[:, 2] > S[5]
        depth = np.where(front, cam[:, 2], 1.0)
        uv = cam[:, :2] / depth[:, None]
    else:
        front = np.ones(len(V), dtype=bool)
        uv = cam[:, :2]
    uv = (uv - S[:2]) / (S[2:4] - S[:2])
    ok = front[F].all(axis=1)
    tri = uv[F]
    e1, e2 = tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]
    signed = np.where(ok, 0.5 * (e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]), 0.0)
    area = np.abs(signed)
    inside = ok & (tri >= 0).all(axis=(1, 2)) & (tri <= 1).all(axis=(1, 2))
    lengths = np.linalg.norm(tri - np.roll(tri, -1, axis=1), axis=2)[ok]
    _, canon = np.unique(V, axis=0, return_inverse=True)
    canon = canon.reshape(-1)
    out[1] = np.log(max(int(canon.max()) + 1, 1))
    out[2] = np.log1p(area.sum())
    out[3] = np.log(area.mean() + 1e-12)
    out[4] = np.log(np.median(area) + 1e-12)
    if lengths.size:
        out[5] = np.log(lengths.mean() + 1e-12)
        out[6] = lengths.std() / (lengths.mean() + 1e-12)
    # contour edges: a face-edge slot is internal iff its undirected edge has
    # exactly two slots, traversed in opposite directions, on faces that face
    # the same way (the edge topology rule of warp_peel/kernels.py:471-536)
    fc = canon[F]
    a = fc.reshape(-1)
    b = np.roll(fc, -1, axis=1).reshape(-1)
    facing = np.repeat(np.sign(signed), 3)
    key = np.minimum(a, b) * (int(canon.max()) + 1) + np.maximum(a, b)
    order = np.argsort(key, kind="stable")
    ks = key[order]
    start = np.r_[True, ks[1:] != ks[:-1]]
    group = np.cumsum(start) - 1
    sizes = np.bincount(group)
    internal = np.zeros(len(key), dtype=bool)
    pair = sizes[group] == 2
    first = order[start & pair]
    second = order[np.r_[start[1:], True] & pair & ~start]
    good = (a[first] != a[second]) & (facing[first] == facing[second]) & (facing[first] != 0)
    internal[first[good]] = True
    internal[second[good]] = True
    contour = ~internal
    slot_len = np.linalg.norm(tri - np.roll(tri, -1, axis=1), axis=2).reshape(-1)
    slot_ok = np.repeat(ok, 3)
    out[7] = np.log1p(contour.sum())
    out[8] = np.log1p(slot_len[contour & slot_ok].sum())
    out[9] = 1.0 - inside.mean()
    out[11] = float((signed < 0).mean())
    out[12:15] = depth_complexity(tri[ok], area[ok])
    return out


def closed_mesh(vertices, faces):
    """True if every undirected edge of the welded mesh has exactly two
    incidences in opposite directions (a closed, consistently oriented
    surface, which may have several components)."""
    V = np.asarray(vertices, dtype=np.float64).reshape(-1, 3)
    F = np.asarray(faces, dtype=np.int64).reshape(-1, 3)
    if len(F) == 0:
        return False
    _, canon = np.unique(V, axis=0, return_inverse=True)
    fc = canon.reshape(-1)[F]
    a, b = fc.reshape(-1), np.roll(fc, -1, axis=1).reshape(-1)
    n = int(canon.max()) + 1
    directed = a * n + b
    undirected = np.minimum(a, b) * n + np.maximum(a, b)
    _, counts = np.unique(undirected, return_counts=True)
    if (counts != 2).any():
        return False
    return np.unique(directed).size == directed.size


def closed_mesh_floor(usage, faces, kinds, quantile=0.995, min_faces=FLOOR_MIN_FACES):
    """Per-face floor of each pool for closed meshes: the `quantile` of
    use / faces over the logged runs of closed-surface generators with at least
    `min_faces` faces. Their per-face use stays in a narrow band whatever the
    depth complexity (each closed surface adds fragments along the contours
    that cross it and one ray hit per covering surface), which lets unseen
    closed meshes (such as nested shells) be admitted where the network,
    trained on shallower scenes, under-predicts."""
    usage, faces, kinds = np.asarray(usage, np.float64), np.asarray(faces), np.asarray(kinds)
    pick = np.isin(kinds, CLOSED_KINDS) & (faces >= min_faces)
    if not pick.any():
        return np.zeros(usage.shape[1])
    return np.quantile(usage[pick] / faces[pick][:, None], quantile, axis=0)


def depth_complexity(tri, area, grid=DEPTH_GRID):
    """(covered fraction, log mean depth, log max depth) of projected faces:
    each face's area is spread evenly over the window cells its bounding box
    touches (a weighted difference array, then prefix sums), and a cell's
    depth is the area it received over its own area."""
    if len(tri) == 0:
        return 0.0, 0.0, 0.0
    lo = np.clip(np.floor(np.clip(tri.min(axis=1), 0.0, 1.0) * grid), 0, grid - 1).astype(np.int64)
    hi = np.clip(np.floor(np.clip(tri.max(axis=1), 0.0, 1.0) * grid), 0, grid - 1).astype(np.int64)
    cells = (hi[:, 0] - lo[:, 0] + 1) * (hi[:, 1] - lo[:, 1] + 1)
    w = area / cells
    diff = np.zeros((grid + 1, grid + 1))
    np.add.at(diff, (lo[:, 1], lo[:, 0]), w)
    np.add.at(diff, (lo[:, 1], hi[:, 0] + 1), -w)
    np.add.at(diff, (hi[:, 1] + 1, lo[:, 0]), -w)
    np.add.at(diff, (hi[:, 1] + 1, hi[:, 0] + 1), w)
    depth = diff.cumsum(axis=0).cumsum(axis=1)[:grid, :grid] * grid * grid
    covered = depth > 1e-6
    if not covered.any():
        return 0.0, 0.0, 0.0
    return float(covered.mean()), float(np.log1p(depth[covered].mean())), float(np.log1p(depth.max()))


def usage_of(info):
    """Per-pool use (len(TARGETS),) from ``EventPeeler.read_status()``."""
    return np.array([info[key] for _, key in TARGETS], dtype=np.float64)


def _device_call(code):
    if code != 0:
        raise RuntimeError(f"pool network kernel failed ({code})")


class PoolModel:
    """Standardisation, network parameters and per-pool log margins."""

    def __init__(self, params, x_mean, x_std, y_mean, y_std, margin, meta=None, floor=None):
        self.params = np.asarray(params, dtype=np.float32)
        self.x_mean, self.x_std = np.asarray(x_mean, np.float64), np.asarray(x_std, np.float64)
        self.y_mean, self.y_std = np.asarray(y_mean, np.float64), np.asarray(y_std, np.float64)
        self.margin = np.asarray(margin, dtype=np.float64)
        self.meta = dict(meta or {})
        self.inputs, self.outputs = len(self.x_mean), len(self.y_mean)
        # per-face use floor of closed meshes (closed_mesh_floor), zeros = none
        self.floor = np.zeros(self.outputs) if floor is None else np.asarray(floor, dtype=np.float64)
        self._device = {}

    def save(self, path):
        np.savez(path, params=self.params, x_mean=self.x_mean, x_std=self.x_std, y_mean=self.y_mean,
                 y_std=self.y_std, margin=self.margin, floor=self.floor,
                 meta_keys=np.array(list(self.meta.keys())),
                 meta_values=np.array([float(v) for v in self.meta.values()]))

    @classmethod
    def load(cls, path=MODEL_PATH):
        with np.load(path, allow_pickle=False) as z:
            meta = dict(zip(z["meta_keys"].tolist(), z["meta_values"].tolist()))
            floor = z["floor"] if "floor" in z.files else None
            return cls(z["params"], z["x_mean"], z["x_std"], z["y_mean"], z["y_std"], z["margin"], meta, floor)

    def _params_on(self, device):
        import torch

        key = str(device)
        if key not in self._device:
            self._device[key] = torch.as_tensor(self.params, device=device)
        return self._device[key]

    def predict_log_usage(self, features, device="cuda:0"):
        """log1p(pool use) predicted for each feature row, on the device. The
        network predicts use per face, log1p(use) - log(faces) (feature 0),
        which carries over to meshes larger than any it was trained on."""
        import torch

        features = np.atleast_2d(features)
        x = (features - self.x_mean) / self.x_std
        xt = torch.as_tensor(x.astype(np.float32), device=device).contiguous()
        y = torch.empty((len(x), self.outputs), dtype=torch.float32, device=device)
        lib = _lib.load()
        stream = torch.cuda.current_stream(device)
        _device_call(lib.pool_mlp_forward(self._params_on(device).data_ptr(), self.inputs, self.outputs,
                                          xt.data_ptr(), len(x), y.data_ptr(), stream.cuda_stream))
        return y.double().cpu().numpy() * self.y_std + self.y_mean + features[:, :1]

    def capacities(self, features, faces, device="cuda:0", closed=False):
        """EventCapacities from the prediction plus the calibrated margin."""
        return self.capacities_many(np.atleast_2d(features), [faces], device, [closed])[0]

    def capacities_many(self, features, faces, device="cuda:0", closed=None):
        """EventCapacities for each feature row (one device launch); a closed
        mesh of at least FLOOR_MIN_FACES faces gets at least `floor` per face
        in every pool."""
        use = np.expm1(self.predict_log_usage(features, device) + self.margin)
        closed = [False] * len(use) if closed is None else closed
        out = []
        for row, n, shut in zip(use, faces, closed):
            if shut and n >= FLOOR_MIN_FACES:
                row = np.maximum(row, self.floor * n)
            caps = {pool: max(MIN_CAPACITY, int(np.ceil(u))) for (pool, _), u in zip(TARGETS, row)}
            out.append(replace(EventCapacities.for_mesh(int(n)), **caps))
        return out


def load_default():
    if not MODEL_PATH.exists():
        raise RuntimeError(f"{MODEL_PATH} is missing: train it with `python -m hip_peel.train_pools`")
    return PoolModel.load(MODEL_PATH)


class PoolTrainer:
    """Full-batch training of the pool network on the device, with the
    library's gradient and Adam kernels; the parameters stay in HBM."""

    def __init__(self, inputs, outputs, device="cuda:0", seed=0):
        import torch

        self.lib = _lib.load()
        self.inputs, self.outputs, self.device = inputs, outputs, device
        count = self.lib.pool_mlp_param_count(inputs, outputs)
        if count < 0:
            raise ValueError("unsupported network shape")
        self.count = count
        self.params = torch.as_tensor(init_params(inputs, outputs, seed), device=device)
        self.m = torch.zeros(count, dtype=torch.float32, device=device)
        self.v = torch.zeros(count, dtype=torch.float32, device=device)
        self.grad = torch.zeros(count, dtype=torch.float32, device=device)
        self.loss = torch.zeros(1, dtype=torch.float32, device=device)
        self.partial = None
        self.steps = 0

    def _stream(self):
        import torch

        return torch.cuda.current_stream(self.device).cuda_stream

    def gradient(self, x, y, under):
        """Fill ``self.grad`` and ``self.loss`` for device tensors x (N, I), y (N, O)."""
        import torch

        n = x.shape[0]
        need = self.lib.pool_mlp_partial_count(self.inputs, self.outputs, n)
        if self.partial is None or self.partial.numel() < need:
            self.partial = torch.empty(need, dtype=torch.float32, device=self.device)
        _device_call(self.lib.pool_mlp_gradient(
            self.params.data_ptr(), self.inputs, self.outputs, x.data_ptr(), y.data_ptr(), n, ctypes.c_float(under),
            self.partial.data_ptr(), self.grad.data_ptr(), self.loss.data_ptr(), self._stream()))

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        self.steps += 1
        b1, b2 = 1.0 - beta1 ** self.steps, 1.0 - beta2 ** self.steps
        _device_call(self.lib.pool_mlp_adam(
            self.params.data_ptr(), self.grad.data_ptr(), self.m.data_ptr(), self.v.data_ptr(), self.count,
            lr, beta1, beta2, eps, b1, b2, self._stream()))

    def fit(self, x, y, steps, lr=3e-3, under=4.0, log_every=0):
        """Adam on the full batch; returns the loss after every ``log_every``
        steps (host reads only then)."""
        import torch

        x = torch.as_tensor(np.asarray(x, np.float32), device=self.device).contiguous()
        y = torch.as_tensor(np.asarray(y, np.float32), device=self.device).contiguous()
        curve = []
        for k in range(steps):
            self.gradient(x, y, under)
            self.adam(lr)
            if log_every and (k + 1) % log_every == 0:
                curve.append(float(self.loss.item()))
        return curve


def init_params(inputs, outputs, seed=0):
    """Glorot-uniform weights, zero biases, flat in the kernels' layout."""
    rng = np.random.default_rng(seed)
    h = 64

    def glorot(rows, cols):
        bound = np.sqrt(6.0 / (rows + cols))
        return rng.uniform(-bound, bound, (rows, cols))

    parts = [glorot(h, inputs).ravel(), np.zeros(h), glorot(h, h).ravel(), np.zeros(h),
             glorot(outputs, h).ravel(), np.zeros(outputs)]
    return np.concatenate(parts).astype(np.float32)


def calibrate_margin(log_target, log_pred, quantile=0.995, slack=0.05):
    """Per-pool log margin: the ``quantile`` of the held-out under-prediction,
    plus a slack, and never below zero."""
    gap = np.asarray(log_target) - np.asarray(log_pred)
    return np.maximum(np.quantile(gap, quantile, axis=0), 0.0) + slack


def mesh_specs(count, seed=0):
    """Seeded recipes (kind, seed, parameters) for ``build_mesh``, drawn over
    the size ranges of the benchmark configurations' scene generators: noisy
    icospheres (levels 1-6, 80 to 81,920 faces) under pinholes, torus-knot
    tubes (up to 256 x 128), identical-footprint stacks (up to 2,048) and
    resolved random triangle soups (2 to 12 triangles)."""
    rng = np.random.default_rng(seed)
    kinds = rng.choice(4, size=count, p=[0.4, 0.3, 0.1, 0.2])
    specs = []
    for kind in kinds:
        s = int(rng.integers(1 << 30))
        if kind == 0:
            specs.append(("icosphere", s, (int(rng.integers(1, 7)), float(rng.uniform(0.0, 0.4)),
                                           float(rng.uniform(2.2, 5.0)))))
        elif kind == 1:
            specs.append(("knot", s, (int(rng.choice([16, 32, 48, 64, 96, 128, 192, 256])),
                                      int(rng.choice([8, 16, 24, 32, 48, 64, 96, 128])),
                                      float(rng.uniform(0.04, 0.1)))))  # thin: the tube never meets itself
        elif kind == 2:
            specs.append(("stack", s, (int(rng.integers(1, 2049)),)))
        else:
            specs.append(("soup", s, (int(rng.integers(2, 13)),)))
    return specs


def build_mesh(spec):
    """(kind, (vertices, faces, camera, settings)) of one recipe."""
    from . import scenes

    kind, s, args = spec
    if kind == "icosphere":
        level, amp, dist = args
        v, f, c, st = scenes.icosphere_instances(1, seed=s, levels=(level, level), amplitude=amp, distance=dist)[0]
        return kind, (v, f, c, st)
    if kind == "knot":
        seg, sides, radius = args
        return kind, scenes.torus_knot(seed=s, segments=seg, sides=sides, radius=radius)
    if kind == "stack":
        return kind, scenes.stack(args[0])
    v, f, c, st, _ = scenes.random_triangles_resolved(args[0], seed=s)
    return kind, (v, f, c, st)


def _build_with_features(spec):
    kind, mesh = build_mesh(spec)
    return kind, mesh, mesh_features(*mesh)


def generated_meshes(count, seed=0):
    """Meshes of ``mesh_specs(count, seed)``, built in order."""
    for spec in mesh_specs(count, seed):
        yield build_mesh(spec)


def collect_runs(count, seed=0, device="cuda:0", progress=None, workers=1):
    """Peel ``count`` generated meshes (conservative capacities, grown until
    nothing overflows) and log features and per-pool use. ``workers`` > 1
    builds the meshes in that many spawned processes while this one peels.
    Returns (features (N, len(FEATURES)), usage (N, len(TARGETS)), kinds,
    faces); meshes that fail for a reason no pool size cures are skipped."""
    from .events import EventPeeler
    from .io import upload

    specs = mesh_specs(count, seed)
    if workers > 1:
        import multiprocessing
        from concurrent.futures import ProcessPoolExecutor

        pool = ProcessPoolExecutor(workers, mp_context=multiprocessing.get_context("spawn"))
        built = pool.map(_build_with_features, specs, chunksize=4)
    else:
        pool = None
        built = map(_build_with_features, specs)
    feats, uses, kinds, faces = [], [], [], []
    skipped = 0
    try:
        for k, (kind, mesh, features) in enumerate(built):
            arrays = upload(*mesh, device=device)
            n = mesh[1].shape[0]
            peeler = None
            for scale in (1, 4, 16, 64):
                cap = EventCapacities.for_mesh(n)
                cap = replace(cap, **{p: getattr(cap, p) * scale for p, _ in TARGETS})
                peeler = EventPeeler(*arrays, capacities=cap)
                peeler.warmup()
                status, _ = peeler._read()
                if status[9] & ~POOL_BITS:
                    peeler = None  # a failure no pool size cures (noding, layers, rounds, input)
                    break
                if status[9] == 0:
                    break
                peeler = None
            if peeler is None:
                skipped += 1
                continue
            info = peeler.read_status()
            del peeler
            feats.append(features)
            uses.append(usage_of(info))
            kinds.append(kind)
            faces.append(n)
            if progress and (k + 1) % progress == 0:
                print(f"runs {k + 1}/{count} (skipped {skipped})", flush=True)
    finally:
        if pool is not None:
            pool.shutdown()
    return np.array(feats), np.array(uses), np.array(kinds), np.array(faces)


def admitted_event_peeler(vertices, faces, camera, settings, model=None, **options):
    """EventPeeler sized by the learned pools, run once; if a pool overflowed,
    rebuilt with the conservative capacities (``EventCapacities.for_mesh``).
    Takes device tensors; returns (peeler, rebuilt)."""
    from .events import EventPeeler

    model = model or load_default()
    host = [t.detach().cpu().numpy() for t in (vertices, faces, camera, settings)]
    n = host[1].reshape(-1, 3).shape[0]
    cap = model.capacities(mesh_features(*host), n, device=vertices.device, closed=closed_mesh(*host[:2]))
    peeler = EventPeeler(vertices, faces, camera, settings, capacities=cap, **options)
    peeler.run()
    status, _ = peeler._read()
    if status[9] & CAPACITY_BITS:
        del peeler
        peeler = EventPeeler(vertices, faces, camera, settings, **options)
        peeler.run()
        return peeler, True
    return peeler, False


CAPACITY_BITS = 4 | 8 | 16 | 32 | 128 | 1024 | 2048
# status bits of the five predicted pools (frontier, output polygons, boundary
# vertices, triangles, ray hits)
POOL_BITS = 4 | 8 | 16 | 32 | 1024
