"""Many meshes, one graph launch.

``BatchPeeler`` captures the pipelines of several peelers (each with its own
mesh, camera, capacities and outputs) into one HIP graph: the first peeler's
stream forks to every other peeler's stream and joins them again, so one
``run()`` replays every extraction and independent meshes may overlap on the
GPU. The reference peels one mesh per graph; this is an extension for callers
with many small meshes (for example 256 icospheres of 320 to 81,920 faces,
each seen by its own pinhole).

``from_meshes`` sizes each workspace with the learned pool sizes
(``pools.PoolModel``, one device inference for the whole batch) and, when a
member overflows, rebuilds that workspace with the reference's conservative
capacities and captures the batch again.
"""

import ctypes

import numpy as np

from . import _lib
from .events import EventCapacities, EventPeeler
from .pools import CAPACITY_BITS


class BatchPeeler:
    def __init__(self, peelers):
        self.peelers = list(peelers)
        if not self.peelers:
            raise ValueError("A batch needs at least one peeler")
        self._lib = _lib.load()
        handles = (ctypes.c_void_p * len(self.peelers))(*[p._handle.value for p in self.peelers])
        out = ctypes.c_void_p()
        for p in self.peelers:
            if not p.warmed:
                p.warmup()
        _lib.check(self._lib.peel_batch_create(handles, len(self.peelers), ctypes.byref(out)))
        self._handle = out
        self.stream = self.peelers[0].stream

    @classmethod
    def from_meshes(cls, meshes, device="cuda:0", pools="learned", model=None, **options):
        """Peelers for (vertices, faces, camera, settings) host arrays, each
        uploaded to ``device``. ``pools="learned"`` sizes every workspace from
        the pool network (``model`` or the trained default), runs the batch
        once, and rebuilds each overflowing member with the conservative
        capacities (``EventCapacities.for_mesh``); ``pools="conservative"``
        uses those from the start. ``rebuilt`` lists the rebuilt members."""
        from . import pools as pool_sizes
        from .io import upload

        if pools not in ("learned", "conservative"):
            raise ValueError("pools must be 'learned' or 'conservative'")
        arrays = [upload(*m[:4], device=device) for m in meshes]
        faces = [a[1].shape[0] for a in arrays]
        if pools == "learned":
            model = model or pool_sizes.load_default()
            features = [pool_sizes.mesh_features(*m[:4]) for m in meshes]
            closed = [pool_sizes.closed_mesh(*m[:2]) for m in meshes]
            caps = model.capacities_many(features, faces, device, closed)
        else:
            caps = [EventCapacities.for_mesh(n) for n in faces]
        peelers = [EventPeeler(*a, capacities=c, **options) for a, c in zip(arrays, caps)]
        batch = cls(peelers)
        batch.rebuilt = []
        if pools == "learned":
            batch.run()
            batch.synchronize()
            redo = [i for i, p in enumerate(batch.peelers) if int(p._read()[0][9]) & CAPACITY_BITS]
            if redo:
                batch.close()
                for i in redo:
                    peelers[i] = EventPeeler(*arrays[i], capacities=EventCapacities.for_mesh(faces[i]), **options)
                batch = cls(peelers)
            batch.rebuilt = redo
        return batch

    def run(self):
        """Asynchronous replay of every member's extraction."""
        for p in self.peelers:
            p._order_inputs()
        _lib.check(self._lib.peel_batch_launch(self._handle))

    def synchronize(self):
        _lib.check(self._lib.peel_batch_synchronize(self._handle))

    def graph_nodes(self):
        """Node counts of the batch graph by type (see DevicePeeler.graph_nodes)."""
        return _lib.census(self._lib.peel_batch_graph_census, self._handle)

    def read_status(self):
        """Synchronise and decode every member's status (raises like the
        members' own read_status on the first failure)."""
        self.synchronize()
        return [p.read_status() for p in self.peelers]

    def close(self):
        handle = getattr(self, "_handle", None)
        if handle is not None and handle.value:
            self._lib.peel_batch_destroy(handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Merged batches: many instances in one workspace


def atlas_scene(meshes, margin=1.0 / 64.0):
    """One orthographic scene holding every instance in its own tile.

    Each instance's vertices go to its camera space; a pinhole instance then
    takes the projective map (x/z, y/z, -1/z), which sends planes to planes and
    keeps the depth order along every ray (the route's own perspective depth
    is -1/z), so an orthographic peel of the mapped mesh ranks it as the
    pinhole peel does. Window coordinates are normalised to [0, 1]^2 and
    instance k is translated to tile (k mod T, k div T) of a T x T atlas with
    pitch 1 + margin; tiles are disjoint, so instances never interact.
    Requires every vertex of a pinhole instance beyond its near plane and every
    instance inside its window (no near or window clipping is then needed).
    Returns (vertices, faces, camera, settings, layout) with layout rows
    (vertex offset, face offset, tile x, tile y)."""
    count = len(meshes)
    tiles = max(1, int(np.ceil(np.sqrt(count))))
    pitch = 1.0 + margin
    vs, fs, layout = [], [], []
    voff = foff = 0
    for k, (v, f, cam, st) in enumerate(meshes):
        v = np.asarray(v, dtype=np.float64).reshape(-1, 3)
        f = np.asarray(f, dtype=np.int64).reshape(-1, 3)
        cam = np.asarray(cam, dtype=np.float64).reshape(4, 4)
        st = np.asarray(st, dtype=np.float64).reshape(6)
        pc = v @ cam[:3, :3].T + cam[:3, 3]
        if st[4]:
            z = pc[:, 2]
            if len(z) and (z <= st[5]).any():
                raise ValueError(f"instance {k}: a merged batch needs every vertex beyond the near plane")
            u, w, d = pc[:, 0] / z, pc[:, 1] / z, -1.0 / z
        else:
            u, w, d = pc[:, 0], pc[:, 1], pc[:, 2]
        un = (u - st[0]) / (st[2] - st[0])
        wn = (w - st[1]) / (st[3] - st[1])
        if len(un) and (un.min() < 0.0 or un.max() > 1.0 or wn.min() < 0.0 or wn.max() > 1.0):
            raise ValueError(f"instance {k}: a merged batch needs every instance inside its window")
        tx, ty = k % tiles, k // tiles
        vs.append(np.stack([un + tx * pitch, wn + ty * pitch, d], axis=1))
        fs.append(f + voff)
        layout.append((voff, foff, tx, ty))
        voff += len(v)
        foff += len(f)
    side = tiles * pitch
    vertices = np.concatenate(vs) if vs else np.zeros((0, 3))
    faces = (np.concatenate(fs) if fs else np.zeros((0, 3), np.int64)).astype(np.int32)
    settings = np.array([0.0, 0.0, side, side, 0.0, 0.01])
    return vertices, faces, np.eye(4), settings, np.array(layout + [(voff, foff, 0, 0)], dtype=np.int64)


class MergedBatch:
    """Instances peeled by one workspace over an atlas of their windows
    (``atlas_scene``): every kernel of the pipeline runs once over all faces,
    instead of once per instance. Pools are the sum of the instances' learned
    sizes (``pools``), rebuilt with the conservative capacities of the whole
    scene if one overflows. Ranks are exact per instance; coordinates go
    through the atlas map, so they agree with single-instance peels to
    rounding, not bit for bit. ``instance(k)`` returns instance k's fragments
    and triangles in its own window and world frames."""

    def __init__(self, meshes, device="cuda:0", pools="learned", model=None, margin=1.0 / 64.0, **options):
        from dataclasses import replace

        from . import pools as pool_sizes
        from .io import upload

        self.meshes = [tuple(np.asarray(a) for a in m[:4]) for m in meshes]
        v, f, cam, st, self.layout = atlas_scene(self.meshes, margin)
        self.pitch = 1.0 + margin
        self.side = float(st[2])
        arrays = upload(v, f, cam, st, device=device)
        self.arrays = arrays
        faces = [len(m[1]) for m in self.meshes]
        cap = EventCapacities.for_mesh(len(f))
        if pools == "learned":
            model = model or pool_sizes.load_default()
            features = [pool_sizes.mesh_features(*m) for m in self.meshes]
            closed = [pool_sizes.closed_mesh(*m[:2]) for m in self.meshes]
            per = model.capacities_many(features, faces, device, closed)
            fields = [p for p, _ in pool_sizes.TARGETS]
            cap = replace(cap, **{p: int(sum(getattr(c, p) for c in per)) for p in fields})
        elif pools != "conservative":
            raise ValueError("pools must be 'learned' or 'conservative'")
        self.peeler = EventPeeler(*arrays, capacities=cap, **options)
        self.rebuilt = False
        if pools == "learned":
            self.peeler.run()
            status, _ = self.peeler._read()
            if status[9] & CAPACITY_BITS:
                del self.peeler
                self.peeler = EventPeeler(*arrays, **options)
                self.rebuilt = True

    @property
    def stream(self):
        return self.peeler.stream

    def run(self):
        self.peeler.run()

    def synchronize(self):
        self.peeler.synchronize()

    def read_status(self):
        return self.peeler.read_status()

    def instance(self, k):
        """Instance k's fragments (rings in its normalised window, local source
        ids, layers) and triangles (uv, world, local sources, layers), read
        from the device."""
        p = self.peeler
        info = p.read_status()
        voff, foff, tx, ty = self.layout[k]
        fend = self.layout[k + 1][1]
        P, T = info["polygons"], info["triangles"]
        src = p.output_sources[:P].cpu().numpy()
        keep = np.nonzero((src >= foff) & (src < fend))[0]
        polys = p.output.numpy(P)
        shift = np.array([tx * self.pitch, ty * self.pitch])
        rings = [polys[i]["v"][: polys[i]["n"]] * self.side - shift for i in keep]
        layers = p.output_layers[:P].cpu().numpy()[keep]
        tsrc = p.triangle_sources[:T].cpu().numpy()
        tkeep = np.nonzero((tsrc >= foff) & (tsrc < fend))[0]
        uv = p.triangles_uv[:T].cpu().numpy()[tkeep] * self.side - shift
        atlas = p.triangles_world[:T].cpu().numpy()[tkeep]
        v, f, cam, st = self.meshes[k]
        u = st[0] + (atlas[..., 0] - tx * self.pitch) * (st[2] - st[0])
        w = st[1] + (atlas[..., 1] - ty * self.pitch) * (st[3] - st[1])
        z = atlas[..., 2]
        if st[4]:
            z = -1.0 / z
            u, w = u * z, w * z
        pc = np.stack([u, w, z], axis=-1)
        world = (pc - cam[:3, 3]) @ np.linalg.inv(cam[:3, :3]).T
        return {
            "rings": rings,
            "sources": src[keep] - foff,
            "layers": layers,
            "triangles_uv": uv,
            "triangles_world": world,
            "triangle_sources": tsrc[tkeep] - foff,
            "triangle_layers": p.triangle_layers[:T].cpu().numpy()[tkeep],
        }
