"""One scene peeled as window tiles.

The route clips sources to the window and treats every edge of a clipped
source as a contour (warp_peel/kernels.py:424-430, 532-533). Peeling with a
tile of the window as the window therefore gives the exact ranks inside that
tile, and the tiles of a partition together give the whole extraction, with
fragments split along the seams. This is how one scene too large for one
workspace is peeled, and how a scene would shard over several devices. Here
the tiles run one after another on one device. Each tile gets only the faces
whose projected bounding box meets it.

Tiles form a quadtree over the normalised window. The tiles of the initial
grid are split until each holds at most ``max_faces`` faces. A tile whose
pools still overflow at the largest ``scales`` entry, or whose edge noding or
subdivision rounds outgrow their fixed bounds, is split again. Pools
grow by ``scales`` as the reference reruns with ``--pool-scale``
(warp_peel/__main__.py:24-47). Tile bounds are dyadic fractions of the window,
so the seam two neighbours share is the same binary64 number in both.
"""

from dataclasses import dataclass, replace

import numpy as np

from .events import EventCapacities, EventPeeler
from .solver import Capacities, DevicePeeler

# per route: peeler, capacities, the pools that grow with the scale, the
# status bits of those pools, and the bits of bounds fixed per edge or per
# tile (edge noding; subdivision rounds), where a smaller tile helps and a
# larger pool does not
ROUTES = {
    "events": (EventPeeler, EventCapacities, ("frontier", "regions", "hits", "vertices", "triangles"),
               4 | 8 | 16 | 32 | 1024, 128 | 2048),
    "layered": (DevicePeeler, Capacities, ("residual", "polygons", "vertices", "triangles", "pieces_per_worker",
                                           "layers"), 2 | 4 | 8 | 16 | 32 | 64, 128),
}


def projected_boxes(vertices, faces, camera, settings):
    """Normalised window bounding box (u0, v0, u1, v1) of every face. A face
    with a vertex at or behind a pinhole's near plane gets an unbounded box."""
    v = np.asarray(vertices, dtype=np.float64).reshape(-1, 3)
    f = np.asarray(faces, dtype=np.int64).reshape(-1, 3)
    c = np.asarray(camera, dtype=np.float64).reshape(4, 4)
    s = np.asarray(settings, dtype=np.float64).reshape(6)
    pc = v @ c[:3, :3].T + c[:3, 3]
    if s[4]:
        front = pc[:, 2] > s[5]
        uv = pc[:, :2] / np.where(front, pc[:, 2], 1.0)[:, None]
        behind = ~front[f].all(axis=1)
    else:
        uv = pc[:, :2]
        behind = np.zeros(len(f), dtype=bool)
    uv = (uv - s[:2]) / (s[2:4] - s[:2])
    tri = uv[f]
    lo, hi = tri.min(axis=1), tri.max(axis=1)
    lo[behind] = -np.inf
    hi[behind] = np.inf
    return np.concatenate([lo, hi], axis=1)


@dataclass
class Tile:
    """One tile's extraction in whole-window normalised coordinates.

    ``rect`` is (u0, v0, u1, v1). Fragments are CSR rings: ``ring_vertices``
    (V, 2) with ``ring_offsets`` (P + 1). ``faces`` lists the global face ids
    peeled in the tile. Sources are global face ids."""

    rect: tuple
    faces: np.ndarray
    scale: int
    info: dict
    ring_vertices: np.ndarray
    ring_offsets: np.ndarray
    sources: np.ndarray
    layers: np.ndarray
    triangles_uv: np.ndarray
    triangles_world: np.ndarray
    triangle_sources: np.ndarray
    triangle_layers: np.ndarray
    device_bytes: int

    def rings(self):
        return np.split(self.ring_vertices, self.ring_offsets[1:-1])


def _to_window(rect, xy):
    u0, v0, u1, v1 = rect
    out = np.array(xy, dtype=np.float64, copy=True)
    out[..., 0] = u0 + out[..., 0] * (u1 - u0)
    out[..., 1] = v0 + out[..., 1] * (v1 - v0)
    return out


class TiledPeeler:
    """Peel host arrays (vertices, faces, camera, settings) tile by tile with
    the peeler of ``route``. ``tiles`` (columns, rows) is the initial grid.

    * ``peel_tiles()`` yields one ``Tile`` at a time, so a caller can consume
      and drop each;
    * ``run()`` keeps them all in ``self.results``;
    * ``fragments()`` and ``triangles()`` concatenate the kept tiles.

    ``route`` is "events" (EventPeeler) or "layered" (DevicePeeler, the
    per-layer clipping route, whose fragments follow only the contours of
    the layer in front, so that it suits scenes whose full contour
    arrangement is dense); ``options`` go to the peeler.
    """

    def __init__(self, vertices, faces, camera, settings, tiles=(1, 1), max_faces=None, device="cuda:0",
                 scales=(1, 4, 16, 64), max_depth=6, route="events", **options):
        if route not in ROUTES:
            raise ValueError(f"route must be one of {sorted(ROUTES)}")
        self.route = route
        self.vertices = np.asarray(vertices, dtype=np.float64).reshape(-1, 3)
        self.faces = np.asarray(faces, dtype=np.int64).reshape(-1, 3)
        self.camera = np.asarray(camera, dtype=np.float64).reshape(4, 4)
        self.settings = np.asarray(settings, dtype=np.float64).reshape(6)
        nx, ny = (int(t) for t in tiles)
        if nx < 1 or ny < 1 or (nx & (nx - 1)) or (ny & (ny - 1)):
            raise ValueError("tile grid sides must be powers of two")
        self.tiles = (nx, ny)
        self.max_faces = max_faces
        self.device = device
        self.scales = tuple(scales)
        self.max_depth = max_depth
        self.options = options
        self.boxes = projected_boxes(self.vertices, self.faces, self.camera, self.settings)
        self.results = []
        self.splits = 0

    def _faces_in(self, rect):
        b = self.boxes
        u0, v0, u1, v1 = rect
        return np.nonzero((b[:, 0] <= u1) & (b[:, 2] >= u0) & (b[:, 1] <= v1) & (b[:, 3] >= v0))[0]

    @staticmethod
    def _quarters(rect):
        u0, v0, u1, v1 = rect
        um, vm = 0.5 * (u0 + u1), 0.5 * (v0 + v1)
        return [(u0, v0, um, vm), (um, v0, u1, vm), (u0, vm, um, v1), (um, vm, u1, v1)]

    def _peel(self, rect, keep):
        """Peel one tile; None when it must be split: its pools overflow at
        every scale, or an edge or the subdivision outgrows its fixed bound."""
        import torch

        from .io import upload

        s = self.settings
        settings = s.copy()
        settings[0] = s[0] + rect[0] * (s[2] - s[0])
        settings[2] = s[0] + rect[2] * (s[2] - s[0])
        settings[1] = s[1] + rect[1] * (s[3] - s[1])
        settings[3] = s[1] + rect[3] * (s[3] - s[1])
        used, local = np.unique(self.faces[keep], return_inverse=True)
        arrays = upload(self.vertices[used], local.reshape(-1, 3), self.camera, settings, device=self.device)
        kind, caps, fields, pool_bits, split_bits = ROUTES[self.route]
        peeler = None
        for scale in self.scales:
            cap = caps.for_mesh(len(keep))
            cap = replace(cap, **{p: getattr(cap, p) * scale for p in fields})
            peeler = kind(*arrays, capacities=cap, **self.options)
            peeler.run()
            bits = int(peeler._read()[0][9])
            if not bits & (pool_bits | split_bits):
                break
            del peeler
            peeler = None
            torch.cuda.synchronize(self.device)
            if not bits & pool_bits:
                break
        if peeler is None:
            return None
        info = peeler.read_status()
        P, T = info["polygons"], info["triangles"]
        n = peeler.output._n[:P]
        xy = peeler.output._xy[:, :P].permute(1, 0, 2)
        mask = torch.arange(xy.shape[1], device=n.device)[None, :] < n[:, None]
        nonempty = (n > 0).cpu().numpy()
        verts = xy[mask].cpu().numpy()
        counts = n.cpu().numpy().astype(np.int64)[nonempty]
        offsets = np.zeros(len(counts) + 1, dtype=np.int64)
        np.cumsum(counts, out=offsets[1:])
        tile = Tile(
            rect=rect,
            faces=keep,
            scale=scale,
            info=info,
            ring_vertices=_to_window(rect, verts),
            ring_offsets=offsets,
            sources=keep[peeler.output_sources[:P].cpu().numpy()[nonempty]],
            layers=peeler.output_layers[:P].cpu().numpy()[nonempty],
            triangles_uv=_to_window(rect, peeler.triangles_uv[:T].cpu().numpy()),
            triangles_world=peeler.triangles_world[:T].cpu().numpy(),
            triangle_sources=keep[peeler.triangle_sources[:T].cpu().numpy()],
            triangle_layers=peeler.triangle_layers[:T].cpu().numpy(),
            device_bytes=peeler.device_bytes,
        )
        del peeler
        torch.cuda.synchronize(self.device)
        return tile

    def peel_tiles(self):
        """Yield the non-empty tiles' extractions, splitting as needed."""
        nx, ny = self.tiles
        stack = [((i / nx, j / ny, (i + 1) / nx, (j + 1) / ny), 0) for j in reversed(range(ny)) for i in reversed(range(nx))]
        while stack:
            rect, depth = stack.pop()
            keep = self._faces_in(rect)
            if len(keep) == 0:
                continue
            if self.max_faces is not None and len(keep) > self.max_faces and depth < self.max_depth:
                self.splits += 1
                stack += [(q, depth + 1) for q in reversed(self._quarters(rect))]
                continue
            tile = self._peel(rect, keep)
            if tile is not None:
                yield tile
            elif depth < self.max_depth:
                self.splits += 1
                stack += [(q, depth + 1) for q in reversed(self._quarters(rect))]
            else:
                raise RuntimeError(f"tile {rect} overflows {self.scales[-1]}x the default pools at the split limit")

    def run(self):
        """Peel every tile and keep the results; returns their status records."""
        self.results = list(self.peel_tiles())
        return [t.info for t in self.results]

    def fragments(self):
        """(ring vertices, ring offsets, sources, layers) of all kept tiles."""
        if not self.results:
            return np.zeros((0, 2)), np.zeros(1, np.int64), np.zeros(0, np.int64), np.zeros(0, np.int32)
        offsets, base = [np.zeros(1, np.int64)], 0
        for t in self.results:
            offsets.append(t.ring_offsets[1:] + base)
            base += t.ring_offsets[-1]
        return (np.concatenate([t.ring_vertices for t in self.results]), np.concatenate(offsets),
                np.concatenate([t.sources for t in self.results]), np.concatenate([t.layers for t in self.results]))

    def triangles(self):
        """(uv, world, sources, layers) of all kept tiles."""
        if not self.results:
            return np.zeros((0, 3, 2)), np.zeros((0, 3, 3)), np.zeros(0, np.int64), np.zeros(0, np.int32)
        return tuple(np.concatenate([getattr(t, k) for t in self.results])
                     for k in ("triangles_uv", "triangles_world", "triangle_sources", "triangle_layers"))

    @property
    def layers(self):
        return max((t.info["layers"] for t in self.results), default=0)
