"""Many meshes, one graph launch.

``BatchPeeler`` captures the pipelines of several peelers (each with its own
mesh, camera, capacities and outputs) into one HIP graph: the first peeler's
stream forks to every other peeler's stream and joins them again, so one
``run()`` replays every extraction and independent meshes may overlap on the
GPU. The reference peels one mesh per graph; this is an extension for callers
with many small meshes (for example 256 icospheres of 320 to 81,920 faces,
each seen by its own pinhole).

``from_meshes`` sizes each workspace from an estimate and, when an estimate
overflows, rebuilds that workspace with the reference's conservative
capacities and captures the batch again.
"""

import ctypes
from dataclasses import replace

from . import _lib
from .events import EventCapacities, EventPeeler

CAPACITY_BITS = 4 | 8 | 16 | 32 | 128 | 1024 | 2048


def estimated_capacities(n):
    """Tighter pools for closed, star-shaped meshes seen from outside: a few
    layers, few cuts. Overflow is detected and redone with
    ``EventCapacities.for_mesh``."""
    cap = EventCapacities.for_mesh(n)
    m = max(16, n)
    return replace(cap, frontier=6 * m, regions=3 * m, hits=8 * m, vertices=12 * m, triangles=8 * m)


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
    def from_meshes(cls, meshes, device="cuda:0", estimate=True, **options):
        """Peelers for (vertices, faces, camera, settings) host arrays, each
        uploaded to ``device``; with ``estimate`` the pools start from
        ``estimated_capacities`` and an overflowing member is rebuilt with the
        conservative defaults."""
        from .io import upload

        arrays = [upload(*m[:4], device=device) for m in meshes]
        peelers = []
        for a in arrays:
            n = a[1].shape[0]
            cap = estimated_capacities(n) if estimate else EventCapacities.for_mesh(n)
            peelers.append(EventPeeler(*a, capacities=cap, **options))
        batch = cls(peelers)
        if estimate:
            batch.run()
            batch.synchronize()
            redo = [i for i, p in enumerate(batch.peelers) if int(p._read()[0][9]) & CAPACITY_BITS]
            if redo:
                batch.close()
                for i in redo:
                    n = arrays[i][1].shape[0]
                    peelers[i] = EventPeeler(*arrays[i], capacities=EventCapacities.for_mesh(n), **options)
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
