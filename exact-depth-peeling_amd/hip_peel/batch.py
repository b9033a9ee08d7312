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
            caps = model.capacities_many(features, faces, device)
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
