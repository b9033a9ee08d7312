"""Device-resident analytic depth peeling on MI355X: workspace, capacities and
the DevicePeeler host interface.

Mirrors warp_peel/solver.py (alecjacobson/exact-depth-peeling): the same
constructor arguments, capacity record, warmup/enqueue/capture/run/read_status
lifecycle, status slots and error messages. The arrays are torch tensors on a
ROCm device instead of Warp arrays; every stage runs in the HIP library
(libpeel355.so) on the workspace's own stream.
"""

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib

ERROR_LABELS = {
    1: "polygon vertices",
    2: "per-worker pieces",
    4: "residual pool",
    8: "output polygon pool",
    16: "boundary vertex pool",
    32: "triangle pool",
    64: "layer limit",
    128: "edge noding capacity",
    256: "invalid input",
    4096: "device wait limit",
}


@dataclass(frozen=True)
class Capacities:
    """warp_peel/solver.py:10-31: residual pieces per layer, output polygons,
    boundary vertices, triangles, worker threads, pieces per worker pool,
    layers and edge-noding candidates."""

    residual: int
    polygons: int
    vertices: int
    triangles: int
    workers: int = 4096
    pieces_per_worker: int = 512
    layers: int = 128
    edge_vertices: int = 128

    @classmethod
    def for_mesh(cls, triangles: int):
        n = max(triangles, 16)
        return cls(
            4 * n,
            16 * n,
            96 * n,
            64 * n,
            workers=min(4096, n),
            pieces_per_worker=512 if n <= 8192 else 2048,
        )


class _View:
    def __init__(self, ptr, shape, typestr):
        self.__cuda_array_interface__ = {
            "shape": tuple(int(s) for s in shape),
            "typestr": typestr,
            "data": (int(ptr), False),
            "version": 3,
            "strides": None,
        }


def device_view(ptr, shape, dtype, device):
    typestr = {torch.float64: "<f8", torch.int32: "<i4", torch.int64: "<i8"}[dtype]
    return torch.as_tensor(_View(ptr, shape, typestr), device=device)


class PolygonPool:
    """Output polygons with the reference's record layout: ``numpy()`` returns a
    structured array with fields ``v`` (32, 2) float64 and ``n`` int32
    (warp_peel/kernels.py:119-122, read by tests/gpu/test_peeler.py:24-34)."""

    dtype = np.dtype([("v", np.float64, (_lib.MAX_VERTICES, 2)), ("n", np.int32)])

    def __init__(self, xy, n):
        self._xy, self._n = xy, n  # xy: (32, R, 2) vertex-major device view

    @property
    def size(self):
        return self._n.shape[0]

    def numpy(self, count=None):
        count = self.size if count is None else int(count)
        out = np.zeros(count, dtype=self.dtype)
        if count:
            out["n"] = self._n[:count].cpu().numpy()
            out["v"] = self._xy[:, :count].permute(1, 0, 2).cpu().numpy()
        return out

    def __len__(self):
        return self.size


def _check_input(t, dtype, shape_tail, name, device):
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name} must be a torch tensor on a ROCm device")
    if t.device != device or t.dtype != dtype or tuple(t.shape[1:]) != shape_tail or not t.is_contiguous():
        raise ValueError(
            f"Expected a contiguous {dtype} tensor of shape (N, {', '.join(map(str, shape_tail))}) on {device} for {name}"
        )


class _Workspace:
    """Shared workspace lifecycle; subclasses choose the route and capacities."""

    _route = _lib.ROUTE_EVENTS

    def _open(self, vertices, faces, camera, settings, caps, layers, flags, workers=1, pieces=1):
        lib = _lib.load()
        if not isinstance(vertices, torch.Tensor) or vertices.device.type != "cuda":
            raise ValueError("Peelers require ROCm device tensors (torch.device('cuda'))")
        self.device = vertices.device
        if faces.ndim == 1 and faces.numel() == 0:
            faces = faces.reshape(0, 3)
        _check_input(vertices, torch.float64, (3,), "vertices", self.device)
        _check_input(faces, torch.int32, (3,), "faces", self.device)
        if camera.ndim == 2:
            camera = camera.reshape(1, 4, 4)
        _check_input(camera, torch.float64, (4, 4), "camera", self.device)
        if camera.shape[0] != 1 or settings.numel() != 6:
            raise ValueError("Expected one camera matrix and six camera settings")
        if settings.dtype != torch.float64 or settings.device != self.device or not settings.is_contiguous():
            raise ValueError(f"Expected a contiguous float64 settings tensor on {self.device}")
        self.vertices, self.faces, self.camera, self.settings = vertices, faces, camera, settings
        self.n = faces.shape[0]
        self.nv = vertices.shape[0]
        self.flags = flags
        desc = _lib.PeelDesc(
            vertices.data_ptr() if self.nv else None,
            self.nv,
            faces.data_ptr() if self.n else None,
            self.n,
            camera.data_ptr(),
            settings.data_ptr(),
            *caps,
            layers,
            self._route,
            flags,
            workers,
            pieces,
        )
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(lib.peel_create(ctypes.byref(desc), ctypes.byref(handle)))
        self._handle = handle
        self._lib = lib
        b = _lib.PeelBuffers()
        _lib.check(lib.peel_get_buffers(handle, ctypes.byref(b)))
        self._buffers = b
        dev = self.device
        R, T = b.regions, b.triangles
        self._output_xy = device_view(b.output_xy, (_lib.MAX_VERTICES, R, 2), torch.float64, dev)
        self._output_n = device_view(b.output_n, (R,), torch.int32, dev)
        self.output = PolygonPool(self._output_xy, self._output_n)
        self.output_sources = device_view(b.output_sources, (R,), torch.int32, dev)
        self.output_layers = device_view(b.output_layers, (R,), torch.int32, dev)
        self.triangles_uv = device_view(b.triangles_uv, (T, 3, 2), torch.float64, dev)
        self.triangles_world = device_view(b.triangles_world, (T, 3, 3), torch.float64, dev)
        self.triangle_sources = device_view(b.triangle_sources, (T,), torch.int32, dev)
        self.triangle_layers = device_view(b.triangle_layers, (T,), torch.int32, dev)
        self.status = device_view(b.status, (16,), torch.int32, dev)
        self.metrics = device_view(b.metrics, (b.metrics_count,), torch.int64, dev)
        self.canonical = device_view(b.canonical, (max(1, self.nv),), torch.int32, dev)
        self.stats = device_view(b.stats, (b.stats_rows, 3), torch.int64, dev)
        self.device_bytes = int(b.device_bytes)
        self.stream = torch.cuda.ExternalStream(b.stream, device=dev)
        self.graph_captured = False
        self.warmed = False
        self.limit = layers
        self.all_layers = int(layers == 0)

    def __del__(self):
        handle = getattr(self, "_handle", None)
        if handle is not None and handle.value:
            try:
                self._lib.peel_destroy(handle)
            except Exception:
                pass
            self._handle = None

    # -- lifecycle (warp_peel/solver.py:447-519) ---------------------------------
    def _order_inputs(self):
        # Inputs written on the caller's current stream must be visible to the
        # workspace stream (asynchronous, no host synchronisation).
        self.stream.wait_stream(torch.cuda.current_stream(self.device))

    def warmup(self):
        """Run every stage once, eagerly, before any capture."""
        if torch.cuda.is_current_stream_capturing():
            raise RuntimeError("Warm-up must run before graph capture")
        self._order_inputs()
        _lib.check(self._lib.peel_warmup(self._handle))
        self.warmed = True

    def enqueue(self):
        """Record the pipeline on ``self.stream``; valid inside an external
        capture on that stream. No host reads, synchronisation or allocation."""
        if not self.warmed:
            raise RuntimeError("Call warmup(), then enqueue() inside a capture on peeler.stream")
        _lib.check(self._lib.peel_enqueue(self._handle))

    def capture(self):
        if not self.warmed:
            self.warmup()
        _lib.check(self._lib.peel_capture(self._handle))
        self.graph_captured = True
        return self

    @property
    def graph(self):
        return self if self.graph_captured else None

    def run(self):
        """Asynchronous graph replay on ``self.stream``."""
        if not self.graph_captured:
            self.capture()
        self._order_inputs()
        _lib.check(self._lib.peel_launch(self._handle))

    def profile(self):
        """Eager run with hipEvents between stages: {stage: ms} in pipeline order."""
        if torch.cuda.is_current_stream_capturing():
            raise RuntimeError("profile() runs eagerly; it cannot be captured")
        self._order_inputs()
        count = ctypes.c_int32(64)
        ms = (ctypes.c_float * 64)()
        names = ctypes.create_string_buffer(4096)
        _lib.check(self._lib.peel_profile(self._handle, ctypes.byref(count), ms, names, 4096))
        keys = names.value.decode().split(",")
        return dict(zip(keys, [float(ms[i]) for i in range(count.value)]))

    @property
    def tuning(self):
        """Launch tuning (``PeelTuning`` fields as a dict, including the
        read-only ``max_blocks_per_cu``)."""
        t = _lib.PeelTuning()
        _lib.check(self._lib.peel_get_tuning(self._handle, ctypes.byref(t)))
        return {name: int(getattr(t, name)) for name, _ in t._fields_}

    def set_tuning(self, **values):
        """Change launch tuning; the graph is captured again at the next run.
        Outputs do not depend on it."""
        if torch.cuda.is_current_stream_capturing():
            raise RuntimeError("Tuning cannot change during a capture")
        current = self.tuning
        unknown = set(values) - set(current) | ({"max_blocks_per_cu"} & set(values))
        if unknown:
            raise ValueError(f"Unknown or read-only tuning fields: {sorted(unknown)}")
        current.update(values)
        t = _lib.PeelTuning(**current)
        _lib.check(self._lib.peel_set_tuning(self._handle, ctypes.byref(t)))
        self.graph_captured = False

    def graph_nodes(self):
        """Node counts of the captured graph by type (``kernel``, ``empty``,
        ``event_record`` ...): the reference checks that its graph holds no
        copies or allocations (tests/gpu/test_peeler.py:197-202)."""
        if not self.graph_captured:
            self.capture()
        return _lib.census(self._lib.peel_graph_census, self._handle)

    def synchronize(self):
        _lib.check(self._lib.peel_synchronize(self._handle))

    def _read(self):
        status = np.zeros(16, dtype=np.int32)
        metrics = np.zeros(self._buffers.metrics_count, dtype=np.int64)
        _lib.check(
            self._lib.peel_read_status(
                self._handle, status.ctypes.data_as(ctypes.c_void_p), metrics.ctypes.data_as(ctypes.c_void_p)
            )
        )
        return status, metrics

    def read_status(self):
        """Synchronise once, copy status and metrics once, and decode them
        (solver.py:488-519)."""
        values, metrics = self._read()
        self._last_metrics = metrics
        return self._decode(values, metrics)

    def _decode(self, values, metrics):
        if values[9]:
            raise RuntimeError(
                "GPU extraction failed: " + ", ".join(v for b, v in ERROR_LABELS.items() if values[9] & b)
            )
        return {
            "residual_patches": int(values[0]),
            "polygons": int(values[2]),
            "triangles": int(values[3]),
            "layers": int(values[4]),
            "peeling_passes": int(values[5]),
            "boundary_vertices": int(values[6]),
            "depth_ties": int(values[8]),
            "error": int(values[9]),
            "max_hidden_pieces": int(values[10]),
            "peak_residual_patches": int(values[12]),
        }


class DevicePeeler(_Workspace):
    """warp_peel.DevicePeeler interface (solver.py:92-519): whole faces with a
    certain centroid rank are emitted directly, every other face is clipped
    layer by layer against the sources and residual pieces that may cover it.

    Vertices float64 (V, 3), faces int32 (F, 3), camera float64 (1, 4, 4) affine
    world-to-camera, settings float64 (6,) = xmin, ymin, xmax, ymax,
    perspective (0/1), near. Outputs use dense source indices into faces;
    ``stats`` holds per layer the targets, candidate tests and fully hidden
    targets.
    """

    _route = _lib.ROUTE_LAYERED

    def __init__(self, vertices, faces, camera, settings, *, capacities=None, layers=0):
        cap = capacities or Capacities.for_mesh(faces.shape[0] if faces.ndim == 2 else faces.numel() // 3)
        if min(vars(cap).values()) < 1 or cap.edge_vertices > 128:
            raise ValueError("Capacities must be positive; edge_vertices must be <= 128")
        if layers < 0 or layers > cap.layers:
            raise ValueError("layers must be zero (all), or fit in the layer capacity")
        self.cap = cap
        caps = (
            cap.residual,
            cap.polygons,
            1,
            cap.vertices,
            cap.triangles,
            1,
            cap.layers,
            cap.edge_vertices,
        )
        self._open(vertices, faces, camera, settings, caps, layers, 0, cap.workers, cap.pieces_per_worker)
