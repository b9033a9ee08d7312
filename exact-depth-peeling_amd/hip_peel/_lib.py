"""ctypes binding of libpeel355.so (declarations: include/peel355.h).

The HIP library is the only compute path: if it is missing or fails to load,
every constructor raises. There is no CPU fallback.
"""

import ctypes
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libpeel355.so"

ROUTE_EVENTS, ROUTE_LAYERED = 0, 1
SHARE_REGIONS, PROPAGATE_RANKS, MERGE_REGIONS, FILTER_FP32 = 1, 2, 4, 8
MAX_VERTICES = 32
METRICS_COUNT = 30


class PeelDesc(ctypes.Structure):
    _fields_ = [
        ("vertices", ctypes.c_void_p),
        ("num_vertices", ctypes.c_int64),
        ("faces", ctypes.c_void_p),
        ("num_faces", ctypes.c_int64),
        ("camera", ctypes.c_void_p),
        ("settings", ctypes.c_void_p),
        ("frontier", ctypes.c_int32),
        ("regions", ctypes.c_int32),
        ("hits", ctypes.c_int32),
        ("vertices_cap", ctypes.c_int32),
        ("triangles", ctypes.c_int32),
        ("rounds", ctypes.c_int32),
        ("layer_capacity", ctypes.c_int32),
        ("edge_vertices", ctypes.c_int32),
        ("layers", ctypes.c_int32),
        ("route", ctypes.c_int32),
        ("flags", ctypes.c_int32),
        ("workers", ctypes.c_int32),
        ("pieces", ctypes.c_int32),
    ]


class PeelBuffers(ctypes.Structure):
    _fields_ = [
        ("output_xy", ctypes.c_void_p),
        ("output_n", ctypes.c_void_p),
        ("output_sources", ctypes.c_void_p),
        ("output_layers", ctypes.c_void_p),
        ("regions", ctypes.c_int32),
        ("max_vertices", ctypes.c_int32),
        ("triangles_uv", ctypes.c_void_p),
        ("triangles_world", ctypes.c_void_p),
        ("triangle_sources", ctypes.c_void_p),
        ("triangle_layers", ctypes.c_void_p),
        ("triangles", ctypes.c_int32),
        ("status", ctypes.c_void_p),
        ("metrics", ctypes.c_void_p),
        ("metrics_count", ctypes.c_int32),
        ("canonical", ctypes.c_void_p),
        ("stream", ctypes.c_void_p),
        ("stats", ctypes.c_void_p),
        ("stats_rows", ctypes.c_int32),
        ("device_bytes", ctypes.c_int64),
    ]


class PeelOutputs(ctypes.Structure):
    """Caller-owned outputs of peel_extract_static (include/peel355.h)."""

    _fields_ = [
        ("output_xy", ctypes.c_void_p),
        ("output_n", ctypes.c_void_p),
        ("output_sources", ctypes.c_void_p),
        ("output_layers", ctypes.c_void_p),
        ("triangles_uv", ctypes.c_void_p),
        ("triangles_world", ctypes.c_void_p),
        ("triangle_sources", ctypes.c_void_p),
        ("triangle_layers", ctypes.c_void_p),
        ("status", ctypes.c_void_p),
        ("metrics", ctypes.c_void_p),
    ]


STATIC_SCRATCH_BYTES = 256 << 20


class PeelTuning(ctypes.Structure):
    _fields_ = [
        ("persistent_blocks_per_cu", ctypes.c_int32),
        ("ray_lanes_target", ctypes.c_int32),
        ("heavy_lanes", ctypes.c_int32),
        ("long_edge_cells", ctypes.c_int32),
        ("leaf_bucket_log2", ctypes.c_int32),
        ("max_blocks_per_cu", ctypes.c_int32),
    ]


EXPORTS = (
    "peel_create",
    "peel_warmup",
    "peel_enqueue",
    "peel_capture",
    "peel_launch",
    "peel_get_buffers",
    "peel_read_status",
    "peel_profile",
    "peel_synchronize",
    "peel_destroy",
    "peel_last_error",
    "peel_version",
    "peel_get_tuning",
    "peel_set_tuning",
    "peel_batch_create",
    "peel_batch_launch",
    "peel_batch_synchronize",
    "peel_batch_destroy",
    "pool_mlp_param_count",
    "pool_mlp_partial_count",
    "pool_mlp_forward",
    "pool_mlp_gradient",
    "pool_mlp_adam",
    "peel_extract_static",
    "peel_graph_census",
    "peel_batch_graph_census",
    "peel_ring_test",
)

_lib = None


def load():
    """Load the HIP library; raise if it is absent (no fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python exact-depth-peeling_amd/build.py`"
        )
    lib = ctypes.CDLL(str(LIB_PATH))
    vp = ctypes.c_void_p
    lib.peel_create.argtypes = [ctypes.POINTER(PeelDesc), ctypes.POINTER(vp)]
    for name in ("peel_warmup", "peel_enqueue", "peel_capture", "peel_launch", "peel_synchronize", "peel_destroy"):
        getattr(lib, name).argtypes = [vp]
        getattr(lib, name).restype = ctypes.c_int
    lib.peel_create.restype = ctypes.c_int
    lib.peel_get_buffers.argtypes = [vp, ctypes.POINTER(PeelBuffers)]
    lib.peel_get_buffers.restype = ctypes.c_int
    lib.peel_read_status.argtypes = [vp, vp, vp]
    lib.peel_read_status.restype = ctypes.c_int
    lib.peel_profile.argtypes = [vp, vp, vp, ctypes.c_char_p, ctypes.c_int32]
    lib.peel_profile.restype = ctypes.c_int
    lib.peel_get_tuning.argtypes = [vp, ctypes.POINTER(PeelTuning)]
    lib.peel_get_tuning.restype = ctypes.c_int
    lib.peel_set_tuning.argtypes = [vp, ctypes.POINTER(PeelTuning)]
    lib.peel_set_tuning.restype = ctypes.c_int
    lib.peel_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int32, ctypes.POINTER(vp)]
    lib.peel_batch_create.restype = ctypes.c_int
    for name in ("peel_batch_launch", "peel_batch_synchronize", "peel_batch_destroy"):
        getattr(lib, name).argtypes = [vp]
        getattr(lib, name).restype = ctypes.c_int
    i32, f32 = ctypes.c_int32, ctypes.c_float
    lib.pool_mlp_param_count.argtypes = [i32, i32]
    lib.pool_mlp_param_count.restype = i32
    lib.pool_mlp_partial_count.argtypes = [i32, i32, i32]
    lib.pool_mlp_partial_count.restype = i32
    lib.pool_mlp_forward.argtypes = [vp, i32, i32, vp, i32, vp, vp]
    lib.pool_mlp_forward.restype = ctypes.c_int
    lib.pool_mlp_gradient.argtypes = [vp, i32, i32, vp, vp, i32, f32, vp, vp, vp, vp]
    lib.pool_mlp_gradient.restype = ctypes.c_int
    lib.pool_mlp_adam.argtypes = [vp, vp, vp, vp, i32, f32, f32, f32, f32, f32, f32, vp]
    lib.pool_mlp_adam.restype = ctypes.c_int
    lib.peel_extract_static.argtypes = [ctypes.POINTER(PeelDesc), ctypes.POINTER(PeelOutputs), vp]
    lib.peel_extract_static.restype = ctypes.c_int
    lib.peel_ring_test.argtypes = [i32, vp, i32, vp, vp, vp, vp]
    lib.peel_ring_test.restype = ctypes.c_int
    for name in ("peel_graph_census", "peel_batch_graph_census"):
        getattr(lib, name).argtypes = [vp, vp, i32]
        getattr(lib, name).restype = ctypes.c_int
    lib.peel_last_error.restype = ctypes.c_char_p
    lib.peel_version.restype = ctypes.c_char_p
    _lib = lib
    return lib


GRAPH_NODE_TYPES = ("kernel", "memcpy", "memset", "host", "graph", "empty", "wait_event", "event_record",
                    "semaphore_signal", "semaphore_wait", "mem_alloc", "mem_free", "memcpy_from_symbol",
                    "memcpy_to_symbol", "batch_mem_op")


def census(fn, handle):
    """Node counts of a captured graph by type name (nonzero types only)."""
    counts = (ctypes.c_int32 * len(GRAPH_NODE_TYPES))()
    check(fn(handle, counts, len(GRAPH_NODE_TYPES)))
    return {name: int(c) for name, c in zip(GRAPH_NODE_TYPES, counts) if c}


def check(code):
    if code != 0:
        message = load().peel_last_error().decode()
        if code > 0:
            raise ValueError(message)
        raise RuntimeError(f"HIP error {-code}: {message}")
