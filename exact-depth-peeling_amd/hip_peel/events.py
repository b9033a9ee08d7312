"""EventPeeler: all-layer contour subdivision followed by shared sorted ray
ranks (warp_peel/events.py:36-845), on gfx950 HIP kernels.

Same constructor switches, capacity record, metrics names and error messages
as the reference. The subdivision rounds of the reference's conditional while
node run here as one persistent work-queue kernel inside the captured graph.
"""

from dataclasses import dataclass

from . import _lib
from .solver import _Workspace

METRIC_NAMES = (
    "target_visits",
    "contour_tests",
    "splits",
    "ray_groups",
    "ray_candidate_tests",
    "ray_hits",
    "rounds",
    "unique_contours",
    "heavy_ray_groups",
    "beam_attempts",
    "certified_beams",
    "beam_candidate_tests",
    "heavy_beams",
    "rank_components",
    "rank_connections",
    "connection_contour_tests",
    "merged_polygons",
    "matched_patch_edges",
    "edge_hash_misses",
)
EXTRA_METRICS = ("filtered_containment", "fp64_containment", "frontier_slots", "max_edge_vertices",
                 "unique_boundary_vertices", "filtered_straddle", "fp64_straddle", "filtered_front_segment",
                 "fp64_front_segment", "filtered_depth_order", "fp64_depth_order")
FILTER_CLASSES = ("containment", "straddle", "front_segment", "depth_order")


@dataclass(frozen=True)
class EventCapacities:
    """warp_peel/events.py:19-33. ``frontier`` counts frontier slots: every
    contour cut takes two (the reference counts live patches per round), so the
    default is 16 n instead of 4 n."""

    frontier: int
    regions: int
    hits: int
    vertices: int
    triangles: int
    rounds: int = 4096
    layers: int = 4096
    edge_vertices: int = 128

    @classmethod
    def for_mesh(cls, n):
        n = max(16, n)
        return cls(16 * n, 8 * n, 32 * n, 32 * n, 32 * n, layers=n)


class EventPeeler(_Workspace):
    """warp_peel.EventPeeler interface. Requires resolved surface intersections.

    ``filter_fp32`` (an extension; the reference has no such switch) decides
    four predicate classes in binary32 under static forward error bounds and
    evaluates only uncertain tests in binary64: the rays' point-in-polygon
    orientations (``containment``), the subdivision's straddle signs
    (``straddle``) and ring-edge tests of ``front_segment``, and the depth
    comparisons that place a patch in its ray's hit order (``depth_order``).
    The outputs are bit-identical to the binary64 route; ``read_status``
    reports ``filtered_<class>`` and ``fp64_<class>`` per class and their sums
    ``filtered_predicates`` and ``fp64_fallbacks``."""

    _route = _lib.ROUTE_EVENTS

    def __init__(
        self,
        vertices,
        faces,
        camera,
        settings,
        *,
        capacities=None,
        layers=0,
        share_regions=True,
        propagate_ranks=True,
        merge_regions=False,
        filter_fp32=False,
    ):
        n = faces.shape[0] if faces.ndim == 2 else faces.numel() // 3
        ec = capacities or EventCapacities.for_mesh(n)
        if min(vars(ec).values()) < 1:
            raise ValueError("Event capacities must be positive")
        if ec.edge_vertices > 128:
            raise ValueError("Capacities must be positive; edge_vertices must be <= 128")
        if layers < 0 or layers > ec.layers:
            raise ValueError("layers must be zero (all), or fit in the layer capacity")
        self.event_cap = ec
        self.share_regions = share_regions
        self.propagate_ranks = propagate_ranks
        self.merge_regions = merge_regions
        self.filter_fp32 = filter_fp32
        flags = (
            (_lib.SHARE_REGIONS if share_regions else 0)
            | (_lib.PROPAGATE_RANKS if propagate_ranks else 0)
            | (_lib.MERGE_REGIONS if merge_regions else 0)
            | (_lib.FILTER_FP32 if filter_fp32 else 0)
        )
        caps = (
            ec.frontier,
            ec.regions,
            ec.hits,
            ec.vertices,
            ec.triangles,
            ec.rounds,
            ec.layers,
            ec.edge_vertices,
        )
        self._open(vertices, faces, camera, settings, caps, layers, flags)

    def _decode(self, status, metrics):
        if status[9] & 1024:
            raise RuntimeError("GPU extraction failed: shared ray hit capacity")
        if status[9] & 2048:
            raise RuntimeError("GPU extraction failed: subdivision round capacity")
        info = super()._decode(status, metrics)
        for name in ("peeling_passes", "max_hidden_pieces"):
            info.pop(name)
        info.update(zip(METRIC_NAMES, map(int, metrics)))
        info.update(zip(EXTRA_METRICS, map(int, metrics[19:30])))
        info["filtered_predicates"] = sum(info[f"filtered_{c}"] for c in FILTER_CLASSES)
        info["fp64_fallbacks"] = sum(info[f"fp64_{c}"] for c in FILTER_CLASSES)
        return info
