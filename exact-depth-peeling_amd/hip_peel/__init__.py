"""MI355X-native analytic depth peeling: the warp_peel API on gfx950 HIP kernels."""

from .batch import BatchPeeler, MergedBatch
from .events import EventCapacities, EventPeeler
from .solver import Capacities, DevicePeeler
from .tiles import TiledPeeler

__all__ = ["BatchPeeler", "MergedBatch", "Capacities", "DevicePeeler", "EventCapacities", "EventPeeler", "TiledPeeler"]
