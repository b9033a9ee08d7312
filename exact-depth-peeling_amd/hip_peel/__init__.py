"""MI355X-native analytic depth peeling: the warp_peel API on gfx950 HIP kernels."""

from .events import EventCapacities, EventPeeler
from .solver import Capacities, DevicePeeler

__all__ = ["Capacities", "DevicePeeler", "EventCapacities", "EventPeeler"]
