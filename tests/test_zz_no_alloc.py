"""peel_extract_static, the entry point that allocates nothing: its outputs
equal those of a workspace created by peel_create (same kernels, pools carved
from a static device array instead of hipMalloc), on two calls in a row, the
second with a smaller input, so the second call runs over what the first
left in the array."""

import ctypes

import numpy as np
import pytest

from peel_checks import fixture_path


def _static_desc(arrays, caps, route=0, flags=0, layers=0, workers=1, pieces=1):
    from hip_peel import _lib

    v, f, c, s = arrays
    return _lib.PeelDesc(v.data_ptr() if v.shape[0] else None, v.shape[0], f.data_ptr() if f.shape[0] else None,
                         f.shape[0], c.data_ptr(), s.data_ptr(), *caps, layers, route, flags, workers, pieces)


def _event_caps(n):
    from hip_peel import EventCapacities

    e = EventCapacities.for_mesh(n)
    return (e.frontier, e.regions, e.hits, e.vertices, e.triangles, e.rounds, e.layers, e.edge_vertices)


def _run_static(arrays, caps, flags):
    import torch

    from hip_peel import _lib

    lib = _lib.load()
    dev = arrays[0].device
    R, T = caps[1], caps[4]
    out = {
        "xy": torch.full((_lib.MAX_VERTICES, R, 2), float("nan"), dtype=torch.float64, device=dev),
        "n": torch.full((R,), -7, dtype=torch.int32, device=dev),
        "src": torch.full((R,), -7, dtype=torch.int32, device=dev),
        "layer": torch.full((R,), -7, dtype=torch.int32, device=dev),
        "uv": torch.full((T, 3, 2), float("nan"), dtype=torch.float64, device=dev),
        "world": torch.full((T, 3, 3), float("nan"), dtype=torch.float64, device=dev),
        "tsrc": torch.full((T,), -7, dtype=torch.int32, device=dev),
        "tlayer": torch.full((T,), -7, dtype=torch.int32, device=dev),
        "status": torch.full((16,), -7, dtype=torch.int32, device=dev),
        "metrics": torch.full((_lib.METRICS_COUNT,), -7, dtype=torch.int64, device=dev),
    }
    o = _lib.PeelOutputs(*(out[k].data_ptr() for k in ("xy", "n", "src", "layer", "uv", "world", "tsrc", "tlayer",
                                                         "status", "metrics")))
    desc = _static_desc(arrays, caps, flags=flags)
    stream = torch.cuda.current_stream(dev)
    print("peel_extract_static: faces", arrays[1].shape[0], flush=True)
    _lib.check(lib.peel_extract_static(ctypes.byref(desc), ctypes.byref(o), ctypes.c_void_p(stream.cuda_stream)))
    stream.synchronize()
    return out


def _records(xy, n, src, layer, count):
    """Multiset of (source, layer, ring bytes) of the valid polygon prefix."""
    xy = xy[:, :count].permute(1, 0, 2).cpu().numpy()
    n, src, layer = n[:count].cpu().numpy(), src[:count].cpu().numpy(), layer[:count].cpu().numpy()
    return sorted((int(s), int(l), xy[i, : n[i]].tobytes()) for i, (s, l) in enumerate(zip(src, layer)))


def _triangles(uv, world, src, layer, count):
    uv, world = uv[:count].cpu().numpy(), world[:count].cpu().numpy()
    src, layer = src[:count].cpu().numpy(), layer[:count].cpu().numpy()
    return sorted((int(s), int(l), uv[i].tobytes(), world[i].tobytes()) for i, (s, l) in enumerate(zip(src, layer)))


def _compare(arrays, flags):
    from hip_peel import EventPeeler, _lib

    caps = _event_caps(arrays[1].shape[0])
    got = _run_static(arrays, caps, flags)
    status = got["status"].cpu().numpy()
    assert status[9] == 0, status
    ref = EventPeeler(*arrays, share_regions=bool(flags & _lib.SHARE_REGIONS),
                      propagate_ranks=bool(flags & _lib.PROPAGATE_RANKS))
    ref.run()
    info = ref.read_status()
    P, T = info["polygons"], info["triangles"]
    assert (int(status[2]), int(status[3]), int(status[4])) == (P, T, info["layers"])
    assert _records(got["xy"], got["n"], got["src"], got["layer"], P) == _records(
        ref._output_xy, ref._output_n, ref.output_sources, ref.output_layers, P)
    assert _triangles(got["uv"], got["world"], got["tsrc"], got["tlayer"], T) == _triangles(
        ref.triangles_uv, ref.triangles_world, ref.triangle_sources, ref.triangle_layers, T)
    # every status slot but the work queue's peak of live tasks (12), which
    # depends on scheduling; of the metrics, the counters that define the
    # partition (ray counts depend on which patch leads a shared rectangle,
    # the lowest output slot, which the order of appends decides)
    keep = [k for k in range(16) if k != 12]
    np.testing.assert_array_equal(status[keep], ref.status.cpu().numpy()[keep])
    defining = [1, 2, 7, 13, 21, 23]  # contour tests, splits, unique contours, components,
    # frontier slots, distinct boundary vertices
    np.testing.assert_array_equal(got["metrics"].cpu().numpy()[defining], ref.metrics.cpu().numpy()[defining])
    return P, T


@pytest.mark.gpu
def test_static_entry_matches_workspace_twice():
    from hip_peel import _lib, scenes
    from hip_peel.io import load_scene, upload

    flags = _lib.SHARE_REGIONS | _lib.PROPAGATE_RANKS
    big = upload(*scenes.torus_knot(seed=3, segments=48, sides=24))
    small = load_scene(fixture_path("perspective_torus"), "cuda:0")[0]
    first = _compare(big, flags)
    second = _compare(small, flags)
    assert first[0] > second[0] > 0


@pytest.mark.gpu
def test_static_entry_rejects_oversized_descriptors():
    import torch

    from hip_peel import _lib, scenes
    from hip_peel.io import upload

    arrays = upload(*scenes.torus_knot(seed=3, segments=48, sides=24))
    caps = list(_event_caps(arrays[1].shape[0]))
    caps[0] = 1 << 24  # frontier slots far beyond the static scratch
    lib = _lib.load()
    dummy = torch.zeros(16, dtype=torch.int64, device="cuda")
    o = _lib.PeelOutputs(*([dummy.data_ptr()] * 10))
    desc = _static_desc(arrays, caps)
    code = lib.peel_extract_static(ctypes.byref(desc), ctypes.byref(o), None)
    assert code == 2
    assert "bytes of scratch" in lib.peel_last_error().decode()
    assert lib.peel_extract_static(ctypes.byref(desc), None, None) == 1
