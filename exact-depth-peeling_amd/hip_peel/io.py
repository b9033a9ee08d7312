"""Host serialisation adapters around a device-resident extraction
(warp_peel/io.py:11-72): JSON scene upload and explicit NPZ download."""

import json
from fractions import Fraction
from pathlib import Path

import numpy as np
import torch


def scene_arrays(data):
    """Scene JSON -> host arrays (vertices (3N, 3) f64, faces (N, 3) i32,
    camera (4, 4) f64, settings (6,) f64, labels (N,) i64). Faces are an
    unwelded soup; rational strings become binary64 (warp_peel/io.py:13-41)."""
    number = lambda x: float(Fraction(x)) if isinstance(x, str) else float(x)
    triangles = data["triangles"]
    vertices = np.array(
        [[[number(x) for x in p] for p in t["vertices"]] for t in triangles], dtype=np.float64
    ).reshape((-1, 3))
    faces = np.arange(3 * len(triangles), dtype=np.int32).reshape((-1, 3))
    camera = np.eye(4, dtype=np.float64)
    cam = data.get("camera", {})
    if cam.get("type", "orthographic") not in ("orthographic", "perspective"):
        raise ValueError("Unknown camera type")
    if "world_to_camera" in cam:
        camera[:3] = [[number(x) for x in row] for row in cam["world_to_camera"]]
    settings = np.array(
        [number(x) for x in data["window"]]
        + [int(cam.get("type") == "perspective"), number(cam.get("near", 0.01))],
        dtype=np.float64,
    )
    if not np.isfinite(vertices).all() or not np.isfinite(camera).all() or not np.isfinite(settings).all():
        raise ValueError("Input must be finite")
    if settings[2] <= settings[0] or settings[3] <= settings[1] or settings[5] <= 0:
        raise ValueError("Invalid image window or near plane")
    labels = np.array([t.get("id", i) for i, t in enumerate(triangles)], dtype=np.int64)
    if len(set(labels.tolist())) != len(labels):
        raise ValueError("Source IDs must be unique")
    return vertices, faces, camera, settings, labels


def upload(vertices, faces, camera, settings, device="cuda:0"):
    """Host arrays -> device tensors in the peelers' input layout."""
    return [
        torch.as_tensor(np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 3), device=device),
        torch.as_tensor(np.ascontiguousarray(faces, dtype=np.int32).reshape(-1, 3), device=device),
        torch.as_tensor(np.ascontiguousarray(camera, dtype=np.float64).reshape(1, 4, 4), device=device),
        torch.as_tensor(np.ascontiguousarray(settings, dtype=np.float64).reshape(6), device=device),
    ]


def load_scene(path, device="cuda:0"):
    """Returns (arrays, labels, data) like warp_peel.io.load_scene."""
    data = json.loads(Path(path).read_text())
    vertices, faces, camera, settings, labels = scene_arrays(data)
    return upload(vertices, faces, camera, settings, device), labels, data


def download_mesh(peeler, labels=None):
    """Explicit host transfer for serialisation; not part of the device pipeline."""
    info = peeler.read_status()
    n = info["triangles"]
    if n == 0:
        return {
            "uv": np.empty((0, 3, 2)),
            "world": np.empty((0, 3, 3)),
            "sources": np.empty(0, dtype=np.int64),
            "layers": np.empty(0, dtype=np.int32),
            "status": info,
        }
    sources = peeler.triangle_sources[:n].cpu().numpy()
    if labels is not None:
        sources = np.asarray(labels)[sources]
    return {
        "uv": peeler.triangles_uv[:n].cpu().numpy(),
        "world": peeler.triangles_world[:n].cpu().numpy(),
        "sources": sources,
        "layers": peeler.triangle_layers[:n].cpu().numpy(),
        "status": info,
    }
