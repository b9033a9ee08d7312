"""Scene files in, layer meshes out: the host side of a device extraction.

The on-disk formats are the reference's (warp_peel/io.py:11-72):
- a scene JSON holds ``triangles`` (each ``vertices`` [3][3], optional ``id``),
  ``window`` [xmin, ymin, xmax, ymax] and an optional ``camera``
  (``type`` orthographic|perspective, ``world_to_camera`` 3x4 rows, ``near``);
  every number may be a JSON number or a rational string such as "1/3";
- the extracted mesh is a dict of ``uv`` (T, 3, 2) f64, ``world`` (T, 3, 3) f64,
  ``sources`` (external labels, int64) and ``layers`` (zero-based, int32).

Triangles are kept as an unwelded soup (three vertices per face); the device
weld recovers shared vertices exactly. Nothing here touches geometry beyond
parsing: every stage of the extraction runs in the HIP library.
"""

import json
from dataclasses import dataclass
from fractions import Fraction
from pathlib import Path

import numpy as np
import torch

CAMERA_TYPES = ("orthographic", "perspective")
DEFAULT_NEAR = 0.01


def _real(token):
    """A JSON number, or an exact rational string rounded once to binary64."""
    if isinstance(token, str):
        return float(Fraction(token))
    return float(token)


@dataclass
class Scene:
    """Host arrays in the peelers' input layout plus the external source ids."""

    vertices: np.ndarray  # (3F, 3) f64, face f owns rows 3f..3f+2
    faces: np.ndarray  # (F, 3) i32
    camera: np.ndarray  # (4, 4) f64 world -> camera, last row [0, 0, 0, 1]
    settings: np.ndarray  # (6,) f64 [xmin, ymin, xmax, ymax, perspective, near]
    labels: np.ndarray  # (F,) i64

    def arrays(self):
        return self.vertices, self.faces, self.camera, self.settings

    def check(self):
        if not all(np.isfinite(a).all() for a in (self.vertices, self.camera, self.settings)):
            raise ValueError("Input must be finite")
        xmin, ymin, xmax, ymax, _, near = self.settings
        if not (xmax > xmin and ymax > ymin and near > 0):
            raise ValueError("Invalid image window or near plane")
        if np.unique(self.labels).size != self.labels.size:
            raise ValueError("Source IDs must be unique")
        return self


def _camera_block(spec):
    kind = spec.get("type", "orthographic")
    if kind not in CAMERA_TYPES:
        raise ValueError(f"Unknown camera type {kind!r}")
    matrix = np.eye(4, dtype=np.float64)
    rows = spec.get("world_to_camera")
    if rows is not None:
        matrix[:3, :] = np.array([[_real(x) for x in row] for row in rows], dtype=np.float64).reshape(3, 4)
    return matrix, float(kind == "perspective"), _real(spec.get("near", DEFAULT_NEAR))


def parse_scene(data):
    """Scene JSON (already decoded) -> Scene, validated."""
    records = data["triangles"]
    count = len(records)
    vertices = np.empty((count, 3, 3), dtype=np.float64)
    labels = np.empty(count, dtype=np.int64)
    for f, record in enumerate(records):
        corners = record["vertices"]
        for k in range(3):
            vertices[f, k] = [_real(x) for x in corners[k]]
        labels[f] = record.get("id", f)
    matrix, perspective, near = _camera_block(data.get("camera", {}))
    window = [_real(x) for x in data["window"]]
    settings = np.array(window + [perspective, near], dtype=np.float64)
    faces = np.arange(3 * count, dtype=np.int32).reshape(count, 3)
    return Scene(vertices.reshape(-1, 3), faces, matrix, settings, labels).check()


def scene_arrays(data):
    """Scene JSON -> (vertices, faces, camera, settings, labels) host arrays."""
    s = parse_scene(data)
    return s.vertices, s.faces, s.camera, s.settings, s.labels


_INPUT_LAYOUT = ((torch.float64, (-1, 3)), (torch.int32, (-1, 3)), (torch.float64, (1, 4, 4)), (torch.float64, (6,)))


def upload(vertices, faces, camera, settings, device="cuda:0"):
    """Host arrays -> the four device tensors a peeler takes (one copy each)."""
    out = []
    for host, (dtype, shape) in zip((vertices, faces, camera, settings), _INPUT_LAYOUT):
        t = torch.as_tensor(np.asarray(host)).to(dtype).reshape(shape).contiguous()
        out.append(t.to(device))
    return out


def load_scene(path, device="cuda:0"):
    """Read a scene file: (device arrays, labels, decoded JSON)."""
    data = json.loads(Path(path).read_text())
    scene = parse_scene(data)
    return upload(*scene.arrays(), device=device), scene.labels, data


_MESH_FIELDS = (("uv", "triangles_uv", np.float64, (3, 2)), ("world", "triangles_world", np.float64, (3, 3)),
                ("sources", "triangle_sources", np.int64, ()), ("layers", "triangle_layers", np.int32, ()))


def download_mesh(peeler, labels=None):
    """Synchronise, then copy the valid triangle prefix to the host.

    This is the one explicit readback of an extraction (serialisation only).
    Source indices are mapped through ``labels`` when given."""
    status = peeler.read_status()
    count = int(status["triangles"])
    mesh = {}
    for key, attr, dtype, tail in _MESH_FIELDS:
        if count == 0:
            mesh[key] = np.empty((0, *tail), dtype=dtype)
        else:
            mesh[key] = getattr(peeler, attr)[:count].cpu().numpy().astype(dtype, copy=False)
    if labels is not None and count:
        mesh["sources"] = np.asarray(labels, dtype=np.int64)[mesh["sources"]]
    mesh["status"] = status
    return mesh
