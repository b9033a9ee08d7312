"""Command line: peel a scene JSON on the GPU and write the layer meshes.

    python -m hip_peel SCENE.json --out layers.npz [--method peel|events]
        [--layers N] [--pieces N] [--pool-scale K] [--device cuda:0]

The flags, the NPZ keys (uv, world, sources, layers) and the JSON sidecar
written next to the NPZ (format "warp-peel-approximate-v1", zero-based layers,
the extraction's status counters under "output") are the reference's on-disk
interface (warp_peel/__main__.py:13-77). Exit status 0 on success; a pool or
input error raises the peeler's RuntimeError/ValueError.
"""

import argparse
import dataclasses
import json
import sys
from pathlib import Path

import numpy as np

from . import Capacities, DevicePeeler, EventCapacities, EventPeeler
from .io import download_mesh, load_scene

SIDECAR_FORMAT = "warp-peel-approximate-v1"

# Pools that grow with --pool-scale, per route. The layered route's worker
# count and per-worker piece pool are set separately (--pieces).
_SCALED = {
    "events": ("frontier", "regions", "hits", "vertices", "triangles"),
    "peel": ("residual", "polygons", "vertices", "triangles"),
}


def capacities(method, faces, pool_scale=1, pieces=None):
    """Default capacities for `faces` triangles with the scalable pools times `pool_scale`."""
    base = (EventCapacities if method == "events" else Capacities).for_mesh(faces)
    grown = {name: getattr(base, name) * pool_scale for name in _SCALED[method]}
    if pieces:
        grown["pieces_per_worker"] = pieces
    return dataclasses.replace(base, **grown)


def _arguments(argv):
    cli = argparse.ArgumentParser(prog="python -m hip_peel", description="All-layer depth peeling on one GPU.")
    cli.add_argument("scene", type=Path, help="scene JSON")
    cli.add_argument("--out", type=Path, required=True, help="output .npz (a .json sidecar is written beside it)")
    cli.add_argument("--method", choices=sorted(_SCALED), default="peel",
                     help="peel: per-layer clipping route; events: contour subdivision and shared ranks")
    cli.add_argument("--layers", type=int, default=0, help="layer limit, 0 for all layers")
    cli.add_argument("--pieces", type=int, help="pieces per worker pool (peel only)")
    cli.add_argument("--pool-scale", type=int, default=1, help="multiply the default pool sizes")
    cli.add_argument("--device", default="cuda:0")
    args = cli.parse_args(argv)
    if args.method == "events" and args.pieces is not None:
        cli.error("--pieces applies only to --method peel")
    return args


def run(args):
    arrays, labels, _ = load_scene(args.scene, args.device)
    faces = int(arrays[1].shape[0])
    peeler_type = EventPeeler if args.method == "events" else DevicePeeler
    cap = capacities(args.method, faces, args.pool_scale, args.pieces)
    peeler = peeler_type(*arrays, capacities=cap, layers=args.layers)
    peeler.run()
    mesh = download_mesh(peeler, labels)
    status = mesh.pop("status")
    args.out.parent.mkdir(parents=True, exist_ok=True)
    np.savez_compressed(args.out, **mesh)
    sidecar = {
        "format": SIDECAR_FORMAT,
        "zero_based_layers": True,
        "exact_geometry": False,
        "method": args.method,
        "scene": str(args.scene),
        "output": status,
    }
    args.out.with_suffix(".json").write_text(json.dumps(sidecar, indent=2) + "\n")
    return status


def main(argv=None):
    status = run(_arguments(argv))
    sys.stdout.write(json.dumps(status) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
