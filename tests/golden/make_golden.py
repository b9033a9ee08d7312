#!/usr/bin/env python3
"""Extract the reference's recorded data files into small committed fixtures.

Run once, in a container that has the reference checkout at /root/reference:

    python tests/golden/make_golden.py

Nothing here imports or executes reference code. It copies the reference's own
test inputs and recorded exact outputs (JSON data files) and keeps only the
fields the tests use:

* ``fixtures/*.json``    scene inputs, verbatim (tests/fixtures/*.json upstream,
                          generated by scripts/generate_fixtures.py:168-272).
* ``exact/<case>.json``  exact arrangement regions per (layer, source) and the
                          oracle cells (interior point + exact front-to-back
                          source order) recorded by the reference's CGAL
                          arrangement oracle (results/examples/<case>/arrangement/
                          exact.json; format written by
                          src/triangulate_export.cpp:98-122, cells by
                          src/arrangement_oracle.cpp:55-69).
* ``layers.json``        exact layer count per fixture recorded by the
                          reference's validation suite (results/fast/validation.json,
                          field audit.oracle_layers).
* ``resolved-crossing.json``  the resolver golden (results/fast/resolved-crossing.json).
* ``fast_records.json``  per fixture, the reference's fast CPU method run on the
                          resolved-mesh contract (results/fast/validation.json,
                          records "<case>-mesh"): its mesh topology counts,
                          its rank classification counts and its layer count.
"""

import json
import shutil
from pathlib import Path

REF = Path("/root/reference")
OUT = Path(__file__).resolve().parent

EXAMPLES = [
    "single",
    "holes",
    "closed_object",
    "components",
    "cover_2",
    "crossing",
    "cyclic",
    "perspective_near",
    "perspective_torus",
]


def main():
    (OUT / "fixtures").mkdir(exist_ok=True)
    for path in sorted((REF / "tests/fixtures").glob("*.json")):
        shutil.copyfile(path, OUT / "fixtures" / path.name)
    (OUT / "exact").mkdir(exist_ok=True)
    for case in EXAMPLES:
        data = json.loads((REF / "results/examples" / case / "arrangement/exact.json").read_text())
        slim = {
            "format": data["format"],
            "window": data["window"],
            "camera": data["camera"],
            "layers": [
                {
                    "layer_index": layer["layer_index"],
                    "regions": [
                        {"source_id": r["source_id"], "domain": r["domain"]}
                        for r in layer["regions"]
                    ],
                }
                for layer in data["layers"]
            ],
            "oracle_cells": [
                {"interior_point": c["interior_point"], "depth_order": c["depth_order"]}
                for c in data["oracle_cells"]
            ],
        }
        (OUT / "exact" / f"{case}.json").write_text(json.dumps(slim, separators=(",", ":")) + "\n")
    validation = json.loads((REF / "results/fast/validation.json").read_text())
    layers = {}
    for record in validation["records"]:
        audit = record.get("audit") or {}
        if audit.get("oracle_layers") is not None:
            layers[record["case"]] = audit["oracle_layers"]
    (OUT / "layers.json").write_text(json.dumps(layers, indent=1, sort_keys=True) + "\n")
    fast = {}
    for record in validation["records"]:
        if not record["case"].endswith("-mesh"):
            continue
        cls = record.get("classification") or {}
        fast[record["case"][: -len("-mesh")]] = {
            "layers": record["layers"],
            "topology": {k: v for k, v in record["topology"].items() if k != "resolved_mesh_contract"},
            "classification": {k: cls[k] for k in ("constant_rank_sources", "sources_requiring_clipping",
                                                   "point_query_candidates") if k in cls},
        }
    (OUT / "fast_records.json").write_text(json.dumps(fast, indent=1, sort_keys=True) + "\n")
    shutil.copyfile(REF / "results/fast/resolved-crossing.json", OUT / "resolved-crossing.json")


if __name__ == "__main__":
    main()
