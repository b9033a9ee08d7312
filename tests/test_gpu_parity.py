"""Bit-level parity of the device route with the CPU FP64 restatement
(oracle/cpu_event.cpp), plus exact-rank audits at full procedural sizes.

Both implement the reference's event route with the same Morton hierarchies,
the same candidate order and the same binary64 operation order (no FMA
contraction on either side), so the multiset of output patches (source, layer,
ring) and of triangles must agree bit for bit; only the output order differs
(atomic appends, warp_peel/README.md:85-87).
"""

import numpy as np
import pytest

import oracle
from peel_checks import check_extraction, fixture_path
from test_gpu_fixtures import CASES


def canonical_patches(rings, sources, layers):
    rows = []
    for r, s, l in zip(rings, sources, layers):
        if len(r):
            rows.append((int(s), int(l), tuple(np.asarray(r).ravel().tolist())))
    return sorted(rows)


def canonical_triangles(uv, world, src, layer):
    rows = [
        (int(src[k]), int(layer[k]), tuple(uv[k].ravel().tolist()), tuple(world[k].ravel().tolist()))
        for k in range(len(src))
    ]
    return sorted(rows)


def gpu_outputs(p):
    info = p.read_status()
    P, T = info["polygons"], info["triangles"]
    polys = p.output.numpy(P)
    rings = [x["v"][: x["n"]] for x in polys]
    return (
        info,
        rings,
        p.output_sources[:P].cpu().numpy(),
        p.output_layers[:P].cpu().numpy(),
        p.triangles_uv[:T].cpu().numpy(),
        p.triangles_world[:T].cpu().numpy(),
        p.triangle_sources[:T].cpu().numpy(),
        p.triangle_layers[:T].cpu().numpy(),
    )


def compare(host, threads=8, propagate=True, share=True, filter_fp32=False):
    import torch  # noqa: F401

    from hip_peel import EventPeeler
    from hip_peel.io import upload

    p = EventPeeler(*upload(*host[:4]), propagate_ranks=propagate, share_regions=share,
                    filter_fp32=filter_fp32)
    p.run()
    info, rings, srcs, lays, uv, world, tsrc, tlay = gpu_outputs(p)
    cpu = oracle.cpu_event_peel(*host[:4], threads=threads, propagate_ranks=propagate, share_regions=share)
    assert cpu.counts["error"] == 0
    # ray counters depend on which patch represents a component; the patch
    # and triangle sets do not
    keys = ["polygons", "triangles", "layers", "splits", "unique_contours", "contour_tests",
            "rank_components", "rank_connections"]
    if not propagate:
        # without propagation every patch samples itself, so the ray groups
        # and certified cells are fixed by the patch set
        keys += ["ray_groups", "ray_hits", "beam_attempts", "certified_beams"]
    for key in keys:
        assert info[key] == cpu.counts[key], (key, info[key], cpu.counts[key])
    assert canonical_patches(rings, srcs, lays) == canonical_patches(cpu.rings, cpu.sources, cpu.layers)
    assert canonical_triangles(uv, world, tsrc, tlay) == canonical_triangles(
        cpu.tri_uv, cpu.tri_world, cpu.tri_src, cpu.tri_layer)
    return p, info


@pytest.mark.gpu
@pytest.mark.parametrize("share", [True, False])
@pytest.mark.parametrize("propagate", [True, False])
@pytest.mark.parametrize("name", CASES)
def test_fixture_bit_parity_with_cpu_route(name, propagate, share):
    print(f"[parity] {name} propagate={propagate} share={share}", flush=True)
    compare(oracle.load_scene(fixture_path(name))[:4], propagate=propagate, share=share)


@pytest.mark.gpu
@pytest.mark.parametrize("grid", [(32, 16), (64, 32), (256, 128)])
def test_torus_knot_bit_parity_and_exact_audit(grid):
    from hip_peel import scenes

    print(f"[parity] torus knot {grid}", flush=True)
    host = scenes.torus_knot(segments=grid[0], sides=grid[1])
    p, info = compare(host)
    # exact ranks at every fragment's interior points and a point grid
    report = check_extraction(p, host, grid=32, triangles=grid[0] <= 64)
    assert report["fragment_rank_mismatches"] == 0


@pytest.mark.gpu
def test_stack_65536_closed_form():
    from hip_peel import EventPeeler, scenes
    from hip_peel.io import upload

    print("[parity] stack 65536", flush=True)
    p = EventPeeler(*upload(*scenes.stack(65536)))
    p.run()
    info = p.read_status()
    n = 65536
    assert info["polygons"] == n and info["layers"] == n and info["triangles"] == 3 * n
    assert info["splits"] == 0 and info["unique_contours"] == 3
    assert info["ray_groups"] == 1 and info["ray_hits"] == n and info["ray_candidate_tests"] == n
    np.testing.assert_array_equal(p.output_layers[:n].cpu().numpy(), p.output_sources[:n].cpu().numpy())


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_fixture_fp32_filter_is_bit_identical(name):
    """The binary32 filters (ray containment, subdivision straddle signs,
    front_segment edge tests, rank depth order) give the binary64 route's
    output bit for bit (checked against the CPU restatement) and count their
    decisions."""
    print(f"[parity] {name} fp32 filter", flush=True)
    p, info = compare(oracle.load_scene(fixture_path(name))[:4], filter_fp32=True)
    if info["ray_candidate_tests"]:
        assert info["filtered_predicates"] + info["fp64_fallbacks"] > 0


@pytest.mark.gpu
@pytest.mark.parametrize("grid", [(64, 32), (256, 128)])
def test_torus_knot_fp32_filter_bit_parity(grid):
    from hip_peel import scenes

    print(f"[parity] torus knot {grid} fp32 filter", flush=True)
    p, info = compare(scenes.torus_knot(segments=grid[0], sides=grid[1]), filter_fp32=True)
    total = info["filtered_predicates"] + info["fp64_fallbacks"]
    assert total > 0 and info["fp64_fallbacks"] < total
    # every filtered class decides tests in binary32; the sign classes decide
    # most of theirs (a front_segment edge that a segment crosses needs its
    # binary64 orientations for the clipped interval whatever the filter says)
    from hip_peel.events import FILTER_CLASSES

    for c in FILTER_CLASSES:
        f, b = info[f"filtered_{c}"], info[f"fp64_{c}"]
        print(f"[parity] {c}: {f} binary32, {b} binary64", flush=True)
        assert f > 0 and (c == "front_segment" or f > b), c


@pytest.mark.gpu
def test_stack_fp32_filter_closed_form():
    """Depth gaps of 2^-20 do not touch the 2D filter: the stack's footprint
    tests are decided in binary32 or fall back, and the closed form holds."""
    from hip_peel import EventPeeler, scenes
    from hip_peel.io import upload

    print("[parity] stack 65536 fp32 filter", flush=True)
    n = 65536
    p = EventPeeler(*upload(*scenes.stack(n)), filter_fp32=True)
    p.run()
    info = p.read_status()
    assert info["polygons"] == n and info["layers"] == n and info["triangles"] == 3 * n
    assert info["ray_hits"] == n and info["filtered_predicates"] + info["fp64_fallbacks"] >= 3 * n
    np.testing.assert_array_equal(p.output_layers[:n].cpu().numpy(), p.output_sources[:n].cpu().numpy())
