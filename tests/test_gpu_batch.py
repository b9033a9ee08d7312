"""Many meshes in one graph launch (BatchPeeler over peel_batch_*): every
member's output equals the same mesh peeled on its own, bit for bit as sets,
and matches the exact oracle; replays reset every member."""

import numpy as np
import pytest

import oracle
from peel_checks import check_extraction
from test_gpu_parity import canonical_patches, canonical_triangles, gpu_outputs

pytestmark = pytest.mark.gpu


def meshes(count=6, levels=(2, 3), seed=3):
    from hip_peel import scenes

    return scenes.icosphere_instances(count, seed=seed, levels=levels)


def test_batch_matches_single_peels():
    from hip_peel import BatchPeeler, EventPeeler
    from hip_peel.io import upload

    print("[batch] 6 icospheres", flush=True)
    ms = meshes()
    batch = BatchPeeler.from_meshes(ms, pools="conservative")
    batch.run()
    infos = batch.read_status()
    for k, (m, p, info) in enumerate(zip(ms, batch.peelers, infos)):
        single = EventPeeler(*upload(*m[:4]), propagate_ranks=False)
        single.run()
        one = single.read_status()
        assert info["layers"] == one["layers"] and info["polygons"] == one["polygons"], k
        assert info["triangles"] == one["triangles"], k
        _, r1, s1, l1, uv1, w1, ts1, tl1 = gpu_outputs(p)
        _, r2, s2, l2, uv2, w2, ts2, tl2 = gpu_outputs(single)
        assert canonical_patches(r1, s1, l1) == canonical_patches(r2, s2, l2), k
        assert canonical_triangles(uv1, w1, ts1, tl1) == canonical_triangles(uv2, w2, ts2, tl2), k
    for m, p in list(zip(ms, batch.peelers))[:2]:
        check_extraction(p, m[:4], grid=32)


def test_batch_replay_and_learned_pools():
    """Learned pools with a conservative rebuild on overflow, and replays
    that reset every member."""
    from hip_peel import BatchPeeler

    print("[batch] learned pools and replay", flush=True)
    ms = meshes(count=8, levels=(2, 4), seed=5)
    batch = BatchPeeler.from_meshes(ms)
    batch.run()
    first = [(i["polygons"], i["triangles"], i["layers"]) for i in batch.read_status()]
    batch.run()
    again = [(i["polygons"], i["triangles"], i["layers"]) for i in batch.read_status()]
    assert first == again
    assert all(layers >= 2 for _, _, layers in first)  # every sphere has a back side


def test_batch_overflow_rebuilds_member():
    """A member whose learned pools overflow is rebuilt with the reference's
    conservative capacities (the rerun-on-overflow protocol): a model whose
    margins shrink every prediction far below use."""
    import numpy as np

    from hip_peel import BatchPeeler, pools

    print("[batch] overflow rebuild", flush=True)
    ms = meshes(count=3, levels=(2, 2), seed=7)
    model = pools.load_default()
    tight = pools.PoolModel(model.params, model.x_mean, model.x_std, model.y_mean, model.y_std,
                            np.full(len(model.margin), -3.0))
    b = BatchPeeler.from_meshes(ms, model=tight)
    assert b.rebuilt == [0, 1, 2]
    b.run()
    for m, p in zip(ms, b.peelers):
        check_extraction(p, m[:4], grid=16)


def test_merged_batch_matches_single_peels():
    """All instances in one workspace over an atlas of their windows: the
    merged scene passes the exact oracle, and per instance the layer count
    equals the single peel's, every source's fragments tile the same area,
    and triangles mapped back to the instance's world lie on their sources."""
    import numpy as np

    from hip_peel import EventPeeler, MergedBatch
    from hip_peel.batch import atlas_scene
    from hip_peel.io import upload
    from peel_checks import polygon_area

    print("[batch] merged atlas", flush=True)
    ms = meshes(count=6, levels=(2, 3), seed=11)
    merged = MergedBatch(ms)
    merged.run()
    v, f, c, s, _ = atlas_scene([m[:4] for m in ms])
    check_extraction(merged.peeler, (v, f, c, s), grid=64)
    for k, m in enumerate(ms):
        single = EventPeeler(*upload(*m[:4]))
        single.run()
        one = single.read_status()
        got = merged.instance(k)
        layers = int(got["layers"].max()) + 1 if len(got["layers"]) else 0
        assert layers == one["layers"], (k, layers, one["layers"])
        nf = len(m[1])
        area = np.zeros(nf)
        for r, src in zip(got["rings"], got["sources"]):
            area[src] += polygon_area(r)
        ref = np.zeros(nf)
        polys = single.output.numpy(one["polygons"])
        for x, src in zip(polys, single.output_sources[: one["polygons"]].cpu().numpy()):
            ref[src] += polygon_area(x["v"][: x["n"]])
        np.testing.assert_allclose(area, ref, rtol=0, atol=1e-9)
        tri = m[0][m[1][got["triangle_sources"]]]
        normal = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
        normal /= np.linalg.norm(normal, axis=1)[:, None]
        dist = np.einsum("tk,tck->tc", normal, got["triangles_world"] - tri[:, :1])
        assert np.abs(dist).max() < 1e-8
    # replays reset and reproduce
    first = merged.read_status()
    merged.run()
    again = merged.read_status()
    assert (first["polygons"], first["triangles"], first["layers"]) == (again["polygons"], again["triangles"], again["layers"])


def test_batch_member_with_huge_hit_group():
    """A member whose single ray group holds more than 2,048 hits (stacks of
    5,000 and 3,000 identical footprints) is sorted one phase per launch in
    the batch graph, where its blocks need not be co-resident; every member
    equals its own peel and the stack's layers are its sources."""
    from hip_peel import BatchPeeler, EventPeeler, scenes
    from hip_peel.io import upload

    print("[batch] huge hit groups", flush=True)
    hosts = [scenes.stack(5000), scenes.stack(3000)] + [m[:4] for m in meshes(count=2, levels=(3, 3), seed=13)]
    batch = BatchPeeler([EventPeeler(*upload(*h)) for h in hosts])
    nodes = batch.graph_nodes()
    assert set(nodes) <= {"kernel", "empty", "event_record", "wait_event"}, nodes
    for _ in range(2):
        batch.run()
        infos = batch.read_status()
        for k, (h, p, info) in enumerate(zip(hosts, batch.peelers, infos)):
            single = EventPeeler(*upload(*h))
            single.run()
            one = single.read_status()
            assert (info["layers"], info["polygons"], info["triangles"]) == (
                one["layers"], one["polygons"], one["triangles"]), k
            _, r1, s1, l1, *_ = gpu_outputs(p)
            _, r2, s2, l2, *_ = gpu_outputs(single)
            assert canonical_patches(r1, s1, l1) == canonical_patches(r2, s2, l2), k
        for n, p in zip((5000, 3000), batch.peelers[:2]):
            _, _, srcs, lays, *_ = gpu_outputs(p)
            assert len(srcs) == n
            np.testing.assert_array_equal(lays, srcs)
