"""The benchmark workloads at their own sizes, HIP path vs the oracle.

Each test runs meshes of one benchmark configuration through the C ABI and
compares every product bit for bit with the oracle (ORC_FAST, itself pinned
to the literal loops by test_oracle_modes.py) on the same inputs:

- the plain large mesh: the (3,7) torus knot, 1,048,576 triangles, Level1
  256, Level2 8;
- the mixed-size batch: 16 members of the 256-mesh stream chosen to span
  its triangle range (1k-200k) and every (Level1, Level2) pair, voxelized as
  one batch;
- the exactness stress batch: stress members of about 110k and 1.8M
  triangles (lattice-snapped, sliver and face-coplanar triangles) at Level1
  512, Level2 16, alone and as one batch;
- the served stream: a burst of requests from the served generators through
  the serving loop, with the torus knot at Level1 512 / Level2 16 among them.

Reference semantics: src/Object.cpp:716-779 (Level1 parity), 2256-2342
(Level1 SAT), 3270-3279 (compaction), 1111-1207 (Level2 parity), 2344-2424
(Level2 SAT), 3353-3378 (counts).
"""
import gc

import numpy as np
import pytest

from conftest import compare_results, say

pytestmark = pytest.mark.gpu

COUNTS = ("nb", "l1_inside", "l1_boundary", "l2_inside", "l2_boundary", "tests_l1", "tests_l2", "pairs",
          "max_tris_per_voxel")


def oracle(orc, m):
    g = orc.grid_setup(m["bmin"], m["bmax"], m["mms"], m["voxel_count"], m["voxel_count2"])
    return orc.voxelize(m["soup"], g, orc.ORC_FAST, 0, bins=False), g


def check(name, grid, want, n2):
    got = grid.download()
    info = grid.info()
    compare_results(name, got, want, n2)
    for k in COUNTS:
        assert info[k] == want[k], f"{name}: {k} {info[k]} != {want[k]}"
    return info


def test_torus_knot_256_8(gv, orc, ctx):
    from gpview_amd import workloads
    m = workloads.plain_large(gv)
    want, g = oracle(orc, m)
    say(f"torus knot: T={len(m['soup'])} grid={list(g.n)} nb={want['nb']}")
    grid = ctx.voxelize(m["soup"], m["bmin"], m["bmax"], m["mms"], 256, 8)
    try:
        info = check("torus-knot-256-8", grid, want, 8)
    finally:
        grid.release()
    assert info["n"] == [252, 256, 104] and info["nb"] == 264289


def mixed_selection(workloads, n=256, k=16):
    """k members of the first n of the mixed stream: for every (Level1,
    Level2) pair its smallest and largest member, then the largest and
    smallest overall and members spread over the size range."""
    specs = [(i,) + workloads.mixed_spec(i) for i in range(n)]
    pick = []
    for vc in (64, 128, 256):
        for vc2 in (4, 8):
            grp = sorted((s for s in specs if s[3] == vc and s[4] == vc2), key=lambda s: s[2])
            pick += [grp[0][0], grp[-1][0]]
    by_size = sorted(specs, key=lambda s: s[2])
    for s in (by_size[0], by_size[-1]) + tuple(by_size[j] for j in range(16, n, 48)):
        if len(pick) >= k:
            break
        if s[0] not in pick:
            pick.append(s[0])
    return sorted(set(pick))[:k]


def test_mixed_stream_batch_vs_oracle(gv, orc, ctx):
    from gpview_amd import workloads
    idx = mixed_selection(workloads)
    meshes = [workloads.mixed_member(gv, i) for i in idx]
    sizes = [len(m["soup"]) for m in meshes]
    pairs = {(m["voxel_count"], m["voxel_count2"]) for m in meshes}
    assert len(pairs) == 6 and min(sizes) < 2000 and max(sizes) > 150000, (pairs, sizes)
    assert {m["kind"] for m in meshes} == {0, 1, 2}
    say(f"mixed stream batch: members {idx}, T={sum(sizes)}")
    grids = ctx.voxelize_batch(meshes)
    try:
        for i, m, grid in zip(idx, meshes, grids):
            want, g = oracle(orc, m)
            check(f"mixed member {i} ({len(m['soup'])} tris, {m['voxel_count']}/{m['voxel_count2']})", grid,
                  want, m["voxel_count2"])
    finally:
        for grid in grids:
            grid.release()


@pytest.mark.timeout(900)
def test_stress_512_16_alone_and_batched(gv, orc, ctx):
    """Stress members 19 (~110k triangles) and 9 (~1.8M): alone and as one
    batch, every grid against the oracle."""
    from gpview_amd import workloads
    meshes = [workloads.stress_member(gv, 19), workloads.stress_member(gv, 9)]
    sizes = [len(m["soup"]) for m in meshes]
    assert 100000 <= sizes[0] < 200000 and 1000000 <= sizes[1] <= 2000000, sizes
    say(f"stress 512/16 batch: T={sizes}")
    batch = ctx.voxelize_batch(meshes)
    try:
        for k, m in enumerate(meshes):
            want, g = oracle(orc, m)
            say(f"stress member {k}: grid={list(g.n)} nb={want['nb']} pairs={want['pairs']}")
            check(f"stress {sizes[k]} batched", batch[k], want, 16)
            grid = ctx.voxelize(m["soup"], m["bmin"], m["bmax"], m["mms"], 512, 16)
            try:
                check(f"stress {sizes[k]} alone", grid, want, 16)
            finally:
                grid.release()
            del want
            gc.collect()
    finally:
        for grid in batch:
            grid.release()


@pytest.mark.timeout(900)
def test_served_burst_vs_oracle(gv, orc):
    """A burst from the served generators through the serving loop (two
    lanes, so requests batch and overlap): the torus knot at Level1 512 /
    Level2 16, a 1.3M-triangle icosphere, and the first requests of the
    seeded stream whose library mesh has at most 300k triangles."""
    from gpview_amd import workloads
    lib = workloads.served_library(gv)
    pick, l1, l2, _ = workloads.served_requests(400)
    reqs = [(0, 512, 16)]
    big = max(range(len(lib)), key=lambda i: len(lib[i]["soup"]))
    reqs.append((big, 128, 4))
    for p, a, b in zip(pick, l1, l2):
        if len(reqs) >= 14:
            break
        if len(lib[p]["soup"]) <= 300000 and (a < 512 or len(lib[p]["soup"]) <= 60000):
            reqs.append((int(p), int(a), int(b)))
    assert len(reqs) == 14 and len(lib[big]["soup"]) >= 1000000
    assert {a for _, a, _ in reqs} >= {64, 128, 256, 512} and {b for _, _, b in reqs} == {4, 8, 16}
    say(f"served burst: {[(len(lib[p]['soup']), a, b) for p, a, b in reqs]}")
    with gv.Server(lanes=2, hbm_budget=200e9) as srv:
        ids = [srv.submit(dict(lib[p], voxel_count=a, voxel_count2=b)) for p, a, b in reqs]
        sizes = []
        for rid, (p, a, b) in zip(ids, reqs):
            grid, st = srv.wait(rid, timeout=600)
            assert st["status"] == 0
            sizes.append(st["batch_size"])
            m = dict(lib[p], voxel_count=a, voxel_count2=b)
            want, g = oracle(orc, m)
            check(f"request lib {p} ({len(m['soup'])} tris) at {a}/{b}", grid, want, b)
            srv.release(rid)
            del want
            gc.collect()
    assert max(sizes) > 1, "the burst must be batched"
