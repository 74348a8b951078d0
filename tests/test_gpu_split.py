"""Batches beyond one run's limits or memory, and serving failure handling.

- gpv_voxelize_batch cuts a batch whose padded Level1 voxels exceed 2^32-1
  into consecutive runs, and splits a run that fails on device memory into
  halves; every grid still equals the mesh voxelized alone.
- The serving loop re-admits the requests of a failed batch one at a time
  (the serving analogue of the reference's capacity rerun,
  src/Object.cpp:3198-3214); only a request that fails alone is reported.
- Each served batch's measured peak of device bytes stays within the sum of
  its requests' predicted footprints (gpv_predict_footprint), the number
  admission charges against the budget.
- A vertex coordinate outside the input domain (not finite, or beyond 2^40
  in magnitude) fails the call with GPV_ERR_INVALID, alone or in a batch,
  holds nothing afterwards and leaves the context usable; 2^40 itself is
  inside.
"""
import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu


def bits_and_info(g):
    return g.download_bits(), g.info()


def same(name, a, b):
    (ba, ia), (bb, ib) = a, b
    for k in ("nb", "pairs", "l1_inside", "l1_boundary", "l2_inside", "l2_boundary", "tests_l1", "tests_l2",
              "max_tris_per_voxel", "n"):
        assert ia[k] == ib[k], f"{name}: {k} {ia[k]} != {ib[k]}"
    for x, y in zip(ba, bb):
        assert np.array_equal(x, y), name


def alone(ctx, m):
    g = ctx.voxelize(m["soup"], m["bmin"], m["bmax"], m["mms"], m["voxel_count"], m["voxel_count2"],
                     m.get("level2", True))
    try:
        return bits_and_info(g)
    finally:
        g.release()


def test_batch_beyond_voxel_limit_is_cut(gv, ctx):
    """34 meshes at Level1 512 hold 34 * 512^3 > 2^32 padded voxels."""
    from gpview_amd import workloads
    meshes = []
    for i in range(34):
        v, f = gv.icosphere(2, 0.3 + 0.004 * i, (0.5, 0.5, 0.5), 0.02, seed=100 + i)
        meshes.append(workloads.mesh_entry(gv, v, f, 512, 2))
    vox = sum(m_n[2] * m_n[1] * ((m_n[0] + 63) // 64) * 64
              for m_n in (gv.grid_derive(m["bmin"], m["bmax"], m["mms"], 512, 2).n for m in meshes))
    assert vox >= 2 ** 32
    say(f"batch of 34 at Level1 512: {vox} padded voxels")
    grids = ctx.voxelize_batch(meshes)
    try:
        got = [bits_and_info(g) for g in grids]
    finally:
        for g in grids:
            g.release()
    for i in (0, 1, 16, 17, 33):
        same(f"mesh {i}", got[i], alone(ctx, meshes[i]))


def test_batch_split_on_memory(gv):
    """A context capped below the batch's working set but above each mesh's:
    the batch splits into runs that fit and gives the same grids; a mesh
    that does not fit alone fails with GPV_ERR_OOM and leaves nothing held."""
    from gpview_amd import workloads
    meshes = workloads.mixed_stream(gv, 8, seed=151, tri_range=(40000, 80000), level1=(64,), level2=(4,))
    with gv.Context(0) as c:
        ref = [alone(c, m) for m in meshes]
        single = 0
        for m in meshes:
            base = c.pool_stats()["live_bytes"]
            with gv.Context(0) as c1:
                g = c1.voxelize(m["soup"], m["bmin"], m["bmax"], m["mms"], 64, 4)
                single = max(single, c1.pool_stats()["peak_live_bytes"])
                g.release()
            assert c.pool_stats()["live_bytes"] == base
        with gv.Context(0) as c2:
            gs = c2.voxelize_batch(meshes)
            whole = c2.pool_stats()["peak_live_bytes"]
            for g in gs:
                g.release()
        limit = int(single * 1.5)
        say(f"memory split: single {single} batch {whole} limit {limit}")
        assert whole > 2 * limit, "the batch must not fit the limit"
        with gv.Context(0) as c3:
            c3.set_mem_limit(limit)
            gs = c3.voxelize_batch(meshes)
            try:
                got = [bits_and_info(g) for g in gs]
            finally:
                for g in gs:
                    g.release()
            for i in range(len(meshes)):
                same(f"split mesh {i}", got[i], ref[i])
            assert c3.pool_stats()["peak_live_bytes"] <= limit
            c3.set_mem_limit(single // 4)
            with pytest.raises(gv.GpvError) as e:
                c3.voxelize_batch(meshes[:3])
            assert e.value.status == gv.GPV_ERR_OOM
            assert c3.pool_stats()["live_bytes"] == 0


def test_serve_readmits_failed_batches(gv, ctx):
    """Lanes capped so one request of a burst cannot fit even alone: the
    batches it lands in fail, their other requests are re-admitted one at a
    time and complete with the grid they have alone, and only the oversized
    request is reported failed (GPV_ERR_OOM)."""
    from gpview_amd import workloads
    meshes = workloads.mixed_stream(gv, 10, seed=153, tri_range=(30000, 70000), level1=(64,), level2=(4,))
    ref = [alone(ctx, m) for m in meshes]
    single = resident = 0
    for m in meshes:
        with gv.Context(0) as c1:
            g = c1.voxelize(m["soup"], m["bmin"], m["bmax"], m["mms"], 64, 4)
            single = max(single, c1.pool_stats()["peak_live_bytes"])
            resident = max(resident, c1.pool_stats()["live_bytes"])
            g.release()
    big = workloads.mixed_member(gv, 3, seed=153, tri_range=(400000, 400001), level1=(256,), level2=(16,))
    with gv.Context(0) as c1:
        g = c1.voxelize(big["soup"], big["bmin"], big["bmax"], big["mms"], 256, 16)
        big_need = c1.pool_stats()["peak_live_bytes"]
        g.release()
    # room for any one small mesh's run plus every small result kept resident
    limit = 2 * single + len(meshes) * resident
    assert big_need > limit, (big_need, limit)
    say(f"serve with lane limit {limit} (small run {single}, big run {big_need})")
    order = meshes[:5] + [big] + meshes[5:]
    with gv.Server(lanes=2, lane_mem_limit=limit) as srv:
        ids = [srv.submit(m) for m in order]
        stats = {}
        for k, rid in enumerate(ids):
            if k == 5:
                with pytest.raises(gv.GpvError) as e:
                    srv.wait(rid, timeout=300)
                assert e.value.status == gv.GPV_ERR_OOM
                continue
            i = k if k < 5 else k - 1
            grid, st = srv.wait(rid, timeout=300)
            stats[i] = st
            same(f"served {i}", bits_and_info(grid), ref[i])
        for rid in ids:
            srv.release(rid)
        cnt = srv.counters()
    assert cnt["readmitted"] > 0 and any(s["retries"] == 1 for s in stats.values())
    assert all(s["status"] == 0 for s in stats.values()) and cnt["charged"] == 0
    assert cnt["completed"] == len(order)


def test_serve_peak_within_predicted_footprint(gv):
    """Requests from the served generators (Level1 64-512, Level2 4-16),
    submitted at once: every batch's measured peak of device bytes stays
    within the footprint admission predicted for it."""
    from gpview_amd import workloads
    lib = workloads.served_library(gv)
    pick, l1, l2, _ = workloads.served_requests(300)
    reqs = [(int(p), int(a), int(b)) for p, a, b in zip(pick, l1, l2) if len(lib[p]["soup"]) <= 250000][:40]
    say(f"footprint check: {len(reqs)} requests")
    with gv.Server(lanes=2) as srv:
        ids = [srv.submit(dict(lib[p], voxel_count=a, voxel_count2=b)) for p, a, b in reqs]
        stats = [srv.wait(rid, timeout=300)[1] for rid in ids]
        for rid in ids:
            srv.release(rid)
    batches = {}
    for s in stats:
        assert s["status"] == 0
        batches[s["batch_id"]] = (s["batch_peak_bytes"], s["batch_predicted_bytes"], s["batch_size"])
    ratios = [peak / pred for peak, pred, _ in batches.values()]
    say(f"{len(batches)} batches, peak/predicted max {max(ratios):.3f} median {np.median(ratios):.3f}")
    for bid, (peak, pred, n) in batches.items():
        assert 0 < peak <= pred, f"batch {bid} of {n}: measured {peak} > predicted {pred}"


def test_out_of_domain_coordinates_refused(gv):
    from gpview_amd import workloads
    v, f = gv.icosphere(3, 0.4, (0.5, 0.5, 0.5), 0.02, seed=11)
    good = workloads.mesh_entry(gv, v, f, 48, 4)
    with gv.Context(0) as c:
        ref = alone(c, good)
        for bad_value in (np.nan, np.inf, -np.inf, np.float32(2.0 ** 41), np.float32(-3e38)):
            m = dict(good)
            soup = good["soup"].copy()
            soup[77, 4] = bad_value
            m["soup"] = soup
            say(f"out of domain: {bad_value}")
            with pytest.raises(gv.GpvError) as e:
                alone(c, m)
            assert e.value.status == gv.GPV_ERR_INVALID
            assert c.pool_stats()["live_bytes"] == 0
            batch = [good, good, good, m, good]
            with pytest.raises(gv.GpvError) as e:
                c.voxelize_batch(batch)
            assert e.value.status == gv.GPV_ERR_INVALID
            assert c.pool_stats()["live_bytes"] == 0
            same(f"after {bad_value}", alone(c, good), ref)
        # the no-allocation entry point refuses it too, and runs the next mesh
        nx, ny, nz = ref[1]["n"]
        w1 = nz * ny * ((nx + 63) // 64)
        nb_cap = ref[1]["nb"] + 8
        w2 = (ref[1]["n2"] ** 3 + 63) // 64
        bufs = [c.upload(np.zeros(n, np.uint8)) for n in (w1 * 8, w1 * 8, nb_cap * 4, nb_cap * w2 * 8,
                                                           nb_cap * w2 * 8)]
        soup = good["soup"].copy()
        soup[5, 0] = np.nan
        dbad, dgood = c.upload(soup), c.upload(good["soup"])
        try:
            args = (good["bmin"], good["bmax"], good["mms"], 48, 4, True, bufs[0], bufs[1], w1, bufs[2], bufs[3],
                    bufs[4], nb_cap)
            with pytest.raises(gv.GpvError) as e:
                c.voxelize_noalloc(dbad, len(soup), *args)
            assert e.value.status == gv.GPV_ERR_INVALID
            assert c.voxelize_noalloc(dgood, len(soup), *args)["nb"] == ref[1]["nb"]
        finally:
            for p in bufs + [dbad, dgood]:
                c.free(p)
        # the domain's edge: a far vertex at exactly 2^40 is voxelized
        m = dict(good)
        soup = good["soup"].copy()
        soup[77, 4] = np.float32(2.0 ** 40)
        m["soup"] = soup
        edge = alone(c, m)
        assert edge[1]["nb"] > 0
