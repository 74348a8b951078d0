"""Batched voxelization (gpv_voxelize_batch) through the C ABI.

Every mesh of a batch must hold exactly what it holds when voxelized alone:
bit-exact against the oracle at sizes the oracle finishes quickly, and
against the single-mesh HIP path (itself pinned to the oracle by
test_gpu_parity.py) at the batch sizes of the mixed-stream workload.  The
reference voxelizes one Object per PerformVoxelization call
(src/GPView.cpp:1378-1380), so "alone" is the reference's own unit.
"""
import numpy as np
import pytest

from conftest import compare_results, say

pytestmark = pytest.mark.gpu

KEYS = ("l1_inside", "l1_boundary", "l2_inside", "l2_boundary", "tests_l1", "tests_l2", "pairs",
        "max_tris_per_voxel", "nb", "n1", "n", "n2")


def mesh_entry(gv, verts, faces, vc, vc2, level2=True, initial_capacity=0):
    soup = gv.flatten(verts, faces)
    bmin, bmax, mms = gv.mesh_bbox(verts)
    return {"soup": soup, "bmin": bmin, "bmax": bmax, "mms": mms, "voxel_count": vc, "voxel_count2": vc2,
            "level2": level2, "initial_capacity": initial_capacity}


def small_mix(gv, seed=100):
    """Eleven meshes: the three generator kinds at mixed sizes and
    resolutions, Level2 off for one, a single triangle and an empty soup."""
    out = []
    specs = [(0, 2000, 32, 4), (1, 3000, 16, 2), (2, 4000, 32, 8), (0, 800, 48, 4), (1, 6000, 24, 4),
             (2, 2500, 20, 2), (0, 1500, 32, 8), (1, 1200, 40, 4)]
    for i, (kind, target, vc, vc2) in enumerate(specs):
        v, f = gv.mixed(kind, target, seed + i)
        out.append(mesh_entry(gv, v, f, vc, vc2))
    v, f = gv.mixed(2, 3000, seed + 50)
    out.append(mesh_entry(gv, v, f, 32, 4, level2=False))
    tri_v = np.array([[0.1, 0.2, 0.3], [0.9, 0.3, 0.5], [0.4, 0.8, 0.7]], np.float32)
    out.append(mesh_entry(gv, tri_v, np.array([[0, 1, 2]], np.int32), 16, 4))
    e = mesh_entry(gv, tri_v, np.array([[0, 1, 2]], np.int32), 16, 4)
    e["soup"] = np.zeros((0, 9), np.float32)
    out.append(e)
    return out


def oracle_of(orc, m):
    g = orc.grid_setup(m["bmin"], m["bmax"], m["mms"], m["voxel_count"],
                       m["voxel_count2"] if m["level2"] else 0)
    return orc.voxelize(m["soup"], g, orc.ORC_FAST, 0), g.n2


def alone(ctx, m):
    grid = ctx.voxelize(m["soup"], m["bmin"], m["bmax"], m["mms"], m["voxel_count"], m["voxel_count2"],
                        m["level2"], m.get("initial_capacity", 0))
    try:
        return grid.download(), grid.info()
    finally:
        grid.release()


def batch_results(ctx, meshes, **kw):
    grids = ctx.voxelize_batch(meshes, **kw)
    out = []
    try:
        for g in grids:
            out.append((g.download(), g.info()))
    finally:
        for g in grids:
            g.release()
    return out


def check_info(name, info, ref):
    for k in KEYS:
        if k in ref:
            assert info[k] == ref[k], f"{name}: {k} {info[k]} != {ref[k]}"


def test_batch_small_mix_vs_oracle(gv, orc, ctx):
    meshes = small_mix(gv)
    say(f"batch small mix: {len(meshes)} meshes, T={sum(len(m['soup']) for m in meshes)}")
    res = batch_results(ctx, meshes)
    for i, (m, (got, info)) in enumerate(zip(meshes, res)):
        want, n2 = oracle_of(orc, m)
        compare_results(f"mesh {i}", got, want, n2)
        check_info(f"mesh {i}", info, want)


def test_batch_high_level2_mix_vs_oracle(gv, orc, ctx):
    """Level2 resolutions above 8 beside small ones in one run: groups of one
    voxel (n2 = 16) and of many (n2 = 4) in the grouped Level2 SAT, the
    Level2 parity's run search (n2 = 9, 12, 16) next to its fixed-loop and
    four-voxel paths, and column-pass waves whose entries span meshes of
    different resolutions.  Also the same meshes with the per-voxel Level2
    SAT (grouping off)."""
    specs = [(0, 6000, 40, 16), (1, 5000, 32, 4), (2, 6000, 36, 12), (0, 3000, 24, 9), (1, 4000, 28, 8),
             (2, 2000, 20, 2), (0, 4000, 30, 16)]
    meshes = []
    for i, (kind, target, vc, vc2) in enumerate(specs):
        v, f = gv.mixed(kind, target, 700 + i)
        meshes.append(mesh_entry(gv, v, f, vc, vc2))
    want = [oracle_of(orc, m) for m in meshes]
    before = ctx.get_tune()
    try:
        for below in (1 << 20, 0):
            ctx.set_tune(l2_group_below=below)
            say(f"batch high Level2 mix: {len(meshes)} meshes, l2_group_below {below}")
            res = batch_results(ctx, meshes)
            for i, (m, (got, info)) in enumerate(zip(meshes, res)):
                compare_results(f"mesh {i} (below {below})", got, want[i][0], want[i][1])
                check_info(f"mesh {i} (below {below})", info, want[i][0])
    finally:
        ctx.set_tune(**before)


def test_batch_order_independent(gv, ctx):
    """Each mesh's grid is the same in any batch position (alone vs in the mix)."""
    meshes = small_mix(gv, seed=300)
    say("batch order: forward")
    fwd = batch_results(ctx, meshes)
    perm = [5, 0, 10, 3, 9, 1, 7, 2, 8, 4, 6]
    say("batch order: permuted")
    rev = batch_results(ctx, [meshes[i] for i in perm])
    for j, i in enumerate(perm):
        compare_results(f"mesh {i} at {j}", rev[j][0], fwd[i][0], fwd[i][1]["n2"])
        check_info(f"mesh {i} at {j}", rev[j][1], fwd[i][1])


def test_batch_of_one_equals_plain(gv, ctx):
    v, f = gv.mixed(0, 20000, 7)
    m = mesh_entry(gv, v, f, 64, 8)
    say("batch of one")
    (got, info), = batch_results(ctx, [m])
    want, winfo = alone(ctx, m)
    compare_results("batch-of-one", got, want, 8)
    check_info("batch-of-one", info, winfo)


def test_batch_capacity_regrow(gv, orc, ctx):
    meshes = small_mix(gv, seed=500)[:6]
    for m in meshes:
        m["initial_capacity"] = 100
    say("batch regrow")
    res = batch_results(ctx, meshes)
    assert res[0][1]["regrows"] >= 1
    for i, (m, (got, info)) in enumerate(zip(meshes, res)):
        want, n2 = oracle_of(orc, m)
        compare_results(f"regrow mesh {i}", got, want, n2)


def test_batch_device_soups(gv, ctx):
    """Device-resident inputs (tris_on_device = 1) give the same grids."""
    meshes = small_mix(gv, seed=700)[:5]
    ptrs = [ctx.upload(m["soup"]) if len(m["soup"]) else None for m in meshes]
    try:
        say("batch device soups")
        dev = batch_results(ctx, meshes, device_ptrs=ptrs)
    finally:
        for p in ptrs:
            if p:
                ctx.free(p)
    host = batch_results(ctx, meshes)
    for i in range(len(meshes)):
        compare_results(f"device mesh {i}", dev[i][0], host[i][0], host[i][1]["n2"])


def test_batch_mixed_stream_vs_plain(gv, ctx):
    """A slice of the mixed-size stream (triangle counts log-uniform in
    [1k, 200k], Level1 {64,128,256}, Level2 {4,8}) against the single-mesh
    path, mesh by mesh."""
    rng = np.random.default_rng(2026)
    meshes = []
    for i in range(12):
        kind = int(rng.integers(0, 3))
        target = int(np.exp(rng.uniform(np.log(1000), np.log(200000))))
        vc = int(rng.choice([64, 128, 256]))
        vc2 = int(rng.choice([4, 8]))
        v, f = gv.mixed(kind, target, 9000 + i)
        meshes.append(mesh_entry(gv, v, f, vc, vc2))
    say(f"mixed stream slice: T={sum(len(m['soup']) for m in meshes)}")
    res = batch_results(ctx, meshes)
    for i, m in enumerate(meshes):
        want, winfo = alone(ctx, m)
        compare_results(f"stream mesh {i}", res[i][0], want, winfo["n2"])
        check_info(f"stream mesh {i}", res[i][1], winfo)


def test_batch_views_and_save(gv, ctx, tmp_path):
    """save_raw of a batch member writes the same files as the mesh alone."""
    meshes = small_mix(gv, seed=900)[:3]
    grids = ctx.voxelize_batch(meshes)
    try:
        (tmp_path / "b").mkdir()
        grids[1].save_raw(5, str(tmp_path / "b"))
    finally:
        for g in grids:
            g.release()
    m = meshes[1]
    g = ctx.voxelize(m["soup"], m["bmin"], m["bmax"], m["mms"], m["voxel_count"], m["voxel_count2"], True)
    try:
        (tmp_path / "a").mkdir()
        g.save_raw(5, str(tmp_path / "a"))
    finally:
        g.release()
    for name in ("VoxelConfig.txt", "Level1InOut.raw", "Level1BoundaryPrefixSum.raw", "Level2InOut.raw"):
        assert (tmp_path / "a" / f"Obj5{name}").read_bytes() == (tmp_path / "b" / f"Obj5{name}").read_bytes()


def test_batch_stress_vs_oracle(gv, orc, ctx):
    """The exactness-stress generators (lattice-snapped vertices p = 0.3, ~5%
    face-coplanar triangles, 10% slivers) at Level2 16, reduced in size so the
    oracle checks every mesh: each batch member bit-exact."""
    from gpview_amd import workloads
    meshes = workloads.stress(gv, 4, seed=91, tri_range=(20000, 60000), vc=96, vc2=16)
    say(f"stress batch: T={sum(len(m['soup']) for m in meshes)} L1 96 L2 16")
    res = batch_results(ctx, meshes)
    for i, (m, (got, info)) in enumerate(zip(meshes, res)):
        want, n2 = oracle_of(orc, m)
        compare_results(f"stress mesh {i}", got, want, n2)
        check_info(f"stress mesh {i}", info, want)


def test_pool_reuses_blocks_across_batch_sizes(gv):
    """Batches of different sizes, repeated: the second pass over the same
    sequence is served from cached blocks (a serving loop's steady state
    allocates nothing), and results are unchanged by block reuse."""
    from gpview_amd import workloads
    ms = workloads.mixed_stream(gv, 12, seed=91, tri_range=(1000, 20000), level1=(32, 64), level2=(4, 8))
    cuts = [(0, 12), (0, 5), (5, 12), (2, 4), (0, 12)]
    with gv.Context(0) as c:
        ptrs = [c.upload(m["soup"]) for m in ms]
        first = {}

        def one_pass():
            for a, b in cuts:
                gs = c.voxelize_batch(ms[a:b], device_ptrs=ptrs[a:b])
                for i, g in zip(range(a, b), gs):
                    d = g.download()
                    got = np.concatenate([d["l1"], d["l2"], d["prefix"].view(np.uint8)])
                    if i in first:
                        assert np.array_equal(got, first[i]), f"mesh {i} changed with block reuse"
                    else:
                        first[i] = got
                    g.release()

        s0 = c.pool_stats()
        one_pass()
        s1 = c.pool_stats()
        one_pass()
        s2 = c.pool_stats()
        say(f"pool: {s0} -> {s1} -> {s2}")
        assert s1["allocations"] > s0["allocations"]
        assert s2["allocations"] - s1["allocations"] <= (s1["allocations"] - s0["allocations"]) // 4
        assert s2["reuses"] > s1["reuses"]
        assert s2["cached_bytes"] <= s2["cap_bytes"]
        for p in ptrs:
            c.free(p)


def test_autotune_keeps_results(gv):
    """gpv_ctx_autotune picks launch shapes by measurement; no shape changes a
    result, so grids before and after tuning (and under hand-set extreme
    shapes) are identical."""
    from gpview_amd import workloads
    ms = workloads.mixed_stream(gv, 6, seed=95, tri_range=(1000, 12000), level1=(32, 64), level2=(4, 8))
    with gv.Context(0) as c:
        def run():
            gs = c.voxelize_batch(ms)
            out = [np.concatenate([d["l1"], d["l2"], d["prefix"].view(np.uint8)]) for d in (g.download() for g in gs)]
            for g in gs:
                g.release()
            return out
        base = run()
        shapes, best_ms = c.autotune(ms, repeats=2)
        say(f"autotune: {shapes} {best_ms:.3f} ms")
        assert shapes == c.get_tune() and best_ms > 0
        assert shapes["expand_blocks_per_cu"] in (4, 8, 16, 32) and shapes["l2_overlap"] in (0, 1)
        assert shapes["sort_items_per_block"] in (1024, 2048, 4096, 8192)
        assert shapes["l2_sat_waves"] in (1, 2, 4) and shapes["l2_parity_waves"] in (1, 2, 4)
        assert shapes["expand_stage_tris"] in (128, 256)
        for a, b in zip(base, run()):
            assert np.array_equal(a, b)
        for extreme in ({"l2_group_below": 0, "l2_overlap": 0, "expand_blocks_per_cu": 1, "sort_items_per_block": 256,
                         "l2_sat_waves": 1, "l2_parity_waves": 1, "expand_stage_tris": 128},
                        {"l2_group_below": 1 << 20, "sort_items_per_block": 1 << 20, "l2_sat_waves": 4,
                         "l2_parity_waves": 2, "expand_stage_tris": 256},
                        {"l2_group_below": 0, "l2_sat_waves": 4, "l2_parity_waves": 4}):
            c.set_tune(**extreme)
            for a, b in zip(base, run()):
                assert np.array_equal(a, b), extreme
        with pytest.raises(gv.GpvError):
            c.set_tune(sort_items_per_block=1)
        with pytest.raises(gv.GpvError):
            c.set_tune(l2_sat_waves=3)


def test_bits_download_matches_states(gv, ctx):
    """gpv_grid_download_bits returns the bitplanes behind gpv_grid_download's
    state bytes: rebuilt states are equal."""
    from gpview_amd import workloads
    ms = workloads.mixed_stream(gv, 3, seed=97, tri_range=(2000, 9000), level1=(40, 64), level2=(4, 8))
    gs = ctx.voxelize_batch(ms)
    try:
        for g in gs:
            i = g.info()
            nx, ny, nz = i["n"]
            nxw = (nx + 63) // 64
            b1, n1, b2, n2b = g.download_bits()
            d = g.download()
            bits = lambda w, n: np.unpackbits(w.view(np.uint8), bitorder="little")[:n]
            bnd = np.unpackbits(b1.view(np.uint8), bitorder="little").reshape(nz, ny, nxw * 64)[:, :, :nx].ravel()
            ins = np.unpackbits(n1.view(np.uint8), bitorder="little").reshape(nz, ny, nxw * 64)[:, :, :nx].ravel()
            assert np.array_equal(np.where(bnd == 1, 2, ins).astype(np.uint8), d["l1"])
            n23 = i["n2"] ** 3
            w2 = (n23 + 63) // 64
            sb = np.unpackbits(b2.view(np.uint8), bitorder="little").reshape(i["nb"], w2 * 64)[:, :n23].ravel()
            pb = np.unpackbits(n2b.view(np.uint8), bitorder="little").reshape(i["nb"], w2 * 64)[:, :n23].ravel()
            assert np.array_equal(np.where(sb == 1, 2, pb).astype(np.uint8), d["l2"])
    finally:
        for g in gs:
            g.release()


def test_stress_level1_512_batch_equals_alone(gv, ctx):
    """Exactness stress at the stress workload's resolution (Level1 512,
    Level2 16; lattice-snapped, sliver and face-coplanar triangles): each
    mesh of a batch has the same bitplanes as the mesh alone (the single-mesh
    path itself is pinned to the oracle by the parity tests)."""
    from gpview_amd import workloads
    ms = workloads.stress(gv, 3, seed=101, tri_range=(60000, 250000))
    say(f"stress 512/16: {[len(m['soup']) for m in ms]} triangles")
    gs = ctx.voxelize_batch(ms)
    try:
        for k, (m, gb) in enumerate(zip(ms, gs)):
            ga = ctx.voxelize(m["soup"], m["bmin"], m["bmax"], m["mms"], 512, 16, True)
            try:
                ib, ia = gb.info(), ga.info()
                for key in ("nb", "pairs", "l1_inside", "l2_inside", "l2_boundary", "max_tris_per_voxel"):
                    assert ib[key] == ia[key], (k, key)
                for x, y in zip(gb.download_bits(), ga.download_bits()):
                    assert np.array_equal(x, y), k
            finally:
                ga.release()
    finally:
        for g in gs:
            g.release()


STAGED_BATCHES = {
    "large": [(1, 900000), (2, 700000), (1, 500000), (1, 300000), (0, 40000)],
    # just above the 2 MiB staging threshold, the small meshes at the cut
    "small": [(0, 30000), (1, 40000), (2, 20000), (0, 900), (1, 500), (2, 3000)],
    "small_mixed_memory": [(0, 30000), (1, 40000), (2, 20000), (0, 900), (1, 500), (2, 3000)],
}


@pytest.mark.parametrize("which", sorted(STAGED_BATCHES))
def test_batch_staged_uploads_equal_resident(gv, ctx, which):
    """Host soups of a batch above 2 MiB go up in two runs, the second
    uploaded under the first's kernels (soups of up to 256 KiB in
    page-locked memory by one gather kernel reading across the bus, the rest
    by a copy each); the grids equal those of the same batch from resident
    soups."""
    meshes = []
    for i, (kind, target) in enumerate(STAGED_BATCHES[which]):
        v, f = gv.mixed(kind, target, 900 + i)
        meshes.append(mesh_entry(gv, v, f, 96 + 32 * (i % 2), 4 + 4 * (i % 2)))
    if which.startswith("small"):
        # an empty soup and a single triangle where the first run ends
        meshes.insert(1, dict(meshes[0], soup=meshes[0]["soup"][:0]))
        meshes.insert(2, dict(meshes[0], soup=meshes[0]["soup"][:1]))
    pinned = [gv.pinned_copy(m["soup"]) for m in meshes]
    host = [dict(m, soup=p.array) for m, p in zip(meshes, pinned)]
    if which == "small_mixed_memory":
        # pageable soups between page-locked ones, and a page-locked soup
        # that starts inside its allocation and ends at its end
        host = [dict(m, soup=m["soup"]) if i % 2 else h for i, (m, h) in enumerate(zip(meshes, host))]
        small = min((k for k, m in enumerate(meshes) if len(m["soup"]) > 100), key=lambda k: len(meshes[k]["soup"]))
        meshes.append(dict(meshes[small], soup=meshes[small]["soup"][7:]))
        host.append(dict(meshes[small], soup=pinned[small].array[7:]))
    total = sum(m["soup"].nbytes for m in meshes)
    assert total > (2 << 20)
    ptrs = [ctx.upload(m["soup"]) for m in meshes]
    try:
        say(f"staged batch: {len(meshes)} meshes, {total >> 20} MiB of soup")
        a = batch_results(ctx, host)
        b = batch_results(ctx, meshes, device_ptrs=ptrs)
    finally:
        for p in ptrs:
            ctx.free(p)
        for p in pinned:
            p.close()
    for i, ((ga, ia), (gb, ib)) in enumerate(zip(a, b)):
        for k in ("l1", "prefix", "bidx", "l2"):
            assert np.array_equal(ga[k], gb[k]), f"member {i}: {k}"
        check_info(f"member {i}", ia, ib)


def test_batch_rewritten_pinned_soups(gv, ctx):
    """Page-locked soups rewritten in place between two batch calls: the
    second call reads the new contents (the gather of small soups must not
    return lines cached from the first)."""
    meshes = []
    for i, (kind, target) in enumerate(STAGED_BATCHES["small"]):
        v, f = gv.mixed(kind, target, 950 + i)
        meshes.append(mesh_entry(gv, v, f, 64, 4))
    pinned = [gv.pinned_copy(m["soup"]) for m in meshes]
    host = [dict(m, soup=p.array) for m, p in zip(meshes, pinned)]

    def resident(ms):
        ptrs = [ctx.upload(m["soup"]) for m in ms]
        try:
            return batch_results(ctx, ms, device_ptrs=ptrs)
        finally:
            for q in ptrs:
                ctx.free(q)

    try:
        say("rewritten pinned soups: first call")
        first = batch_results(ctx, host)
        want_first = resident(meshes)
        # shrink every mesh about its box center, in the same buffers
        shrunk = []
        for m, p in zip(meshes, pinned):
            c = (np.asarray(m["bmin"], np.float32) + np.asarray(m["bmax"], np.float32)) * np.float32(0.5)
            t = m["soup"].reshape(-1, 3, 3)
            new = (c + (t - c) * np.float32(0.75)).astype(np.float32).reshape(-1, 9)
            p.array[...] = new
            shrunk.append(dict(m, soup=new))
        say("rewritten pinned soups: second call")
        second = batch_results(ctx, host)
        want_second = resident(shrunk)
    finally:
        for p in pinned:
            p.close()
    for name, got, want in (("first", first, want_first), ("second", second, want_second)):
        for i, ((ga, ia), (gb, ib)) in enumerate(zip(got, want)):
            for k in ("l1", "prefix", "bidx", "l2"):
                assert np.array_equal(ga[k], gb[k]), f"{name} call, member {i}: {k}"
            check_info(f"{name} call, member {i}", ia, ib)
    assert any(not np.array_equal(a[0]["l1"], b[0]["l1"]) for a, b in zip(first, second))
