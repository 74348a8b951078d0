"""Serving loop (include/gpv_serve.h) through the C ABI.

The contract under test: a request's grid is the same whether it ran alone
(gpv_voxelize) or was admitted into a batch with whatever else was waiting,
under any arrival pattern and memory budget.  Integer occupancy, so "the
same" is bit-exact.
"""
import time

import numpy as np
import pytest

from conftest import compare_results, say

pytestmark = pytest.mark.gpu


def stream(gv, n, seed):
    from gpview_amd import workloads
    return workloads.mixed_stream(gv, n, seed=seed, tri_range=(1000, 60000), level1=(32, 64, 128),
                                  level2=(4, 8))


def alone(ctx, m):
    g = ctx.voxelize(m["soup"], m["bmin"], m["bmax"], m["mms"], m["voxel_count"], m["voxel_count2"], True)
    try:
        return g.download(), g.info()
    finally:
        g.release()


def check_same(name, grid, want, winfo):
    got = grid.download()
    info = grid.info()
    compare_results(name, got, want, winfo["n2"])
    for k in ("l1_inside", "l1_boundary", "l2_inside", "l2_boundary", "tests_l1", "pairs", "max_tris_per_voxel"):
        assert info[k] == winfo[k], f"{name}: {k} {info[k]} != {winfo[k]}"


def test_serve_burst_alone_vs_mix(gv, ctx):
    """24 requests submitted at once: they are batched together across two
    lanes, and each result equals the mesh voxelized alone."""
    meshes = stream(gv, 24, seed=41)
    ref = [alone(ctx, m) for m in meshes]
    say("serve burst: 24 requests, 2 lanes")
    with gv.Server(lanes=2) as srv:
        ids = [srv.submit(m) for m in meshes]
        sizes = []
        for i, rid in enumerate(ids):
            grid, st = srv.wait(rid, timeout=120)
            sizes.append(st["batch_size"])
            assert st["status"] == 0 and st["t_done"] >= st["t_admit"] >= st["t_submit"]
            check_same(f"burst {i}", grid, *ref[i])
            srv.release(rid)
        c = srv.counters()
    assert max(sizes) > 1, "no request was batched with another"
    assert c["completed"] == 24 and c["charged"] == 0


def test_serve_poisson_device_resident(gv, ctx):
    """Poisson arrivals at a fixed rate with device-resident soups; results
    equal the alone runs, and every charge is returned."""
    meshes = stream(gv, 8, seed=43)
    ref = [alone(ctx, m) for m in meshes]
    rng = np.random.default_rng(7)
    picks = rng.integers(0, len(meshes), 60)
    gaps = rng.exponential(1 / 400.0, 60)
    say("serve poisson: 60 requests at 400/s")
    with gv.Server(lanes=2) as srv:
        ptrs = [srv_ctx_upload(gv, srv, ctx, m) for m in meshes]
        try:
            ids = []
            for k, gap in zip(picks, gaps):
                time.sleep(gap)
                ids.append((int(k), srv.submit(meshes[k], device_ptr=ptrs[k])))
            lat = []
            for j, (k, rid) in enumerate(ids):
                grid, st = srv.wait(rid, timeout=120)
                lat.append(st["t_done"] - st["t_submit"])
                check_same(f"poisson {j} (mesh {k})", grid, *ref[k])
                srv.release(rid)
            c = srv.counters()
        finally:
            for p in ptrs:
                ctx.free(p)
    assert c["completed"] == 60 and c["charged"] == 0
    say(f"poisson p50 {np.percentile(lat, 50) * 1e3:.2f} ms p99 {np.percentile(lat, 99) * 1e3:.2f} ms")


def srv_ctx_upload(gv, srv, ctx, m):
    """Device copy of a soup for tris_on_device requests (any context's pool
    gives a device pointer every context on the device can read)."""
    return ctx.upload(m["soup"])


def test_serve_budget_limits_admission(gv, ctx):
    """A budget of about three requests' predicted footprint: the charge never
    exceeds it, requests still all complete and stay exact."""
    meshes = stream(gv, 10, seed=47)
    ref = [alone(ctx, m) for m in meshes]
    pred = [gv.predict_footprint(m) for m in meshes]
    budget = int(sorted(pred)[-1] * 3)
    say(f"serve budget {budget} bytes")
    with gv.Server(lanes=2, hbm_budget=budget) as srv:
        ids = [srv.submit(m) for m in meshes]
        for i, rid in enumerate(ids):
            grid, st = srv.wait(rid, timeout=120)
            check_same(f"budget {i}", grid, *ref[i])
            srv.release(rid)
        c = srv.counters()
    assert c["peak_charged"] <= budget and c["charged"] == 0
