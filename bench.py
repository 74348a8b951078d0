"""Benchmark of the plain (one mesh per call) voxelizer on one MI355X.

Workload: the seeded (3,7) torus-knot tube, 2,048 segments x 256 ring
vertices (1,048,576 FP32 triangles), scaled to 0.9 of the unit cube;
Level1 voxelCount 256, Level2 8^3 per boundary voxel.  One step = one
gpv_voxelize call on the device-resident soup (all stages, counts, sync).

Prints ONE JSON line: meshes/s (value), triangle-box tests/s, voxels/s,
p50/p99 step latency, per-stage device times, the roofline of the dominant
kernel (HIP events on the library's stream) and the CPU oracle timed on this
host on the same mesh.

Also measured (key "batched"): the mixed-size stream of 256 meshes, one
gpv_voxelize per mesh against one gpv_voxelize_batch for all of them; and
(key "served") a Poisson request stream through the serving loop, with
p50/p99 latency at the stated arrival rate.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--batch-meshes N]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
VALU_PEAK_OPS = 78.65e12       # MI355X FP32 vector peak 157.3 TFLOP/s counts an FMA as 2: lane-ops per second
STAGES = ["setup", "l1_sat", "l1_parity", "compaction_bins", "l2_parity", "l2_sat"]


def make_workload(gv):
    v, f = gv.torus_knot(p=3, q=7, segments=2048, ring=256, tube=0.3, jitter=0.01, extent=0.9, seed=2)
    return v, f, gv.flatten(v, f)


def bench_batched(gv, ctx, n_meshes, steps, warmup, check, small=False, autotune=False):
    """The mixed-size stream: n_meshes meshes (triangles log-uniform in
    [1k, 200k], Level1 {64,128,256}, Level2 {4,8}), device-resident soups.
    Plain = one gpv_voxelize per mesh; batched = one gpv_voxelize_batch over
    all of them.  Both timed over whole passes, grids released each pass.
    small=True: the same generators at 1k-20k triangles, Level1 64, Level2 4
    (meshes whose single-mesh run is dominated by per-call costs)."""
    from gpview_amd import workloads
    if small:
        meshes = workloads.mixed_stream(gv, n_meshes, seed=7, tri_range=(1000, 20000), level1=(64,), level2=(4,))
    else:
        meshes = workloads.mixed_stream(gv, n_meshes)
    ptrs = [ctx.upload(m["soup"]) for m in meshes]
    T = sum(len(m["soup"]) for m in meshes)
    tune = None
    if autotune:
        shapes, tms = ctx.autotune(meshes, device_ptrs=ptrs, repeats=2)
        tune = {"shapes": shapes, "device_ms": tms}

    def plain_pass():
        for m, p in zip(meshes, ptrs):
            ctx.voxelize(m["soup"], m["bmin"], m["bmax"], m["mms"], m["voxel_count"], m["voxel_count2"], True,
                         device_ptr=p).release()

    def batch_pass():
        gs = ctx.voxelize_batch(meshes, device_ptrs=ptrs)
        info = gs[0].info()
        for g in gs:
            g.release()
        return info

    for _ in range(warmup):
        plain_pass()
        batch_pass()
    t0 = time.perf_counter()
    for _ in range(steps):
        plain_pass()
    t_plain = (time.perf_counter() - t0) / steps
    lat = []
    binfo = []
    t0 = time.perf_counter()
    for _ in range(steps):
        s0 = time.perf_counter()
        binfo.append(batch_pass())
        lat.append(time.perf_counter() - s0)
    t_batch = (time.perf_counter() - t0) / steps

    # totals and the per-mesh bit-exact check (outside the timed passes)
    gs = ctx.voxelize_batch(meshes, device_ptrs=ptrs)
    tests = 0
    voxels = 0
    tot = {"tests_l1": 0, "tests_l2": 0, "pairs": 0, "nb": 0, "n1": 0}
    exact = True if check else None
    for m, p, g in zip(meshes, ptrs, gs):
        i = g.info()
        tests += i["tests_l1"] + i["tests_l2"]
        voxels += i["n1"] + i["nb"] * i["n2"] ** 3
        for k in tot:
            tot[k] += i[k]
        if check:
            a = g.download()
            ga = ctx.voxelize(m["soup"], m["bmin"], m["bmax"], m["mms"], m["voxel_count"], m["voxel_count2"],
                              True, device_ptr=p)
            b = ga.download()
            ga.release()
            exact = exact and all(np.array_equal(a[k], b[k]) for k in ("l1", "prefix", "bidx", "l2"))
        g.release()
    for p in ptrs:
        ctx.free(p)
    stage_ms = {s: float(np.median([i["ms_stage"][k] for i in binfo])) for k, s in enumerate(STAGES)}
    desc = ("log-uniform 1k-20k; icospheres, tori, drilled boxes; L1 64, L2 4" if small else
            "log-uniform 1k-200k; icospheres, tori, drilled boxes; L1 {64,128,256}, L2 {4,8}")
    return {
        "workload": f"{'small-mesh' if small else 'mixed-size'} stream, {n_meshes} meshes, {T} triangles ({desc})",
        "meshes": n_meshes, "triangles": T,
        "plain_meshes_per_s": n_meshes / t_plain,
        "batched_meshes_per_s": n_meshes / t_batch,
        "speedup": t_plain / t_batch,
        "batched_ms": t_batch * 1e3, "plain_ms": t_plain * 1e3,
        "batched_tri_box_tests_per_s": tests / t_batch,
        "batched_voxels_per_s": voxels / t_batch,
        "batched_stage_ms": stage_ms,
        "totals": tot,
        "bit_exact_vs_plain": exact,
        "tune": tune,
    }


def bench_serve(gv, n_requests, rate, lanes, seed=5, n_library=48, plain_sample=300):
    """The served stream: Poisson arrivals at `rate` requests/s (0: all at
    once, the saturated case), each drawing
    a mesh from a library made by the mixed-stream generators (triangles
    log-uniform in [1k, 2M]) and a resolution (Level1 {64,128,256,512},
    Level2 {4,8,16}).  Library soups are device-resident; results stay
    resident until the client releases them (right after it has waited)."""
    import threading
    from gpview_amd import workloads
    lib_meshes = workloads.mixed_stream(gv, n_library, seed=seed, tri_range=(1000, 2000000))
    rng = np.random.default_rng(seed)
    pick = rng.integers(0, n_library, n_requests)
    l1 = rng.choice([64, 128, 256, 512], n_requests, p=[0.3, 0.3, 0.3, 0.1])
    l2 = rng.choice([4, 8, 16], n_requests, p=[0.45, 0.45, 0.1])
    gaps = rng.exponential(1.0 / rate, n_requests) if rate > 0 else np.zeros(n_requests)
    srv = gv.Server(lanes=lanes, hbm_budget=200e9)
    ctx = gv.Context(0)
    ptrs = [ctx.upload(m["soup"]) for m in lib_meshes]
    ids = [None] * n_requests
    ready = threading.Semaphore(0)
    stats = [None] * n_requests
    tris = np.zeros(n_requests, np.int64)

    def submitter():
        t_next = time.perf_counter()
        for i in range(n_requests):
            t_next += gaps[i]
            d = t_next - time.perf_counter()
            if d > 0:
                time.sleep(d)
            m = dict(lib_meshes[pick[i]], voxel_count=int(l1[i]), voxel_count2=int(l2[i]))
            tris[i] = len(m["soup"])
            ids[i] = srv.submit(m, device_ptr=ptrs[pick[i]])
            ready.release()

    # the same request sequence through the single-mesh path, one call after
    # another (a bounded prefix), for the served-vs-plain comparison
    n_plain = min(n_requests, plain_sample)
    t0 = time.perf_counter()
    for i in range(n_plain):
        m = lib_meshes[pick[i]]
        ctx.voxelize(m["soup"], m["bmin"], m["bmax"], m["mms"], int(l1[i]), int(l2[i]), True,
                     device_ptr=ptrs[pick[i]]).release()
    plain_rate = n_plain / (time.perf_counter() - t0) if n_plain else None
    th = threading.Thread(target=submitter)
    t0 = time.perf_counter()
    th.start()
    for i in range(n_requests):
        ready.acquire()
        _, st = srv.wait(ids[i], timeout=600)
        stats[i] = st
        srv.release(ids[i])
    th.join()
    wall = time.perf_counter() - t0
    c = srv.counters()
    pool = srv.pool_stats()
    srv.close()
    for p in ptrs:
        ctx.free(p)
    ctx.close()
    lat = np.array([s["t_done"] - s["t_submit"] for s in stats])
    span = max(s["t_done"] for s in stats) - min(s["t_submit"] for s in stats)
    sizes = np.array([s["batch_size"] for s in stats])
    pred = np.array([s["predicted_bytes"] for s in stats], np.float64)
    res = np.array([s["resident_bytes"] for s in stats], np.float64)
    return {
        "workload": f"served stream: {n_requests} " + (f"Poisson requests at {rate:g}/s" if rate > 0 else
                                                          "requests submitted at once") + f" over a {n_library}-mesh "
                    "device-resident library (1k-2M triangles), L1 {64,128,256,512}, L2 {4,8,16}",
        "requests": n_requests, "rate_per_s": rate, "lanes": lanes, "hbm_budget_bytes": 200e9,
        "meshes_per_s": n_requests / span, "wall_s": wall, "triangles": int(tris.sum()),
        "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "plain_meshes_per_s": plain_rate, "plain_sample": n_plain,
        "mean_batch": float(sizes.mean()), "batches": c["batches"],
        "peak_charged_bytes": c["peak_charged"],
        "pool": {k: pool[k] for k in ("held_bytes", "allocations", "reuses", "evictions")},
        "predicted_over_resident": float(np.median(pred / np.maximum(res, 1))),
    }


# stage -> its kernel in the single-mesh pipeline (each stage's event pair
# brackets exactly these launches)
STAGE_KERNEL = {"l2_sat": "k_l2_sat_voxel<false>", "l2_parity": "k_l2_parity<false, 8>", "l1_sat": "k_sat_expand<false>"}
PROFILE = os.path.join(ROOT, "profiles", "traffic_torus_knot_plain.json")


def measured(kernel):
    """HBM traffic (calibrated FETCH_SIZE + WRITE_SIZE) and VALU counters of
    `kernel` on this workload, from the committed counter summary."""
    if not kernel or not os.path.exists(PROFILE):
        return None
    with open(PROFILE) as f:
        prof = json.load(f)
    k = prof.get("kernels", {}).get(kernel)
    if not k:
        return None
    return dict(k, source=os.path.relpath(PROFILE, ROOT))


def copy_calibration():
    """One device copy of exactly 1 GiB (read 1 GiB + write 1 GiB): the
    reference point that scales the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": a.numel() * 4}), flush=True)


def kernel_bytes(stage, info, T):
    """Algorithmic HBM bytes of one launch of the stage's main kernel."""
    B, n2, P = info["nb"], info["n2"], info["pairs"]
    n23 = n2 ** 3
    if stage == "l2_sat":
        # bins (4 B/pair) + bin offsets and boundary ids (8 B/voxel) + every
        # triangle once (36 B) + the Level2 boundary bitplane (n2^3/8 B/voxel)
        return 4 * P + 8 * B + 36 * T + B * n23 // 8
    if stage == "l2_parity":
        return 8 * P + 8 * B + 36 * T + B * n23 // 8
    if stage == "l1_sat":
        return 36 * T + 8 * P + (info["n1"] + 7) // 8
    return 36 * T + 2 * ((info["n1"] + 7) // 8)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-threads", type=int, default=0, help="oracle threads (0: all host cores)")
    ap.add_argument("--batch-meshes", type=int, default=256, help="meshes of the mixed-stream batch (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--no-check", action="store_true", help="skip the batched-vs-plain bit-exact check")
    ap.add_argument("--autotune", action="store_true",
                    help="pick launch shapes with gpv_ctx_autotune on each workload before timing it")
    ap.add_argument("--serve-requests", type=int, default=1000, help="served-stream requests (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=100.0, help="served-stream arrivals per second")
    ap.add_argument("--serve-lanes", type=int, default=2)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the 1 GiB device copy that calibrates the traffic counters")
    ap.add_argument("--serve-only", action="store_true",
                    help="only the served stream (--serve-requests at --serve-rate); one JSON line")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (meshes shard "
              "independently across GPUs; see the sharding notes). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import gpview_amd as gv
    if args.serve_only:
        sv = bench_serve(gv, args.serve_requests, args.serve_rate, args.serve_lanes)
        print(json.dumps({"metric": "served meshes/s; p50/p99 request latency", "value": sv["meshes_per_s"],
                          "unit": "meshes/s", "higher_is_better": True, "n_gpus": 1, "served": sv}), flush=True)
        return
    v, f, soup = make_workload(gv)
    T = len(soup)
    bmin, bmax, mms = gv.mesh_bbox(v)
    ctx = gv.Context(0)
    dsoup = ctx.upload(soup)
    tune = None
    if args.autotune:
        shapes, tms = ctx.autotune([{"soup": soup, "bmin": bmin, "bmax": bmax, "mms": mms, "voxel_count": 256,
                                     "voxel_count2": 8}], device_ptrs=[dsoup], repeats=3)
        tune = {"shapes": shapes, "device_ms": tms}

    def step():
        return ctx.voxelize(soup, bmin, bmax, mms, 256, 8, True, device_ptr=dsoup)

    for _ in range(args.warmup):
        step().release()
    lat = []
    infos = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        g = step()
        lat.append(time.perf_counter() - s0)
        infos.append(g.info())
        g.release()
    wall = time.perf_counter() - t0

    # one more run outside the timed region, kept for the correctness check
    g = step()
    info = g.info()
    got = g.download()
    g.release()
    ctx.free(dsoup)
    batched = batched_small = None
    if args.batch_meshes > 0:
        batched = bench_batched(gv, ctx, args.batch_meshes, args.batch_steps, 1, not args.no_check,
                                autotune=args.autotune)
        batched_small = bench_batched(gv, ctx, args.batch_meshes, args.batch_steps, 1, not args.no_check, small=True,
                                      autotune=args.autotune)
    ctx.close()
    served = None
    if args.serve_requests > 0:
        served = bench_serve(gv, args.serve_requests, args.serve_rate, args.serve_lanes)

    meshes_per_s = args.steps / wall
    tests = info["tests_l1"] + info["tests_l2"]
    n2 = info["n2"]
    voxels = info["n1"] + info["nb"] * n2 ** 3
    stage_ms = {s: float(np.median([i["ms_stage"][k] for i in infos])) for k, s in enumerate(STAGES)}
    dom = max(STAGES, key=lambda s: stage_ms[s])
    t_dom = stage_ms[dom] / 1e3
    bytes_dom = kernel_bytes(dom, info, T)
    meas = measured(STAGE_KERNEL.get(dom))
    rl = {"bound": "hbm", "achieved": bytes_dom / t_dom / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
          "frac": bytes_dom / t_dom / 1e9 / HBM_PEAK_GBS,
          "traffic": meas["traffic_bytes"] if meas else None, "kernel": STAGE_KERNEL.get(dom, dom),
          "algorithmic_bytes": bytes_dom, "ms": stage_ms[dom],
          "traffic_source": meas["source"] if meas else None}
    # The path is not a contraction (no MFMA); its binding resource is VALU
    # issue.  Measured from the SQ counters of the same kernel and workload:
    # lane-ops/s = VALU wave instructions x 64 / launch time, against the
    # published FP32 vector peak in lane-operations (FMA counted once).
    valu = None
    if meas and meas.get("valu_insts"):
        lane_ops = meas["valu_insts"] * 64 / t_dom
        valu = {"bound": "valu", "achieved": lane_ops / 1e12, "peak": VALU_PEAK_OPS / 1e12, "unit": "Tlane-op/s",
                "frac": lane_ops / VALU_PEAK_OPS, "kernel": rl["kernel"], "source": meas["source"]}

    cpu = None
    exact = None
    if not args.no_cpu:
        from oracle import orc
        threads = args.cpu_threads or os.cpu_count()
        g0 = orc.grid_setup(bmin, bmax, mms, 256, 8)
        c0 = time.perf_counter()
        want = orc.voxelize(soup, g0, orc.ORC_FAST, threads)
        cdt = time.perf_counter() - c0
        cpu = {"value": 1.0 / cdt, "unit": "meshes/s", "cores": threads, "kind": "port",
               "sample": "the full workload mesh (1 mesh), oracle column-factorised mode",
               "seconds": cdt, "tests_per_s": (want["tests_l1"] + want["tests_l2"]) / cdt,
               "stage_s": dict(zip(["l1_parity", "l1_sat", "compaction", "l2_parity", "l2_sat"],
                                   want["t_stage"]))}
        exact = all(np.array_equal(got[k], want[k]) for k in ("l1", "prefix", "bidx", "l2"))

    out = {
        "metric": "meshes/s & tri-box tests/s at 256^3 L1 + 8^3 L2 (bit-exact); p50/p99 latency",
        "value": meshes_per_s,
        "unit": "meshes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "torus-knot (3,7) tube, 2048x256, 1,048,576 tris, L1 256 nominal, L2 8^3",
                   "triangles": T, "level1": info["n"], "level2": n2, "boundary_voxels": info["nb"],
                   "parallelism": "single GPU"},
        "tri_box_tests_per_s": tests / (wall / args.steps),
        "voxels_per_s": voxels / (wall / args.steps),
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "device_ms_total": float(np.median([i["ms_total"] for i in infos])),
        "stage_ms": stage_ms,
        "tests_l1": info["tests_l1"], "tests_l2": info["tests_l2"], "pairs": info["pairs"],
        "bit_exact_vs_cpu": exact,
        "roofline": rl,
        "roofline_valu": valu,
        "cpu_baseline": cpu,
        "tune": tune,
        "batched": batched,
        "batched_small_meshes": batched_small,
        "served": served,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
