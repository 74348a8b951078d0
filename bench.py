"""Benchmark of the voxelizer on one MI355X.

Headline (value): the plain path, one gpv_voxelize call per step on the
seeded (3,7) torus-knot tube, 2,048 segments x 256 ring vertices
(1,048,576 FP32 triangles) scaled to 0.9 of the unit cube, Level1
voxelCount 256, Level2 8^3 per boundary voxel, the soup resident in HBM.
One step = all stages, counts and the final synchronisation.

Prints ONE JSON line with, beside the headline:
- tri_box_tests_per_s: the tri-box tests/s metric as defined on the
  reference's loop counts (every voxel of every triangle's index AABB at
  Level1, |bin| * n2^3 at Level2); the kernels execute a pruned subset whose
  outcome is the same, so the figure is reference-equivalent;
- pcie_inclusive: the headline step with the soup uploaded from
  page-locked host memory in every step;
- stencil_fill: the headline step with Level1 inside/outside from the
  GL-stencil-equivalent fill (gpv_params.l1_stencil) instead of ray parity;
- roofline (HBM) and roofline_valu of the dominant kernel, from HIP events
  of this run and the committed counter summary of the same kernel;
- cpu_baseline: the oracle in the CPU baseline's composition (column-
  factorised parity, literal SAT loops) on the full mesh over all host
  cores, with the single-thread run and the literal all-triangle
  composition on the 5,120-triangle icosphere beside it;
- batched: 256 meshes of the mixed-size stream, one gpv_voxelize per mesh
  against one gpv_voxelize_batch, with the soups uploaded from pinned host
  memory inside the timed region (and resident, for comparison);
- stress: members of the 64-mesh exactness-stress stream at Level1 512 /
  Level2 16, batched in runs, soups uploaded in the timed region, with a
  bit-exact check against the oracle on a stated sample;
- served: 10,000 Poisson requests through the serving loop at a stated rate
  (0.8 of the measured saturated capacity), soups uploaded per request,
  p50/p99 latency, beside the saturated capacity itself.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] ...
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
# FP32 vector issue without packed instructions (the library is built with
# packed FP32 off), measured on the library's probe kernel of known VALU
# count (bench.py --valu-probe; one wave64 instruction per 2 cycles per SIMD);
# the committed counter summary carries the same figure
VALU_PEAK_OPS = 77.1e12
STAGES = ["setup", "l1_sat", "l1_parity", "compaction_bins", "l2_parity", "l2_sat"]
TESTS_BASIS = ("reference loop counts: sum of index-AABB volumes (Level1) + |bin| * n2^3 (Level2); "
               "the kernels run a pruned subset with the same outcome")


def stage_ms(infos):
    return {s: float(np.median([i["ms_stage"][k] for i in infos])) for k, s in enumerate(STAGES)}


def bench_batched(gv, ctx, meshes, steps, warmup, check, desc):
    """Plain (one gpv_voxelize per mesh, resident soups) against
    gpv_voxelize_batch, the latter both from pinned host soups (upload in the
    timed region) and from resident soups."""
    ptrs = [ctx.upload(m["soup"]) for m in meshes]
    pinned = [gv.pinned_copy(m["soup"]) for m in meshes]
    host = [dict(m, soup=p.array) for m, p in zip(meshes, pinned)]
    T = sum(len(m["soup"]) for m in meshes)
    n = len(meshes)

    def plain_pass():
        for m, p in zip(meshes, ptrs):
            ctx.voxelize(m["soup"], m["bmin"], m["bmax"], m["mms"], m["voxel_count"], m["voxel_count2"], True,
                         device_ptr=p).release()

    def batch_pass(resident):
        gs = ctx.voxelize_batch(meshes, device_ptrs=ptrs) if resident else ctx.voxelize_batch(host)
        info = gs[0].info()
        for g in gs:
            g.release()
        return info

    def timed(fn, *a):
        t0 = time.perf_counter()
        out = [fn(*a) for _ in range(steps)]
        return (time.perf_counter() - t0) / steps, out

    for _ in range(warmup):
        plain_pass()
        batch_pass(False)
        batch_pass(True)
    t_plain, _ = timed(plain_pass)
    t_host, _ = timed(batch_pass, False)
    # stage and device times from the resident-soup batch: one run over every
    # mesh (from host soups a large batch runs in two, its uploads overlapped)
    t_res, binfo = timed(batch_pass, True)

    # totals and the per-mesh bit-exact check against the single-mesh path
    gs = ctx.voxelize_batch(host)
    tests = voxels = 0
    tot = {"tests_l1": 0, "tests_l2": 0, "pairs": 0, "nb": 0, "n1": 0}
    exact = True if check else None
    for m, p, g in zip(meshes, ptrs, gs):
        i = g.info()
        tests += i["tests_l1"] + i["tests_l2"]
        voxels += i["n1"] + i["nb"] * i["n2"] ** 3
        for k in tot:
            tot[k] += i[k]
        if check:
            a = g.download_bits()
            ga = ctx.voxelize(m["soup"], m["bmin"], m["bmax"], m["mms"], m["voxel_count"], m["voxel_count2"],
                              True, device_ptr=p)
            b = ga.download_bits()
            ga.release()
            exact = exact and all(np.array_equal(x, y) for x, y in zip(a, b))
        g.release()
    for p in ptrs:
        ctx.free(p)
    for p in pinned:
        p.close()
    return {
        "workload": f"{desc}: {n} meshes, {T} triangles",
        "meshes": n, "triangles": T, "soup_bytes": 36 * T,
        "plain_meshes_per_s": n / t_plain,
        "batched_meshes_per_s": n / t_host,
        "batched_resident_meshes_per_s": n / t_res,
        "speedup": t_plain / t_host,
        "speedup_resident": t_plain / t_res,
        "plain_ms": t_plain * 1e3, "batched_ms": t_host * 1e3, "batched_resident_ms": t_res * 1e3,
        "batched_tri_box_tests_per_s": tests / t_host,
        "batched_voxels_per_s": voxels / t_host,
        "batched_stage_ms": stage_ms(binfo),
        "batched_device_ms": float(np.median([i["ms_total"] for i in binfo])),
        "totals": tot,
        "bit_exact_vs_plain": exact,
        "note": "plain uses resident soups; batched_meshes_per_s uploads every soup from pinned host memory "
                "inside the timed region; batched_stage_ms and batched_device_ms are those of the resident-soup "
                "batch (one run)",
    }


def bench_stress(gv, ctx, n_meshes, per_run, check_members):
    """Stress members 0..n_meshes-1 (100k-2M triangles, lattice-snapped /
    sliver / face-coplanar) at Level1 512 / Level2 16: gpv_voxelize_batch
    over runs of per_run meshes from pinned host soups, grids released after
    each run; the plain path mesh by mesh beside it.  Returns the summary and
    the batched results of the sampled members (for the CPU leg's check)."""
    from gpview_amd import workloads
    t0 = time.perf_counter()
    meshes = [workloads.stress_member(gv, i) for i in range(n_meshes)]
    t_gen = time.perf_counter() - t0
    pinned = [gv.pinned_copy(m["soup"]) for m in meshes]
    host = [dict(m, soup=p.array) for m, p in zip(meshes, pinned)]
    T = sum(len(m["soup"]) for m in meshes)
    runs = [host[a:a + per_run] for a in range(0, n_meshes, per_run)]
    ctx.voxelize_batch(runs[0][:1])[0].release()   # warm-up
    infos = []
    t0 = time.perf_counter()
    for r in runs:
        gs = ctx.voxelize_batch(r)
        infos += [g.info() for g in gs]
        for g in gs:
            g.release()
    t_batch = time.perf_counter() - t0
    t0 = time.perf_counter()
    for m in host:
        ctx.voxelize(m["soup"], m["bmin"], m["bmax"], m["mms"], 512, 16).release()
    t_plain = time.perf_counter() - t0
    tests = sum(i["tests_l1"] + i["tests_l2"] for i in infos)
    voxels = sum(i["n1"] + i["nb"] * 16 ** 3 for i in infos)
    samples = []
    for k in check_members:
        gs = ctx.voxelize_batch(host[(k // per_run) * per_run:(k // per_run + 1) * per_run])
        samples.append((k, meshes[k], gs[k % per_run].download()))
        for x in gs:
            x.release()
    for p in pinned:
        p.close()
    return {
        "workload": f"exactness-stress stream, members 0-{n_meshes - 1} of 64 ({T} triangles, 100k-2M each; "
                    "lattice-snapped p=0.3, 10% slivers, 5% face-coplanar), Level1 512, Level2 16",
        "meshes": n_meshes, "triangles": T, "per_run": per_run,
        "batched_meshes_per_s": n_meshes / t_batch, "plain_meshes_per_s": n_meshes / t_plain,
        "batched_s": t_batch, "plain_s": t_plain,
        "batched_tri_box_tests_per_s": tests / t_batch, "batched_voxels_per_s": voxels / t_batch,
        "boundary_voxels": int(sum(i["nb"] for i in infos)),
        "generation_s": t_gen,
        "note": "soups uploaded from pinned host memory inside the timed region; grids released after each run; "
                "sampled members are checked against the oracle in cpu_baseline.stress_sample",
    }, samples


def bench_serve(gv, n_requests, rate, lanes, seed=5, n_library=48, plain_sample=300):
    """The served stream: requests drawn from a library of the torus knot and
    mixed-stream meshes of 1k-2M triangles, each at Level1 {64,128,256,512}
    and Level2 {4,8,16}; soups in pinned host memory, uploaded per request
    (tris_on_device = 0); results released by the client right after its
    wait.  rate 0 submits everything at once (saturated)."""
    import threading
    from gpview_amd import workloads
    lib_meshes = workloads.served_library(gv, n_library, seed=seed)
    pinned = [gv.pinned_copy(m["soup"]) for m in lib_meshes]
    lib_host = [dict(m, soup=p.array) for m, p in zip(lib_meshes, pinned)]
    pick, l1, l2, gaps = workloads.served_requests(n_requests, n_library, rate, seed=seed)
    srv = gv.Server(lanes=lanes, hbm_budget=200e9)
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
            m = dict(lib_host[pick[i]], voxel_count=int(l1[i]), voxel_count2=int(l2[i]))
            tris[i] = len(m["soup"])
            ids[i] = srv.submit(m)
            ready.release()

    plain_rate = None
    if plain_sample:
        # the same request sequence through the single-mesh path, one call
        # after another (a bounded prefix), for the served-vs-plain ratio
        ctx = gv.Context(0)
        n_plain = min(n_requests, plain_sample)
        t0 = time.perf_counter()
        for i in range(n_plain):
            m = lib_host[pick[i]]
            ctx.voxelize(m["soup"], m["bmin"], m["bmax"], m["mms"], int(l1[i]), int(l2[i]), True).release()
        plain_rate = n_plain / (time.perf_counter() - t0)
        ctx.close()
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
    for p in pinned:
        p.close()
    lat = np.array([s["t_done"] - s["t_submit"] for s in stats])
    span = max(s["t_done"] for s in stats) - min(s["t_submit"] for s in stats)
    sizes = np.array([s["batch_size"] for s in stats])
    batches = {}
    for s in stats:
        batches[s["batch_id"]] = (s["batch_peak_bytes"], s["batch_predicted_bytes"])
    ratio = [pk / pr for pk, pr in batches.values() if pr > 0]
    return {
        "workload": f"served stream: {n_requests} " + (f"Poisson requests at {rate:g}/s" if rate > 0 else
                                                          "requests submitted at once") +
                    f" over a {n_library}-mesh library (the torus knot + mixed meshes of 1k-2M triangles), "
                    "Level1 {64,128,256,512}, Level2 {4,8,16}; soups uploaded per request from pinned host memory",
        "requests": n_requests, "rate_per_s": rate, "lanes": lanes, "hbm_budget_bytes": 200e9,
        "meshes_per_s": n_requests / span, "wall_s": wall, "triangles": int(tris.sum()),
        "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "plain_meshes_per_s": plain_rate, "plain_sample": min(n_requests, plain_sample),
        "mean_batch": float(sizes.mean()), "max_batch": int(sizes.max()), "batches": c["batches"],
        "readmitted": c["readmitted"], "failed": int(sum(s["status"] != 0 for s in stats)),
        "peak_charged_bytes": c["peak_charged"],
        "batch_peak_over_predicted": {"max": float(max(ratio)), "median": float(np.median(ratio))} if ratio else None,
        "pool": {k: pool[k] for k in ("held_bytes", "allocations", "reuses", "evictions")},
    }


# stage -> its kernel family in the single-mesh pipeline (each stage's event
# pair brackets exactly these launches); the template arguments after the
# first are the launch shape of the context's tuning
STAGE_KERNEL = {"l2_sat": "k_l2_sat_voxel<false", "l2_parity": "k_l2_parity<false", "l1_sat": "k_sat_expand<false"}
PROFILE = os.path.join(ROOT, "profiles", "traffic_torus_knot_plain.json")


def measured(family):
    """HBM traffic (calibrated FETCH_SIZE + WRITE_SIZE) and VALU counters of
    the instantiation of kernel `family` that ran on this workload, from the
    committed counter summary."""
    if not family or not os.path.exists(PROFILE):
        return None
    with open(PROFILE) as f:
        prof = json.load(f)
    ks = prof.get("kernels", {})
    names = [n for n in ks if n.startswith(family + ",") or n.startswith(family + ">")]
    if not names:
        return None
    name = max(names, key=lambda n: ks[n].get("avg_us_trace", 0.0))
    return dict(ks[name], name=name, source=os.path.relpath(PROFILE, ROOT), valu_peak=prof.get("valu_peak_lane_ops"))


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


def valu_probe():
    """The FP32 issue rate measured on a kernel of known VALU instruction
    count (gpv_valu_probe): wave instructions / device time, x 64 lanes."""
    import gpview_amd as gv
    out = {"note": "gpv_valu_probe: 256-thread blocks, 32 v_add_f32 per loop iteration per thread"}
    with gv.Context(0) as c:
        for blocks, iters in ((4096, 4096), (16384, 2048), (65536, 512)):
            c.valu_probe(blocks, iters)
            best = None
            for _ in range(3):
                ms, wi = c.valu_probe(blocks, iters)
                best = ms if best is None or ms < best else best
            out[f"{blocks}x{iters}"] = {"ms": best, "wave_insts": wi, "wave_insts_per_s": wi / (best / 1e3),
                                        "lane_ops_per_s": wi * 64 / (best / 1e3)}
    out["peak_lane_ops_per_s"] = max(v["lane_ops_per_s"] for k, v in out.items() if isinstance(v, dict))
    print(json.dumps(out), flush=True)


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


def progress(msg):
    """One line per finished leg on stderr, so a long run shows its steps."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(gv, orc, soup, bmin, bmax, mms, threads, mixed=None, stress_samples=()):
    """The oracle on this host: the CPU baseline's composition on the full
    headline mesh with all cores (value) and one thread; the literal
    composition (all-triangle parity) against it on the 5,120-triangle
    icosphere; the baseline composition on a stated subset of the mixed
    stream; and the oracle's fast mode on sampled stress members, whose
    results double as the bit-exact check of the GPU's batched grids."""
    from gpview_amd import workloads
    g0 = orc.grid_setup(bmin, bmax, mms, 256, 8)

    def run(tris, g, mode, th):
        c0 = time.perf_counter()
        r = orc.voxelize(tris, g, mode, th, bins=False)
        return time.perf_counter() - c0, r

    t_all, want = run(soup, g0, orc.ORC_PLAN, threads)
    progress(f"cpu baseline, {threads} threads: {t_all:.2f} s")
    t_one, _ = run(soup, g0, orc.ORC_PLAN, 1)
    progress(f"cpu baseline, 1 thread: {t_one:.2f} s")
    s = workloads.plain_small(gv)
    g1 = orc.grid_setup(s["bmin"], s["bmax"], s["mms"], 64, 4)
    legs = {}
    for name, mode in (("literal", orc.ORC_EXACT), ("column_factorised", orc.ORC_PLAN)):
        for th in (threads, 1):
            t, _ = run(s["soup"], g1, mode, th)
            legs[f"icosphere_5120_{name}_{th}t"] = {"seconds": t, "meshes_per_s": 1.0 / t, "cores": th}
            progress(f"cpu icosphere {name}, {th} threads: {t:.2f} s")
    mixed_leg = None
    if mixed:
        c0 = time.perf_counter()
        for m in mixed:
            gm = orc.grid_setup(m["bmin"], m["bmax"], m["mms"], m["voxel_count"], m["voxel_count2"])
            orc.voxelize(m["soup"], gm, orc.ORC_PLAN, threads, bins=False)
        tm = time.perf_counter() - c0
        mixed_leg = {"meshes": len(mixed), "triangles": int(sum(len(m["soup"]) for m in mixed)),
                     "seconds": tm, "meshes_per_s": len(mixed) / tm, "cores": threads,
                     "sample": f"members 0-{len(mixed) - 1} of the 256-mesh mixed stream, baseline composition"}
    stress_leg = []
    for k, m, got in stress_samples:
        gs = orc.grid_setup(m["bmin"], m["bmax"], m["mms"], 512, 16)
        ts, ws = run(m["soup"], gs, orc.ORC_FAST, threads)
        ok = all(np.array_equal(got[key], ws[key]) for key in ("l1", "prefix", "bidx", "l2"))
        progress(f"cpu stress member {k}: {ts:.2f} s, bit-exact {ok}")
        stress_leg.append({"member": k, "triangles": len(m["soup"]), "nb": int(ws["nb"]), "seconds": ts,
                           "cores": threads, "mode": "oracle fast mode (factored SAT sweeps)",
                           "gpu_batched_bit_exact": ok})
        del ws
    tests = want["tests_l1"] + want["tests_l2"]
    out = {"value": 1.0 / t_all, "unit": "meshes/s", "cores": threads, "kind": "port",
           "sample": "the full headline mesh (1 mesh): oracle in the baseline composition (column-factorised "
                     "parity, literal SAT loops over every index-AABB voxel and bin), -O2 -ffp-contract=off",
           "seconds": t_all, "tri_box_tests_per_s": tests / t_all,
           "single_thread": {"value": 1.0 / t_one, "seconds": t_one, "cores": 1,
                             "tri_box_tests_per_s": tests / t_one},
           "icosphere_5120_64_4": legs,
           "mixed_subset": mixed_leg,
           "stress_sample": stress_leg,
           "stage_s": dict(zip(["l1_parity", "l1_sat", "compaction", "l2_parity", "l2_sat"], want["t_stage"]))}
    return out, want


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
    ap.add_argument("--stress-meshes", type=int, default=16, help="stress-stream members (0: skip; 64: all)")
    ap.add_argument("--stress-run", type=int, default=8, help="stress meshes per gpv_voxelize_batch")
    ap.add_argument("--serve-requests", type=int, default=10000, help="served-stream requests (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=-1.0,
                    help="served-stream arrivals per second (-1: 0.8 of the measured saturated capacity)")
    ap.add_argument("--serve-lanes", type=int, default=2)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the 1 GiB device copy that calibrates the traffic counters")
    ap.add_argument("--serve-only", action="store_true", help="only the served stream; one JSON line")
    ap.add_argument("--valu-probe", action="store_true",
                    help="only measure the FP32 VALU issue rate on the library's probe kernel")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.valu_probe:
        valu_probe()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (meshes shard "
              "independently across GPUs; see the sharding notes). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import gpview_amd as gv
    from gpview_amd import workloads

    def served():
        if args.serve_requests <= 0:
            return None, None
        sat = bench_serve(gv, min(2000, args.serve_requests), 0.0, args.serve_lanes, plain_sample=300)
        rate = args.serve_rate if args.serve_rate >= 0 else round(0.8 * sat["meshes_per_s"], -1)
        return bench_serve(gv, args.serve_requests, rate, args.serve_lanes, plain_sample=0), sat

    if args.serve_only:
        sv, sat = served()
        print(json.dumps({"metric": "served meshes/s; p50/p99 request latency", "value": sv["meshes_per_s"],
                          "unit": "meshes/s", "higher_is_better": True, "n_gpus": 1, "served": sv,
                          "served_saturated": sat}), flush=True)
        return
    m = workloads.plain_large(gv)
    soup, bmin, bmax, mms = m["soup"], m["bmin"], m["bmax"], m["mms"]
    T = len(soup)
    ctx = gv.Context(0)
    dsoup = ctx.upload(soup)

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
    progress(f"headline: {args.steps} steps, {wall / args.steps * 1e3:.3f} ms per step")

    # the same step with the soup handed over in page-locked host memory:
    # the PCIe-inclusive rate, reported beside the headline, never as its value
    pin = gv.pinned_copy(soup)
    for _ in range(2):
        ctx.voxelize(pin.array, bmin, bmax, mms, 256, 8, True).release()
    n_pc = max(5, args.steps // 2)
    t1 = time.perf_counter()
    for _ in range(n_pc):
        ctx.voxelize(pin.array, bmin, bmax, mms, 256, 8, True).release()
    t_pc = time.perf_counter() - t1
    pin.close()
    pcie = {"meshes_per_s": n_pc / t_pc, "ms_per_step": t_pc / n_pc * 1e3, "steps": n_pc,
            "soup_bytes": int(soup.nbytes), "note": "soup uploaded from page-locked host memory in every step"}

    # the same step with Level1 inside/outside from the GL-stencil-equivalent
    # fill (gpv_params.l1_stencil) in place of the ray parity
    def stencil_step():
        return ctx.voxelize(soup, bmin, bmax, mms, 256, 8, True, device_ptr=dsoup, l1_stencil=True)

    for _ in range(2):
        stencil_step().release()
    n_st = max(5, args.steps // 2)
    st_infos = []
    t2 = time.perf_counter()
    for _ in range(n_st):
        gs = stencil_step()
        st_infos.append(gs.info())
        gs.release()
    t_st = time.perf_counter() - t2

    # one more run outside the timed region, kept for the correctness check
    g = step()
    info = g.info()
    got = g.download()
    g.release()
    ctx.free(dsoup)
    stencil = {"meshes_per_s": n_st / t_st, "ms_per_step": t_st / n_st * 1e3, "steps": n_st,
               "l1_parity_stage_ms": float(np.median([i["ms_stage"][2] for i in st_infos])),
               "l1_inside": int(st_infos[-1]["l1_inside"]), "l1_inside_ray_parity": int(info["l1_inside"]),
               "note": "gpv_params.l1_stencil: Level1 inside/outside by the GL-stencil-equivalent fill "
                       "(ClassifyInOut2x) on the headline mesh, resident soup; the headline uses the ray parity"}
    progress(f"stencil fill: {stencil['ms_per_step']:.3f} ms per step")
    batched = batched_small = stress = None
    if args.batch_meshes > 0:
        batched = bench_batched(gv, ctx, workloads.mixed_stream(gv, args.batch_meshes), args.batch_steps, 1,
                                not args.no_check,
                                "mixed-size stream (log-uniform 1k-200k triangles; icospheres, tori, drilled "
                                "boxes; Level1 {64,128,256}, Level2 {4,8})")
        batched_small = bench_batched(
            gv, ctx, workloads.mixed_stream(gv, args.batch_meshes, seed=7, tri_range=(1000, 20000), level1=(64,),
                                            level2=(4,)), args.batch_steps, 1, False,
            "small-mesh stream (log-uniform 1k-20k triangles, Level1 64, Level2 4)")
        progress(f"batched: {batched['batched_meshes_per_s']:.0f} meshes/s (plain {batched['plain_meshes_per_s']:.0f})")
    stress_samples = []
    if args.stress_meshes > 0:
        n = min(args.stress_meshes, 64)
        sizes = [workloads.stress_spec(i)[1] for i in range(n)]
        sample = sorted({int(np.argmin(sizes)), int(np.argmax(sizes))}) if not args.no_cpu else []
        stress, stress_samples = bench_stress(gv, ctx, n, args.stress_run, sample)
        progress(f"stress: {stress['batched_meshes_per_s']:.1f} meshes/s")
    ctx.close()
    served_stream, served_sat = served()
    if served_stream:
        progress(f"served: {served_stream['meshes_per_s']:.0f} meshes/s at {served_stream['rate_per_s']:.0f}/s")

    meshes_per_s = args.steps / wall
    tests = info["tests_l1"] + info["tests_l2"]
    n2 = info["n2"]
    voxels = info["n1"] + info["nb"] * n2 ** 3
    sms = stage_ms(infos)
    # the dominant kernel: the Level2 SAT has the most device time in the
    # committed trace of this workload, and its stage's event pair brackets
    # exactly its one launch (the Level2 parity stage spans two kernels and
    # the column sort's wait, so its time is not one kernel's)
    dom = "l2_sat"
    t_dom = sms[dom] / 1e3
    bytes_dom = kernel_bytes(dom, info, T)
    meas = measured(STAGE_KERNEL.get(dom))
    rl = {"bound": "hbm", "achieved": bytes_dom / t_dom / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
          "frac": bytes_dom / t_dom / 1e9 / HBM_PEAK_GBS,
          "traffic": meas["traffic_bytes"] if meas else None,
          "kernel": meas["name"] if meas else STAGE_KERNEL.get(dom, dom),
          "algorithmic_bytes": bytes_dom, "ms": sms[dom],
          "traffic_source": meas["source"] if meas else None,
          "trace_avg_us": meas.get("avg_us_trace") if meas else None,
          "alone_us": meas.get("avg_us_alone") if meas else None}
    # The path is not a contraction (no MFMA); its binding resource is VALU
    # issue: VALU wave instructions of the same kernel and workload (SQ
    # counters, committed summary) x 64 lanes / this run's launch time,
    # against the non-packed FP32 issue peak.
    valu = None
    if meas and meas.get("valu_insts"):
        peak = meas.get("valu_peak") or VALU_PEAK_OPS
        lane_ops = meas["valu_insts"] * 64 / t_dom
        valu = {"bound": "valu", "achieved": lane_ops / 1e12, "peak": peak / 1e12, "unit": "Tlane-op/s",
                "frac": lane_ops / peak, "kernel": rl["kernel"], "source": meas["source"],
                "valu_insts": meas["valu_insts"]}

    cpu = None
    exact = None
    if not args.no_cpu:
        from oracle import orc
        threads = args.cpu_threads or os.cpu_count()
        cpu, want = cpu_baseline(gv, orc, soup, bmin, bmax, mms, threads,
                                 mixed=workloads.mixed_stream(gv, 16) if args.batch_meshes > 0 else None,
                                 stress_samples=stress_samples)
        stress_samples = None
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
        "tri_box_tests_basis": TESTS_BASIS,
        "voxels_per_s": voxels / (wall / args.steps),
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "device_ms_total": float(np.median([i["ms_total"] for i in infos])),
        "stage_ms": sms,
        "tests_l1": info["tests_l1"], "tests_l2": info["tests_l2"], "pairs": info["pairs"],
        "bit_exact_vs_cpu": exact,
        "pcie_inclusive": pcie,
        "stencil_fill": stencil,
        "roofline": rl,
        "roofline_valu": valu,
        "cpu_baseline": cpu,
        "batched": batched,
        "batched_small_meshes": batched_small,
        "stress": stress,
        "served": served_stream,
        "served_saturated": served_sat,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
