"""Writes the golden vectors for the two geometric predicates of the voxelizer.

The expected outputs come from the reference's own source files
(src/TriBoxIntersection.cpp, src/TriRayIntersection.cpp), compiled unmodified
into oracle/_ref/libgpvref.so by oracle/Makefile (g++ -O2 -ffp-contract=off,
x86-64 SSE2, the reference's Release flags).  Inputs are seeded and mix random
cases with the adversarial families that bit-exactness is fragile on.

Run in a tree where /root/reference is mounted:
    make -C oracle && python tests/golden/make_golden.py
Outputs (committed): tests/golden/sat_cases.npz, tests/golden/mt_cases.npz,
tests/golden/normal_cases.npz (from includes/FloatVector.h through
oracle/ref_vec_shim.cpp), tests/golden/sat_nonfinite.npz (SAT on inputs
outside the pipeline's domain: NaN, infinities, overflowing products),
tests/golden/manifest.json.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import orc  # noqa: E402

F32 = np.float32


def _f(a):
    return np.asarray(a, dtype=F32)


def sat_cases(rng):
    cases = {}
    n = 12000
    # random: triangles scattered around boxes of mixed size
    c = _f(rng.uniform(-1, 1, (n, 3)))
    h = _f(rng.uniform(1e-3, 0.2, (n, 3)))
    tv = _f(c[:, None, :] + rng.normal(0, 1, (n, 3, 3)) * h[:, None, :] * rng.uniform(0.5, 4, (n, 1, 1)))
    cases["random"] = (c, h, tv.reshape(n, 9))
    # vertices snapped onto the box face planes c +- h (touching contacts)
    c = _f(rng.uniform(0, 1, (n, 3)))
    h = _f(rng.uniform(1e-3, 0.05, (n, 3)))
    tv = _f(c[:, None, :] + rng.uniform(-2, 2, (n, 3, 3)) * h[:, None, :])
    snap = rng.random((n, 3, 3)) < 0.5
    side = np.where(rng.random((n, 3, 3)) < 0.5, -1, 1)
    planes = _f(c[:, None, :] + side * h[:, None, :])
    tv = np.where(snap, planes, tv).astype(F32)
    cases["face_snapped"] = (c, h, tv.reshape(n, 9))
    # a lattice like the voxel grid: boxes at (i+0.5)*g + min, vertices on
    # lattice lines (i*g + min) and at box centers (gives exact +-0 terms)
    g = F32(1.0 / 64)
    mn = F32(-0.0123)
    idx = rng.integers(0, 64, (n, 3))
    c = _f((idx + 0.5) * np.float64(g) + np.float64(mn))
    h = np.full((n, 3), F32(g / 2), F32)
    vi = idx[:, None, :] + rng.integers(-2, 3, (n, 3, 3))
    lat = _f(vi * g + mn)
    ctr = _f((vi + 0.5) * np.float64(g) + np.float64(mn))
    pick = rng.integers(0, 3, (n, 3, 3))
    off = _f(rng.uniform(-1, 1, (n, 3, 3)) * g)
    tv = np.where(pick == 0, lat, np.where(pick == 1, ctr, _f(lat + off))).astype(F32)
    cases["lattice"] = (c, h, tv.reshape(n, 9))
    # slivers: aspect ratio 1e3..1e6, random orientation near the box
    c = _f(rng.uniform(0, 1, (n, 3)))
    h = _f(rng.uniform(1e-3, 0.02, (n, 3)))
    a = c + rng.normal(0, 1, (n, 3)) * h
    d1 = rng.normal(0, 1, (n, 3))
    d1 /= np.linalg.norm(d1, axis=1, keepdims=True)
    L = rng.uniform(1, 20, (n, 1)) * h.max(1, keepdims=True)
    b = a + d1 * L
    w = L / 10 ** rng.uniform(3, 6, (n, 1))
    d2 = rng.normal(0, 1, (n, 3))
    cc = a + d1 * L * rng.uniform(0, 1, (n, 1)) + d2 * w
    tv = _f(np.stack([a, b, cc], 1))
    cases["sliver"] = (c, h, tv.reshape(n, 9))
    # axis-parallel and face-coplanar triangles (one coordinate shared)
    c = _f(rng.uniform(0, 1, (n, 3)))
    h = _f(rng.uniform(1e-3, 0.05, (n, 3)))
    tv = _f(c[:, None, :] + rng.uniform(-2, 2, (n, 3, 3)) * h[:, None, :])
    ax = rng.integers(0, 3, n)
    val = np.where(rng.random(n) < 0.5, c[np.arange(n), ax] + h[np.arange(n), ax] * rng.choice([-1, 1], n),
                   tv[np.arange(n), 0, ax])
    tv[np.arange(n), :, ax] = _f(val)[:, None]
    cases["axis_plane"] = (c, h, tv.reshape(n, 9))
    # degenerate triangles: repeated vertices, collinear, all equal
    c = _f(rng.uniform(0, 1, (n, 3)))
    h = _f(rng.uniform(1e-3, 0.05, (n, 3)))
    p = _f(c[:, None, :] + rng.uniform(-1.5, 1.5, (n, 3, 3)) * h[:, None, :])
    kind = rng.integers(0, 3, n)
    p[kind == 0, 1] = p[kind == 0, 0]
    t = rng.uniform(-1, 2, (n, 1))
    col = _f(p[:, 0] + t * (p[:, 1] - p[:, 0]))
    p[kind == 1, 2] = col[kind == 1]
    p[kind == 2, 1] = p[kind == 2, 0]
    p[kind == 2, 2] = p[kind == 2, 0]
    cases["degenerate"] = (c, h, p.reshape(n, 9))
    return cases


def mt_cases(rng):
    cases = {}
    n = 12000
    D = np.tile(_f([0, 0, 1]), (n, 1))
    # random +z rays
    o = _f(rng.uniform(0, 1, (n, 3)))
    tv = _f(o[:, None, :] + rng.normal(0, 0.1, (n, 3, 3)) + np.array([0, 0, 0.05]))
    cases["random_z"] = (tv.reshape(n, 9), o, D)
    # |det| near 1e-6: xy-projected area scaled to det in [0.3e-6, 3e-6]
    o = _f(rng.uniform(0, 1, (n, 3)))
    base = rng.normal(0, 1, (n, 3, 3)) * 0.01
    e1 = base[:, 1, :2] - base[:, 0, :2]
    e2 = base[:, 2, :2] - base[:, 0, :2]
    det = np.abs(e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]) + 1e-30
    target = 10 ** rng.uniform(np.log10(0.3e-6), np.log10(3e-6), n)
    s = np.sqrt(target / det)
    xy = base[:, :, :2] * s[:, None, None]
    tv = np.concatenate([xy, base[:, :, 2:] + o[:, None, 2:] + 0.05], 2)
    wb = rng.dirichlet([2, 2, 2], n)
    tv[:, :, :2] += o[:, None, :2] - (wb[:, :, None] * xy).sum(1)[:, None, :]
    cases["det_edge"] = (_f(tv).reshape(n, 9), o, D)
    # t near 1e-6: ray origin just below/above the plane hit point
    o = _f(rng.uniform(0, 1, (n, 3)))
    tri = o[:, None, :] + rng.normal(0, 0.05, (n, 3, 3))
    w = rng.dirichlet([1, 1, 1], n)
    hit = (w[:, :, None] * tri).sum(1)
    tri[:, :, :2] += o[:, None, :2] - hit[:, None, :2]
    hz = (w[:, :, None] * tri).sum(1)[:, 2]
    dz = rng.uniform(-3e-6, 3e-6, n)
    o[:, 2] = _f(hz - 1e-6 + dz)
    cases["t_edge"] = (_f(tri).reshape(n, 9), o, D)
    # ray origin on a vertex / edge / edge extension (u, v, u+v on 0 or 1)
    tri = _f(rng.uniform(0, 1, (n, 3, 3)))
    tri[:, :, 2] += 1.0
    which = rng.integers(0, 6, n)
    a_, b_ = rng.integers(0, 3, n), rng.integers(0, 3, n)
    fr = rng.uniform(-0.5, 1.5, n)
    pa = tri[np.arange(n), a_]
    pb = tri[np.arange(n), b_]
    pt = np.where((which < 3)[:, None], pa, pa + fr[:, None] * (pb - pa))
    o = _f(np.concatenate([pt[:, :2], np.zeros((n, 1))], 1))
    cases["on_edge"] = (tri.reshape(n, 9), o, D)
    # slivers seen nearly edge-on by the ray (large edges, tiny projection)
    o = _f(rng.uniform(0, 1, (n, 3)))
    a = o + rng.normal(0, 0.3, (n, 3))
    d1 = rng.normal(0, 1, (n, 3))
    d2 = rng.normal(0, 1, (n, 3))
    d2[:, :2] = d1[:, :2] * rng.uniform(-2, 2, (n, 1)) + rng.normal(0, 1, (n, 2)) * 10 ** rng.uniform(-7, -4, (n, 1))
    tv = _f(np.stack([a, a + d1 * 0.5, a + d2 * 0.5], 1))
    o2 = o.copy()
    o2[:, :2] = _f(a[:, :2] + rng.uniform(-0.3, 1.3, (n, 1)) * d1[:, :2] * 0.5)
    o2[:, 2] = _f(np.minimum(tv[:, :, 2].min(1), o[:, 2]) - 0.1)
    cases["sliver_edge_on"] = (tv.reshape(n, 9), _f(o2), D)
    # general directions (the reference routine takes any D)
    o = _f(rng.uniform(0, 1, (n, 3)))
    d = rng.normal(0, 1, (n, 3))
    d = _f(d / np.linalg.norm(d, axis=1, keepdims=True))
    tv = _f(o[:, None, :] + d[:, None, :] * rng.uniform(0, 0.5, (n, 1, 1)) + rng.normal(0, 0.1, (n, 3, 3)))
    cases["any_dir"] = (tv.reshape(n, 9), o, d)
    return cases


def nonfinite_cases(rng):
    """SAT inputs whose projections are NaN or infinite: vertex coordinates
    replaced by NaN, +-inf or values whose products overflow (inf - inf is
    NaN), around finite boxes.  The reference's edge macros order the two
    projections per macro (AXISTEST_Z12 compares p2 before p1), and with a
    NaN among them that order decides the interval, so these cases fix the
    operand order of every edge predicate."""
    cases = {}
    n = 20000
    c = _f(rng.uniform(-1, 1, (n, 3)))
    h = _f(rng.uniform(1e-3, 0.2, (n, 3)))
    tv = _f(c[:, None, :] + rng.normal(0, 1, (n, 3, 3)) * h[:, None, :] * rng.uniform(0.5, 4, (n, 1, 1)))
    tv = tv.reshape(n, 9)
    special = _f([np.nan, np.inf, -np.inf, 1e20, -1e20, 3e38, -3e38])
    k = rng.integers(1, 4, n)
    for i in range(n):
        pos = rng.choice(9, k[i], replace=False)
        tv[i, pos] = special[rng.integers(0, len(special), k[i])]
    cases["special"] = (c, h, tv)
    # one vertex far away: edge components and projections overflow to
    # +-inf, and differences of them to NaN, on some axes only
    c = _f(rng.uniform(-1, 1, (n, 3)))
    h = _f(rng.uniform(1e-3, 0.2, (n, 3)))
    tv = _f(c[:, None, :] + rng.normal(0, 1, (n, 3, 3)) * h[:, None, :] * 2).reshape(n, 9)
    v = rng.integers(0, 3, n)
    ax = rng.integers(0, 3, (n, 2))
    mag = _f(10 ** rng.uniform(19, 38.5, (n, 2)) * np.where(rng.random((n, 2)) < 0.5, -1, 1))
    for j in range(2):
        tv[np.arange(n), 3 * v + ax[:, j]] = mag[:, j]
    cases["overflow"] = (c, h, tv)
    return cases


def normal_cases(rng, sat):
    """Face normals of every SAT-family triangle (degenerate ones included:
    zero cross products stay zero), VectorNormalize on vectors from 1e-30 to
    1e30 plus zeros and denormals, and Level1 averages over lists of 1..80
    unit normals (the bins' sum / count / normalize)."""
    tv = np.concatenate([v[2] for v in sat.values()])
    n = 20000
    v = _f(rng.normal(0, 1, (n, 3)) * 10 ** rng.uniform(-30, 30, (n, 1)))
    v[:50] = 0
    v[50:100] = _f(rng.normal(0, 1, (50, 3)) * 1e-40)
    v[100:150, 1:] = 0
    sizes = rng.integers(1, 81, 1000)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    fn = orc.face_normals(_f(rng.normal(0, 1, (int(off[-1]), 9))), use_ref=True)
    avg = np.zeros((len(sizes), 3), F32)
    orc.ref().ref_average_normals(len(sizes), off, fn.ravel(), avg.ravel())
    # face_n follows the triangles of sat_cases.npz, families in file order
    return {"face_n": orc.face_normals(tv, use_ref=True),
            "norm_v": v, "norm_out": orc.normalize(v, use_ref=True),
            "avg_off": off, "avg_fn": fn, "avg_out": avg}


def main():
    R = orc.ref()
    if R is None:
        sys.exit("oracle/_ref/libgpvref.so missing: run `make -C oracle` with the reference tree mounted")
    rng = np.random.default_rng(20261202)
    sat = sat_cases(rng)
    out = {}
    for k, (c, h, tv) in sat.items():
        res = orc.tri_box(c, h, tv, use_ref=True)
        out[f"{k}_c"], out[f"{k}_h"], out[f"{k}_tv"], out[f"{k}_res"] = c, h, tv, res
    np.savez_compressed(os.path.join(HERE, "sat_cases.npz"), **out)
    mt = mt_cases(rng)
    out2 = {}
    for k, (tv, o, d) in mt.items():
        hit, t = orc.tri_ray(tv, o, d, use_ref=True)
        out2[f"{k}_tv"], out2[f"{k}_o"], out2[f"{k}_d"] = tv, o, d
        out2[f"{k}_hit"], out2[f"{k}_t"] = hit, t
    np.savez_compressed(os.path.join(HERE, "mt_cases.npz"), **out2)
    nc = normal_cases(np.random.default_rng(20261203), sat)
    np.savez_compressed(os.path.join(HERE, "normal_cases.npz"), **nc)
    nf = nonfinite_cases(np.random.default_rng(20261204))
    out3 = {}
    for k, (c, h, tv) in nf.items():
        out3[f"{k}_c"], out3[f"{k}_h"], out3[f"{k}_tv"] = c, h, tv
        out3[f"{k}_res"] = orc.tri_box(c, h, tv, use_ref=True)
    np.savez_compressed(os.path.join(HERE, "sat_nonfinite.npz"), **out3)
    src = "/root/reference/src"
    man = {
        "generator": "tests/golden/make_golden.py",
        "seed": 20261202,
        "expected_outputs_from": ["src/TriBoxIntersection.cpp", "src/TriRayIntersection.cpp"],
        "compiled_with": "g++ -O2 -ffp-contract=off -fno-fast-math (oracle/Makefile)",
        "sha256": {f: hashlib.sha256(open(os.path.join(src, f), "rb").read()).hexdigest()
                   for f in ("TriBoxIntersection.cpp", "TriRayIntersection.cpp")},
        "normals_from": "includes/FloatVector.h (Float3, VectorCrossProduct, VectorNormalize) via "
                        "oracle/ref_vec_shim.cpp; normal_cases.npz uses seed 20261203",
        "sha256_includes": {"FloatVector.h": hashlib.sha256(
            open("/root/reference/includes/FloatVector.h", "rb").read()).hexdigest()},
        "sat_families": {k: int(len(v[0])) for k, v in sat.items()},
        "sat_hit_rate": {k: float(out[f"{k}_res"].mean()) for k in sat},
        "mt_families": {k: int(len(v[0])) for k, v in mt.items()},
        "mt_hit_rate": {k: float(out2[f"{k}_hit"].mean()) for k in mt},
        "sat_nonfinite": {"seed": 20261204, "families": {k: int(len(v[0])) for k, v in nf.items()},
                          "hit_rate": {k: float(out3[f"{k}_res"].mean()) for k in nf}},
    }
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(man, f, indent=1)
    print(json.dumps(man, indent=1))


if __name__ == "__main__":
    main()
