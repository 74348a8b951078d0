"""The plane-slab filter (gpview_amd/csrc/gpv_slab.hpp) may only skip boxes
the SAT rejects: checked against the reference's own TriBoxOverlap (golden
families, and live on dense borderline cases when oracle/_ref is built)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

NATIVE = os.path.join(ROOT, "tests", "native")


@pytest.fixture(scope="module")
def slab():
    so = os.path.join(NATIVE, "libslabcheck.so")
    if not os.path.exists(so):
        subprocess.run(["make", "-s", "-C", NATIVE], check=True)
    L = ctypes.CDLL(so)
    f32 = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    L.slab_keep_batch.argtypes = [ctypes.c_int64, f32, f32, f32, np.ctypeslib.ndpointer(dtype=np.uint8)]

    def keep(tv, c, h):
        tv = np.ascontiguousarray(tv, np.float32).reshape(-1, 9)
        out = np.zeros(len(tv), np.uint8)
        L.slab_keep_batch(len(tv), tv.ravel(), np.ascontiguousarray(c, np.float32).ravel(),
                          np.ascontiguousarray(h, np.float32).ravel(), out)
        return out
    return keep


@pytest.mark.parametrize("fam", ["random", "face_snapped", "lattice", "sliver", "axis_plane", "degenerate"])
def test_slab_never_drops_a_sat_hit_golden(slab, fam):
    z = np.load(os.path.join(GOLDEN, "sat_cases.npz"), allow_pickle=False)
    res = z[f"{fam}_res"].astype(bool)
    keep = slab(z[f"{fam}_tv"], z[f"{fam}_c"], z[f"{fam}_h"]).astype(bool)
    assert not np.any(res & ~keep), f"{fam}: {np.count_nonzero(res & ~keep)} SAT hits outside the slab"


def _plane_borderline(rng, n):
    """Boxes whose center sits within a few ulps-to-1e-3 of the slab edge
    |N.(c-t0)| = R, on a voxel-like lattice, with long and sliver triangles."""
    g = np.float32(1 / 256)
    h = np.full((n, 3), np.float32(g / 2), np.float32)
    c = ((rng.integers(0, 256, (n, 3)) + 0.5) * np.float64(g)).astype(np.float32)
    nrm = rng.normal(0, 1, (n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    R = (np.abs(nrm) * h).sum(1)
    off = R * (1 + rng.choice([-1, 1], n) * 10 ** rng.uniform(-9, -3, n)) * rng.choice([-1, 1], n)
    p0 = c - nrm * off[:, None]
    u = np.cross(nrm, rng.normal(0, 1, (n, 3)))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    w = np.cross(nrm, u)
    L = 10 ** rng.uniform(-3, -0.5, (n, 1))
    a = rng.uniform(-1, 1, (n, 3, 2))
    a[:, :, 1] *= np.where(rng.random((n, 1)) < 0.3, 1e-3, 1.0)
    tri = p0[:, None, :] + L[:, :, None] * (a[:, :, :1] * u[:, None, :] + a[:, :, 1:] * w[:, None, :])
    return tri.astype(np.float32).reshape(n, 9), c, h


def test_slab_never_drops_a_sat_hit_borderline(slab, orc):
    rng = np.random.default_rng(123)
    use_ref = orc.ref() is not None
    kept_miss = 0
    hits = 0
    for _ in range(5):
        tv, c, h = _plane_borderline(rng, 200000)
        res = orc.tri_box(c, h, tv, use_ref=use_ref).astype(bool)
        keep = slab(tv, c, h).astype(bool)
        assert not np.any(res & ~keep), f"{np.count_nonzero(res & ~keep)} SAT hits outside the slab"
        hits += int(res.sum())
        kept_miss += int((keep & ~res).sum())
    assert hits > 50000


def test_slab_prunes_far_boxes(slab):
    """Sanity of the other direction: boxes a few voxels off the plane are
    dropped (the filter has to be useful, not only safe)."""
    rng = np.random.default_rng(5)
    n = 20000
    tv = rng.uniform(0.2, 0.8, (n, 9)).astype(np.float32)
    t = tv.reshape(n, 3, 3).astype(np.float64)
    N = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 1])
    N /= np.linalg.norm(N, axis=1, keepdims=True)
    h = np.full((n, 3), 1 / 512, np.float32)
    c = (t[:, 0] + N * 5 / 256).astype(np.float32)
    assert slab(tv, c, h).sum() == 0


def _index_box(tv, bmin, bmax, n):
    """Index AABB of src/Object.cpp:2272-2318 in float32: int((v-min)/(max-min)*n),
    pulled back at the max plane, clipped to the grid."""
    v = tv.reshape(-1, 3, 3).astype(np.float32)
    mn = np.asarray(bmin, np.float32)
    mx = np.asarray(bmax, np.float32)
    r = (v - mn) / (mx - mn)
    s = np.clip(r * np.asarray(n, np.float32), -1, np.asarray(n, np.float32) + 1)
    i = np.trunc(s).astype(np.int64)
    i = np.where((i == np.asarray(n)) & (v == mx), i - 1, i)
    lo = np.maximum(i.min(1), 0)
    hi = np.minimum(i.max(1), np.asarray(n) - 1)
    return lo.astype(np.int32), hi.astype(np.int32)


def _big_triangles(rng, bmin, bmax, g, n_each):
    ext = np.asarray(bmax, np.float64) - bmin
    out = []
    # random, spanning up to ~12 voxels each way
    c = bmin + rng.uniform(0.2, 0.8, (n_each, 3)) * ext
    out.append(c[:, None, :] + rng.uniform(-12, 12, (n_each, 3, 3)) * np.asarray(g))
    # vertices on Level1 lattice planes / centers
    k = rng.integers(8, 56, (n_each, 3, 3))
    lat = bmin + k * np.asarray(g) + rng.choice([0.0, 0.5], (n_each, 3, 3)) * np.asarray(g)
    out.append(lat)
    # planes parallel to a grid plane (one shared coordinate), incl. exact layer planes
    t = c[:, None, :] + rng.uniform(-12, 12, (n_each, 3, 3)) * np.asarray(g)
    ax = rng.integers(0, 3, n_each)
    val = bmin[ax] + rng.integers(8, 56, n_each) * np.asarray(g)[ax] * rng.choice([1.0, 1.0 + 1e-7], n_each)
    t[np.arange(n_each), :, ax] = val[:, None]
    out.append(t)
    # long slivers
    a = c + rng.uniform(-10, 10, (n_each, 3)) * np.asarray(g)
    d = rng.normal(0, 1, (n_each, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    b = a + d * 20 * np.asarray(g).max()
    w = rng.normal(0, 1, (n_each, 3)) * 1e-3 * np.asarray(g).max()
    out.append(np.stack([a, b, (a + b) / 2 + w], 1))
    tv = np.concatenate(out).astype(np.float32).reshape(-1, 9)
    lo = np.asarray(bmin, np.float32)
    hi = np.asarray(bmax, np.float32)
    return np.clip(tv.reshape(-1, 3, 3), lo, hi).reshape(-1, 9)


def test_level1_layers_never_drop_a_sat_hit(orc):
    """gpv::slab_layers (the Level1 sweep's per-column layer interval for
    triangles spanning several layers): every layer where the reference SAT
    hits lies inside the interval; and the interval is useful (prunes most
    of the index-AABB layers of large triangles)."""
    so = os.path.join(NATIVE, "libslabcheck.so")
    if not os.path.exists(so):
        subprocess.run(["make", "-s", "-C", NATIVE], check=True)
    L = ctypes.CDLL(so)
    f32 = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    i32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
    i64 = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
    L.slab_layers_batch.argtypes = [ctypes.c_int64, f32, i32, i32, f32, f32, f32, f32, i64, i32, i32]
    use_ref = orc.ref() is not None
    rng = np.random.default_rng(2024)
    bmin = np.array([-0.0123, 0.01, 0.3], np.float32)
    bmax = np.array([0.98, 1.02, 1.25], np.float32)
    g = orc.grid_setup(bmin, bmax, float(np.max(bmax - bmin)), 64, 4)
    n = np.array(list(g.n))
    tabs = orc.tables(g)
    c1 = [np.ascontiguousarray(x, np.float32) for x in tabs["c1"]]
    h = np.ascontiguousarray(g.ext1, np.float32)
    tv = _big_triangles(rng, bmin.astype(np.float64), bmax.astype(np.float64), np.array(list(g.g)), 300)
    lo, hi = _index_box(tv, bmin, bmax, n)
    keep = np.all(hi - lo + 1 >= 1, axis=1) & (hi[:, 2] - lo[:, 2] + 1 >= 3)
    tv, lo, hi = tv[keep], np.ascontiguousarray(lo[keep]), np.ascontiguousarray(hi[keep])
    cols = (hi[:, 0] - lo[:, 0] + 1) * (hi[:, 1] - lo[:, 1] + 1)
    off = np.concatenate([[0], np.cumsum(cols)]).astype(np.int64)
    za = np.zeros(off[-1], np.int32)
    zb = np.zeros(off[-1], np.int32)
    L.slab_layers_batch(len(tv), np.ascontiguousarray(tv), lo.ravel(), hi.ravel(), c1[0], c1[1], c1[2], h, off, za, zb)
    kept = total = hits = 0
    for i in range(len(tv)):
        p = np.arange(lo[i, 0], hi[i, 0] + 1)
        q = np.arange(lo[i, 1], hi[i, 1] + 1)
        z = np.arange(lo[i, 2], hi[i, 2] + 1)
        Q, P, Z = np.meshgrid(q, p, z, indexing="ij")          # column order (q, p), then z
        ctr = np.stack([c1[0][P], c1[1][Q], c1[2][Z]], -1).reshape(-1, 3)
        res = orc.tri_box(ctr, np.tile(h, (len(ctr), 1)), np.tile(tv[i], (len(ctr), 1)),
                          use_ref=use_ref).reshape(len(q), len(p), len(z)).astype(bool)
        a = za[off[i]:off[i + 1]].reshape(len(q), len(p), 1)
        b = zb[off[i]:off[i + 1]].reshape(len(q), len(p), 1)
        inside = (Z >= a) & (Z < b)
        assert not np.any(res & ~inside), f"triangle {i}: SAT hit outside its layer interval"
        hits += int(res.sum())
        kept += int(inside.sum())
        total += inside.size
    assert hits > 50000
    assert kept < 0.5 * total, f"pruning too weak: kept {kept} of {total}"


def test_float_slab_never_drops_a_sat_hit(orc):
    """The Level2 float slab (gpv::float_slab over a range of centers, tested
    with gpv::float_slab_keeps): borderline planes at single centers, and
    large triangles whose slab spans a wide range of centers around the box."""
    so = os.path.join(NATIVE, "libslabcheck.so")
    if not os.path.exists(so):
        subprocess.run(["make", "-s", "-C", NATIVE], check=True)
    L = ctypes.CDLL(so)
    f32 = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    f64 = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    L.float_slab_keep_batch.argtypes = [ctypes.c_int64, f32, f32, f32, f64, np.ctypeslib.ndpointer(np.uint8)]
    use_ref = orc.ref() is not None
    rng = np.random.default_rng(77)
    hits = 0
    for rnd in range(4):
        tv, c, h = _plane_borderline(rng, 200000)
        c = np.ascontiguousarray(c, np.float32)
        if rnd % 2:
            # a wide centers range around c (up to 40 boxes each way)
            w = rng.uniform(0, 40, (len(c), 1)) * h.astype(np.float64)
            lohi = np.concatenate([c - w * rng.random((len(c), 3)), c + w * rng.random((len(c), 3))], 1)
        else:
            lohi = np.concatenate([c, c], 1).astype(np.float64)
        keep = np.zeros(len(c), np.uint8)
        L.float_slab_keep_batch(len(c), np.ascontiguousarray(tv), c, np.ascontiguousarray(h, np.float32),
                                np.ascontiguousarray(lohi, np.float64), keep)
        res = orc.tri_box(c, h, tv, use_ref=use_ref).astype(bool)
        assert not np.any(res & ~keep.astype(bool)), f"{np.count_nonzero(res & ~keep.astype(bool))} hits dropped"
        hits += int(res.sum())
    z = np.load(os.path.join(GOLDEN, "sat_cases.npz"), allow_pickle=False)
    for fam in ("random", "face_snapped", "lattice", "sliver", "axis_plane", "degenerate"):
        c = np.ascontiguousarray(z[f"{fam}_c"], np.float32)
        h = np.ascontiguousarray(z[f"{fam}_h"], np.float32)
        tv = np.ascontiguousarray(z[f"{fam}_tv"], np.float32)
        lohi = np.ascontiguousarray(np.concatenate([c, c], 1), np.float64)
        keep = np.zeros(len(c), np.uint8)
        L.float_slab_keep_batch(len(c), tv, c, h, lohi, keep)
        res = z[f"{fam}_res"].astype(bool)
        assert not np.any(res & ~keep.astype(bool)), fam
    assert hits > 50000
