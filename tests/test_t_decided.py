"""The Level2 parity's voxel-wide t decision (gpv::t_decided in
gpview_amd/csrc/gpv_slab.hpp) may only settle pairs whose outcome the ray
test itself fixes: when it says every origin of a box passes t > 1e-6, each
origin whose det/u/v test passes must be a hit of the literal Moller-Trumbore
test (src/TriRayIntersection.cpp:78-131); when it says none passes, no origin
may hit.  Checked against the oracle's MT (and oracle/_ref when built) on the
golden +z families and on boxes placed just above and below the plane of
steep, small and far-from-origin triangles; with its rounding terms removed
the same sample must show contradictions, so the terms are load-bearing."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

NATIVE = os.path.join(ROOT, "tests", "native")
EPS_ABOVE = np.uint32(0x358637BE).view(np.float32)


@pytest.fixture(scope="module")
def tdec():
    so = os.path.join(NATIVE, "libslabcheck.so")
    subprocess.run(["make", "-s", "-C", NATIVE], check=True)
    L = ctypes.CDLL(so)
    f32 = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    L.t_decided_batch.argtypes = [ctypes.c_int64, f32, f32, ctypes.c_float, ctypes.c_double,
                                  np.ctypeslib.ndpointer(dtype=np.int8)]

    def dec(tv, box, slack=1.0):
        tv = np.ascontiguousarray(tv, np.float32).reshape(-1, 9)
        out = np.zeros(len(tv), np.int8)
        L.t_decided_batch(len(tv), tv.ravel(), np.ascontiguousarray(box, np.float32).reshape(-1, 6).ravel(),
                          float(EPS_ABOVE), float(slack), out)
        return out
    return dec


def _zray(orc, tv, o):
    """(hit, det/u/v pass) of the +z ray from each origin; the second from
    the same column far below the triangle, where t is large."""
    use_ref = orc.ref() is not None
    d = np.tile(np.float32([0, 0, 1]), (len(o), 1))
    hit, _ = orc.tri_ray(tv, o, d, use_ref=use_ref)
    low = o.copy()
    tz = tv.reshape(-1, 3, 3)[:, :, 2]
    low[:, 2] = (tz.min(1).astype(np.float64) - 1.0 - np.abs(tz).max(1)).astype(np.float32)
    uv, _ = orc.tri_ray(tv, low, d, use_ref=use_ref)
    return hit.astype(bool), uv.astype(bool)


def _contradictions(dec, hit, uv):
    """cases (rows) where a decision disagrees with the literal test"""
    bad_hit = (dec[:, None] == 1) & uv & ~hit
    bad_miss = (dec[:, None] == -1) & hit
    return np.nonzero((bad_hit | bad_miss).any(1))[0]


@pytest.mark.parametrize("fam", ["random_z", "det_edge", "t_edge", "on_edge", "sliver_edge_on"])
def test_t_decided_golden_points(orc, tdec, fam):
    """one-origin boxes at the reference's own golden rays"""
    z = np.load(os.path.join(GOLDEN, "mt_cases.npz"), allow_pickle=False)
    tv, o, d, want = z[f"{fam}_tv"], z[f"{fam}_o"], z[f"{fam}_d"], z[f"{fam}_hit"].astype(bool)
    plusz = np.all(d == np.float32([0, 0, 1]), axis=1)
    tv, o, want = tv[plusz], o[plusz], want[plusz]
    assert len(tv) > 1000
    box = np.repeat(o, 2, axis=1)   # (x, x, y, y, z, z)
    dec = tdec(tv, box)
    hit, uv = _zray(orc, tv, o)
    assert np.array_equal(hit, want)
    assert not np.any((dec == -1) & want), f"{fam}: decided miss on a golden hit"
    assert not np.any((dec == 1) & uv & ~want), f"{fam}: decided hit on a golden miss"
    assert np.count_nonzero(np.isin(dec, (1, -1))) > len(dec) // 4


def _boxes(rng, n, n2, nzl=None, ztall=0.0):
    """triangles of every slope, size and offset with an n2^3 box of origins:
    mode 0 just below the plane, 1 just above, 2 across it, 3 below with the
    highest origin's t a hair above 1e-6, 4 with the lowest origin's t a hair
    below 1e-6 (the two thresholds the decision compares against).  nzl
    layers (default n2) spaced up to 10^ztall times the xy spacing: the tall
    boxes the parity's column headers decide over (layer 0 up to a layer)."""
    nzl = n2 if nzl is None else nzl
    scale = 10 ** rng.uniform(-1, 2, n)
    base = rng.uniform(-1, 1, (n, 3)) * scale[:, None]
    nrm = rng.normal(0, 1, (n, 3))
    nrm[:, 2] = np.sign(nrm[:, 2]) * 10 ** rng.uniform(-3, 0, n)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    u = np.cross(nrm, rng.normal(0, 1, (n, 3)))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    w = np.cross(nrm, u)
    L = 10 ** rng.uniform(-1.5, 0.5, n)
    a = rng.uniform(-1, 1, (n, 3, 2))
    tri = (base[:, None, :] + L[:, None, None] * (a[:, :, :1] * u[:, None, :] + a[:, :, 1:] * w[:, None, :]))
    tri = tri.astype(np.float32)
    t64 = tri.astype(np.float64)
    X = (rng.dirichlet([1, 1, 1], n)[:, :, None] * t64).sum(1)
    N = np.cross(t64[:, 1] - t64[:, 0], t64[:, 2] - t64[:, 0])
    g2 = L * 10 ** rng.uniform(-4, -1, n)
    k = np.arange(n2) - (n2 - 1) / 2
    ox = (X[:, :1] + k * g2[:, None]).astype(np.float32)
    oy = (X[:, 1:2] + k * g2[:, None]).astype(np.float32)
    # plane height over the box's corners
    cx = np.stack([ox[:, 0], ox[:, -1], ox[:, 0], ox[:, -1]], 1).astype(np.float64)
    cy = np.stack([oy[:, 0], oy[:, 0], oy[:, -1], oy[:, -1]], 1).astype(np.float64)
    zp = t64[:, :1, 2] - (N[:, :1] * (cx - t64[:, :1, 0]) + N[:, 1:2] * (cy - t64[:, :1, 1])) / N[:, 2:3]
    delta = 10 ** rng.uniform(-7, -1, n)
    rho = 10 ** rng.uniform(-9, -2, n)
    mode = rng.integers(0, 5, n)
    gz = g2 * 10 ** rng.uniform(0, ztall, n)
    top = (nzl - 1) * gz
    z0 = np.select([mode == 0, mode == 1, mode == 2, mode == 3],
                   [zp.min(1) - delta - top, zp.max(1) + delta, zp.mean(1) - top / 2,
                    zp.min(1) - 1e-6 * (1 + rho) - top],
                   zp.max(1) - 1e-6 * (1 - rho))
    oz = (z0[:, None] + np.arange(nzl) * gz[:, None]).astype(np.float32)
    for arr in (ox, oy, oz):
        assert np.all(np.diff(arr, axis=1) >= 0)
    return tri.reshape(n, 9), ox, oy, oz


def _box_rays(tv, ox, oy, oz):
    n, n2 = ox.shape
    P = np.stack(np.meshgrid(np.arange(n2), np.arange(n2), np.arange(oz.shape[1]), indexing="ij"), -1).reshape(-1, 3)
    o = np.stack([ox[:, P[:, 0]], oy[:, P[:, 1]], oz[:, P[:, 2]]], -1).reshape(-1, 3)
    return np.repeat(tv, len(P), axis=0), o, len(P)


@pytest.mark.parametrize("nzl,ztall", [(4, 0.0), (12, 2.0)])
def test_t_decided_boxes_near_the_plane(orc, tdec, nzl, ztall):
    rng = np.random.default_rng(2024 + nzl)
    n, n2 = 20000 if nzl == 4 else 8000, 4
    tv, ox, oy, oz = _boxes(rng, n, n2, nzl, ztall)
    box = np.stack([ox[:, 0], ox[:, -1], oy[:, 0], oy[:, -1], oz[:, 0], oz[:, -1]], 1)
    rt, ro, m = _box_rays(tv, ox, oy, oz)
    hit, uv = _zray(orc, rt, ro)
    hit, uv = hit.reshape(n, m), uv.reshape(n, m)
    dec = tdec(tv, box)
    bad = _contradictions(dec, hit, uv)
    assert len(bad) == 0, f"{len(bad)} boxes decided against the MT, e.g. case {bad[:5]}"
    # the sample exercises both decisions, on boxes that rays actually reach
    reach = uv.any(1)
    assert np.count_nonzero((dec == 1) & reach) > n // 10
    assert np.count_nonzero((dec == -1) & reach) > n // 10
    assert np.count_nonzero((dec == 0) & reach) > n // 20
    # without its rounding terms the bound decides boxes the MT contradicts
    loose = tdec(tv, box, slack=0.0)
    assert len(_contradictions(loose, hit, uv)) > 0
