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
