"""Pins the oracle's predicates.

1. Golden vectors: tests/golden/{sat,mt}_cases.npz hold inputs and the
   outputs of the reference's own TriBoxOverlap / triangle_ray_intersection
   (src/TriBoxIntersection.cpp, src/TriRayIntersection.cpp compiled
   unmodified, see tests/golden/make_golden.py).  The oracle must match every
   case bit for bit (hit flag and the float t).  sat_nonfinite.npz does the
   same for SAT inputs with NaN and infinite projections.
2. Live cross-check against oracle/_ref when the reference tree is mounted.
3. Mathematics the golden set leaves open: the float thresholds standing in
   for the double 1e-6 compares, the column factorisation of the +z ray, and
   the conservative column bound (no accepted ray may fall outside it).
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN


def _load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


SAT_FAMILIES = ["random", "face_snapped", "lattice", "sliver", "axis_plane", "degenerate"]
MT_FAMILIES = ["random_z", "det_edge", "t_edge", "on_edge", "sliver_edge_on", "any_dir"]


@pytest.mark.parametrize("fam", SAT_FAMILIES)
def test_sat_golden(orc, fam):
    z = _load("sat_cases.npz")
    got = orc.tri_box(z[f"{fam}_c"], z[f"{fam}_h"], z[f"{fam}_tv"])
    want = z[f"{fam}_res"]
    assert len(want) >= 10000
    assert 0.05 < want.mean() < 0.95, "family must exercise both outcomes"
    assert np.array_equal(got, want), f"{fam}: {np.count_nonzero(got != want)} mismatches"


@pytest.mark.parametrize("fam", ["special", "overflow"])
@pytest.mark.parametrize("factored", [False, True])
def test_sat_nonfinite_golden(orc, fam, factored):
    """Inputs outside the pipeline's domain (NaN, infinities, overflowing
    products), where the reference's per-macro order of the two projections
    decides the outcome (AXISTEST_Z12 compares p2 before p1): the oracle's
    literal and factored SAT take the reference's result on every case."""
    z = _load("sat_nonfinite.npz")
    got = orc.tri_box(z[f"{fam}_c"], z[f"{fam}_h"], z[f"{fam}_tv"], factored=factored)
    want = z[f"{fam}_res"]
    assert len(want) >= 10000 and 0.05 < want.mean() < 0.95
    tv = z[f"{fam}_tv"]
    assert np.all(np.any(~(np.abs(tv) <= 2.0 ** 40), axis=1)), "every case lies outside the pipeline's domain"
    assert np.array_equal(got, want), f"{fam}: {np.count_nonzero(got != want)} mismatches"


@pytest.mark.parametrize("fam", MT_FAMILIES)
def test_mt_golden(orc, fam):
    z = _load("mt_cases.npz")
    hit, t = orc.tri_ray(z[f"{fam}_tv"], z[f"{fam}_o"], z[f"{fam}_d"])
    assert np.array_equal(hit, z[f"{fam}_hit"]), f"{fam}: hit mismatch"
    assert np.array_equal(t.view(np.uint32), z[f"{fam}_t"].view(np.uint32)), f"{fam}: t mismatch"


@pytest.mark.parametrize("fam", MT_FAMILIES[:-1])
def test_mt_split_equals_literal(orc, fam):
    """det/u/v per column + t per voxel == the literal MT for +z rays."""
    z = _load("mt_cases.npz")
    hit = orc.tri_zray_split(z[f"{fam}_tv"], z[f"{fam}_o"])
    assert np.array_equal(hit, z[f"{fam}_hit"])


def test_live_reference_random(orc):
    if orc.ref() is None:
        pytest.skip("oracle/_ref not built (reference tree not mounted here)")
    rng = np.random.default_rng(99)
    n = 300000
    c = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    h = rng.uniform(1e-3, 0.1, (n, 3)).astype(np.float32)
    tv = (c[:, None, :] + rng.normal(0, 0.08, (n, 3, 3))).astype(np.float32).reshape(n, 9)
    assert np.array_equal(orc.tri_box(c, h, tv), orc.tri_box(c, h, tv, use_ref=True))
    d = np.tile(np.float32([0, 0, 1]), (n, 1))
    hit1, t1 = orc.tri_ray(tv, c, d)
    hit2, t2 = orc.tri_ray(tv, c, d, use_ref=True)
    assert np.array_equal(hit1, hit2) and np.array_equal(t1.view(np.uint32), t2.view(np.uint32))


def test_epsilon_thresholds():
    """(double)x > 1e-6 <=> x >= 0x358637BE and (double)x < 1e-6 <=> x <= 0x358637BD,
    the float constants the device predicates use for the reference's double
    EPSILON (src/TriRayIntersection.cpp:38, 96, 122)."""
    above = np.uint32(0x358637BE).view(np.float32)
    below = np.uint32(0x358637BD).view(np.float32)
    assert float(below) < 1e-6 < float(above)
    assert np.nextafter(below, np.float32(1)) == above
    bits = np.arange(0x35863700, 0x35863900, dtype=np.uint32)
    xs = bits.view(np.float32)
    for x in xs:
        assert (float(x) > 1e-6) == (x >= above)
        assert (float(x) < 1e-6) == (x <= below)
        assert (float(-x) > -1e-6) == (-x >= -below)


def _column_bound_check(orc, tv, o, use_ref):
    """Every ray the reference MT accepts must have its column inside the bound."""
    d = np.tile(np.float32([0, 0, 1]), (len(o), 1))
    hit, _ = orc.tri_ray(tv, o, d, use_ref=use_ref)
    bmin = np.float32([-1, -1, -1])
    bmax = np.float32([2, 2, 2])
    bad = 0
    for i in np.nonzero(hit)[0]:
        ok, lo, hi = orc.column_bounds(tv[i], bmin, bmax, safety=1.0)
        assert ok
        x, y = float(o[i, 0]), float(o[i, 1])
        if not (lo[0] <= x <= hi[0] and lo[1] <= y <= hi[1]):
            bad += 1
    return int(hit.sum()), bad


def _ill_conditioned(rng, n, det_lo, det_hi):
    """Long, nearly parallel xy edges with |det| in [det_lo, det_hi]: the
    regime where float MT accepts rays beyond the triangle."""
    a = rng.uniform(0.2, 0.8, (n, 3))
    L = 10 ** rng.uniform(-1.5, -0.3, (n, 1))
    e1 = rng.normal(0, 1, (n, 3)) * L
    e2 = e1 * rng.uniform(0.2, 1.5, (n, 1))
    nrm = np.stack([-e1[:, 1], e1[:, 0]], 1)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    target = rng.choice([-1, 1], (n, 1)) * 10 ** rng.uniform(np.log10(det_lo), np.log10(det_hi), (n, 1))
    e2[:, :2] += nrm * (target / np.linalg.norm(e1[:, :2], axis=1, keepdims=True))
    return np.stack([a, a + e1, a + e2], 1).astype(np.float32), L


def test_column_bound_is_conservative(orc):
    """Rays near the vertices of ill-conditioned slivers: float MT accepts some
    of them outside the triangle's own xy bbox; every accepted ray must lie in
    the bound (evaluated with safety factor 1, the kernels use 4)."""
    rng = np.random.default_rng(11)
    use_ref = orc.ref() is not None
    spill = 0
    hits = 0
    for trial in range(4):
        n = 200000
        tv, L = _ill_conditioned(rng, n, 1e-6, 1e-5)
        k = rng.integers(0, 3, n)
        base = tv[np.arange(n), k, :2].astype(np.float64)
        dirn = rng.normal(0, 1, (n, 2))
        dirn /= np.linalg.norm(dirn, axis=1, keepdims=True)
        o = np.zeros((n, 3), np.float32)
        o[:, :2] = (base + dirn * L * 10 ** rng.uniform(-7, -1, (n, 1))).astype(np.float32)
        o[:, 2] = -0.5
        d = np.tile(np.float32([0, 0, 1]), (n, 1))
        hit, _ = orc.tri_ray(tv.reshape(n, 9), o, d, use_ref=use_ref)
        lo = tv[:, :, :2].min(1)
        hi = tv[:, :, :2].max(1)
        out = np.nonzero(hit.astype(bool) & ((o[:, :2] < lo) | (o[:, :2] > hi)).any(1))[0]
        hits += int(hit.sum())
        spill += len(out)
        for i in out:
            ok, blo, bhi = orc.column_bounds(tv[i].reshape(9), np.float32([-1, -1, -1]), np.float32([2, 2, 2]), 1.0)
            assert ok and blo[0] <= o[i, 0] <= bhi[0] and blo[1] <= o[i, 1] <= bhi[1], \
                f"accepted ray {o[i]} outside bound {blo} {bhi}"
    assert hits > 100000
    assert spill > 500, "the sample must exercise acceptances outside the triangle"


def test_column_bound_rejects_det(orc):
    tv = np.float32([0.1, 0.1, 0.1, 0.1, 0.1, 0.9, 0.2, 0.1, 0.5])  # vertical: det = 0
    ok, _, _ = orc.column_bounds(tv, np.float32([0, 0, 0]), np.float32([1, 1, 1]))
    assert not ok
