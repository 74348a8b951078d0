"""Pins the oracle's fast modes against its literal mode.

ORC_FAST is what the GPU parity tests compare with at large sizes, so every
shortcut it takes must give the literal loops' result:

- the SAT sweeps screen sub-boxes by the SAT's own predicate groups (the
  box-axis tests per axis, the edge tests along e x X / e x Y / e x Z per
  pair of center components) and decide survivors by the plane test
  (orc_tri_box_factored): equal to orc_tri_box_overlap on every golden
  family and on dense random cases near touching configurations;
- the Level1 parity runs per column with a conservative column bound;
- the Level2 parity tests det/u/v once per (candidate, sub-column) and finds
  each candidate's passing layers by a binary search when t is provably
  monotone along the column, one layer at a time otherwise.

ORC_PLAN, the CPU baseline's composition (column-factorised parity with a
t test per layer, literal SAT loops), is checked the same way.  The whole
pipeline in every mode must agree on every output (states,
prefix, boundary list, bins, counts) for meshes that exercise all of it:
lattice-snapped, sliver and face-coplanar triangles at Level2 up to 16, the
reference's sample mesh, and the CUDA-kernel Level2 candidate rule.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN

SAT_FAMILIES = ["random", "face_snapped", "lattice", "sliver", "axis_plane", "degenerate"]


@pytest.mark.parametrize("fam", SAT_FAMILIES)
def test_factored_sat_equals_golden(orc, fam):
    z = np.load(os.path.join(GOLDEN, "sat_cases.npz"), allow_pickle=False)
    got = orc.tri_box(z[f"{fam}_c"], z[f"{fam}_h"], z[f"{fam}_tv"], factored=True)
    assert np.array_equal(got, z[f"{fam}_res"]), f"{fam}: {np.count_nonzero(got != z[f'{fam}_res'])} mismatches"


def test_factored_sat_dense_touching(orc):
    """Triangles whose vertices sit on, just inside and just outside the box
    planes (the strict-inequality boundary of every predicate group)."""
    rng = np.random.default_rng(404)
    n = 400000
    c = rng.uniform(0.2, 0.8, (n, 3)).astype(np.float32)
    h = np.repeat(rng.uniform(1e-3, 0.05, (n, 1)), 3, 1).astype(np.float32)
    # vertices snapped to box faces, edges and corners, then nudged by 0-2 ulps
    sel = rng.integers(-1, 2, (n, 3, 3)).astype(np.float32)
    tv = c[:, None, :] + sel * h[:, None, :] + rng.normal(0, 1, (n, 3, 3)).astype(np.float32) * h[:, None, :] * \
        (rng.random((n, 3, 3)) < 0.5)
    tv = tv.astype(np.float32)
    nud = rng.integers(-2, 3, (n, 3, 3))
    tv = np.nextafter(tv, np.where(nud > 0, np.inf, -np.inf).astype(np.float32))
    tv = np.where(nud == 0, c[:, None, :] + sel * h[:, None, :], tv).astype(np.float32).reshape(n, 9)
    full = orc.tri_box(c, h, tv)
    fact = orc.tri_box(c, h, tv, factored=True)
    assert 0.05 < full.mean() < 0.95
    assert np.array_equal(full, fact), f"{np.count_nonzero(full != fact)} mismatches"
    if orc.ref() is not None:
        assert np.array_equal(full, orc.tri_box(c, h, tv, use_ref=True))


def _meshes(gv):
    out = []
    v, f = gv.icosphere(3, 0.4, (0.5, 0.5, 0.5), 0.02, seed=3)
    out.append(("icosphere-1280-24-8", v, f, 24, 8))
    v, f = gv.mixed(2, 3000, 61)
    vs, fs = gv.stress(v, f, 20, 16, p_snap=0.3, f_sliver=0.1, f_coplanar=0.05, seed=62)
    out.append(("stress-box-20-16", vs, fs, 20, 16))
    v, f = gv.mixed(1, 2500, 63)
    vs, fs = gv.stress(v, f, 28, 5, p_snap=0.5, f_sliver=0.2, f_coplanar=0.1, seed=64)
    out.append(("stress-torus-28-5", vs, fs, 28, 5))
    v, f = gv.mixed(0, 1500, 65)
    vs, fs = gv.stress(v, f, 16, 12, p_snap=0.3, f_sliver=0.1, f_coplanar=0.05, seed=66)
    out.append(("stress-ico-16-12", vs, fs, 16, 12))
    v, f, _ = gv.read_obj(os.path.join(GOLDEN, "cessna.obj"))
    out.append(("cessna-8-4", v, f, 8, 4))
    return out


@pytest.mark.parametrize("cuda_l2", [False, True])
def test_fast_equals_exact_pipeline(gv, orc, cuda_l2):
    for name, v, f, vc, vc2 in _meshes(gv):
        soup = gv.flatten(v, f)
        bmin, bmax, mms = gv.mesh_bbox(v)
        g = orc.grid_setup(bmin, bmax, mms, vc, vc2, cuda_l2=cuda_l2)
        a = orc.voxelize(soup, g, orc.ORC_EXACT)
        assert a["nb"] > 50, name
        for mode in (orc.ORC_FAST, orc.ORC_PLAN):
            b = orc.voxelize(soup, g, mode)
            for k in ("l1", "prefix", "bidx", "l2", "list_off", "list_tri"):
                assert np.array_equal(a[k], b[k]), f"{name} mode {mode}: {k} differs ({np.count_nonzero(a[k] != b[k])})"
            for k in ("pairs", "tests_l1", "tests_l2", "l1_inside", "l1_boundary", "l2_inside", "l2_boundary",
                      "max_tris_per_voxel"):
                assert a[k] == b[k], f"{name} mode {mode}: {k} {a[k]} != {b[k]}"
        assert 0 < a["l2_inside"] and 0 < a["l2_boundary"], name


def test_fast_is_thread_count_independent(gv, orc):
    v, f = gv.mixed(2, 8000, 67)
    soup = gv.flatten(v, f)
    bmin, bmax, mms = gv.mesh_bbox(v)
    g = orc.grid_setup(bmin, bmax, mms, 48, 8)
    a = orc.voxelize(soup, g, orc.ORC_FAST, 1)
    b = orc.voxelize(soup, g, orc.ORC_FAST, 7)
    for k in ("l1", "prefix", "bidx", "l2", "list_off", "list_tri"):
        assert np.array_equal(a[k], b[k]), k
