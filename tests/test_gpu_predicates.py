"""The device predicates on the reference's golden vectors, on the GPU.

tests/golden/{sat,mt}_cases.npz hold inputs and the outputs of the
reference's own TriBoxOverlap and triangle_ray_intersection
(src/TriBoxIntersection.cpp, src/TriRayIntersection.cpp compiled
unmodified).  A test-only gfx950 library (tests/native/device_predicates.hip,
built with the product's flags) evaluates gpview_amd/csrc/gpv_device.hpp on
every case:

- tri_box_overlap with all 13 predicates, and the three pruned forms the
  kernels use: the box axes established by sat_axis_ok, the e x z edges by
  xy_edges_ok, and both;
- the +z ray test split at the column seam: mt_setup (det), mt_column
  (u, v), mt_t_hit (t).

Each must equal the reference's result on every case of every family.
"""
import ctypes
import os

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, say

pytestmark = pytest.mark.gpu

SAT_FAMILIES = ["random", "face_snapped", "lattice", "sliver", "axis_plane", "degenerate"]
MT_ZRAY_FAMILIES = ["random_z", "det_edge", "t_edge", "on_edge", "sliver_edge_on"]
VARIANTS = ["full SAT", "box axes by sat_axis_ok", "e x z edges by xy_edges_ok", "both pruned"]


@pytest.fixture(scope="module")
def dp():
    so = os.path.join(ROOT, "tests", "native", "libdevpredicates.so")
    if not os.path.exists(so):
        pytest.fail(f"{so} missing: run `make -C tests/native` (hipcc, gfx950) before the GPU tests")
    L = ctypes.CDLL(so)
    f32 = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    u8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
    L.dp_sat.argtypes = [ctypes.c_int64, f32, f32, f32, u8]
    L.dp_mt.argtypes = [ctypes.c_int64, f32, f32, u8]
    return L


def _f(a, w):
    return np.ascontiguousarray(a, np.float32).reshape(-1, w)


@pytest.mark.parametrize("fam", SAT_FAMILIES)
def test_device_sat_golden(dp, fam):
    z = np.load(os.path.join(GOLDEN, "sat_cases.npz"), allow_pickle=False)
    c, h, tv = _f(z[f"{fam}_c"], 3), _f(z[f"{fam}_h"], 3), _f(z[f"{fam}_tv"], 9)
    want = z[f"{fam}_res"].astype(np.uint8)
    out = np.zeros(4 * len(c), np.uint8)
    say(f"device SAT {fam}: {len(c)} cases")
    assert dp.dp_sat(len(c), c.ravel(), h.ravel(), tv.ravel(), out) == 0
    out = out.reshape(-1, 4)
    for k, name in enumerate(VARIANTS):
        bad = np.nonzero(out[:, k] != want)[0]
        assert len(bad) == 0, f"{fam}/{name}: {len(bad)} of {len(c)} differ, first case {bad[0]}"


@pytest.mark.parametrize("fam", ["special", "overflow"])
def test_device_sat_nonfinite(dp, fam):
    """NaN and infinite projections: the pipeline refuses such coordinates,
    but the device predicates still take the reference's result, because
    each edge test compares its two projections in the order of the
    reference's macro."""
    z = np.load(os.path.join(GOLDEN, "sat_nonfinite.npz"), allow_pickle=False)
    c, h, tv = _f(z[f"{fam}_c"], 3), _f(z[f"{fam}_h"], 3), _f(z[f"{fam}_tv"], 9)
    want = z[f"{fam}_res"].astype(np.uint8)
    out = np.zeros(4 * len(c), np.uint8)
    say(f"device SAT nonfinite {fam}: {len(c)} cases")
    assert dp.dp_sat(len(c), c.ravel(), h.ravel(), tv.ravel(), out) == 0
    out = out.reshape(-1, 4)
    for k, name in enumerate(VARIANTS):
        bad = np.nonzero(out[:, k] != want)[0]
        assert len(bad) == 0, f"{fam}/{name}: {len(bad)} of {len(c)} differ, first case {bad[0]}"


@pytest.mark.parametrize("fam", MT_ZRAY_FAMILIES)
def test_device_mt_golden(dp, fam):
    z = np.load(os.path.join(GOLDEN, "mt_cases.npz"), allow_pickle=False)
    tv, o, d = _f(z[f"{fam}_tv"], 9), _f(z[f"{fam}_o"], 3), _f(z[f"{fam}_d"], 3)
    assert np.all(d == np.float32([0, 0, 1])), "the device test is the +z ray"
    want = z[f"{fam}_hit"].astype(np.uint8)
    hit = np.zeros(len(tv), np.uint8)
    say(f"device MT {fam}: {len(tv)} cases")
    assert dp.dp_mt(len(tv), tv.ravel(), o.ravel(), hit) == 0
    assert 0.02 < want.mean() < 0.98
    bad = np.nonzero(hit != want)[0]
    assert len(bad) == 0, f"{fam}: {len(bad)} of {len(tv)} differ, first case {bad[0]}"
