"""The oracle's normals (Level1Normal.raw / Level2Normal.raw semantics) pinned
against the reference's own vector arithmetic and against closed forms.

- face normal, VectorNormalize and the Level1 list average: golden vectors
  written by tests/golden/make_golden.py from includes/FloatVector.h compiled
  into oracle/_ref (bit-exact), and live against oracle/_ref when present;
- an axis-aligned box mesh: voxels touching one face have the exact face
  normal, voxels on an edge the normalized sum of the two (sign and axis
  mistakes fail here);
- a single triangle: every boundary voxel's normal is the triangle's;
- the .raw byte uint8(n * 256/3 + 127) at n = 0, +-1, +-0.5 worked by hand.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN

FAMILIES = ["random", "face_snapped", "lattice", "sliver", "axis_plane", "degenerate"]


@pytest.fixture(scope="module")
def golden():
    z = np.load(os.path.join(GOLDEN, "normal_cases.npz"), allow_pickle=False)
    s = np.load(os.path.join(GOLDEN, "sat_cases.npz"), allow_pickle=False)
    tv = np.concatenate([s[f"{f}_tv"] for f in FAMILIES])
    return z, tv


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_face_normals_golden(orc, golden):
    z, tv = golden
    assert np.array_equal(bits(orc.face_normals(tv)), bits(z["face_n"]))


def test_normalize_golden(orc, golden):
    z, _ = golden
    assert np.array_equal(bits(orc.normalize(z["norm_v"])), bits(z["norm_out"]))


def test_level1_average_golden(orc, golden):
    """The bin average of src/Object.cpp:3226-3250 (sum in order, divide by
    the count, normalize) against the reference's own Float3 operators."""
    z, _ = golden
    got = orc.average_normals(z["avg_off"], z["avg_fn"])
    assert np.array_equal(bits(got), bits(z["avg_out"]))


def test_live_against_reference(orc):
    R = orc.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (no reference tree)")
    rng = np.random.default_rng(11)
    tv = (rng.normal(0, 1, (50000, 9)) * 10 ** rng.uniform(-6, 6, (50000, 1))).astype(np.float32)
    assert np.array_equal(bits(orc.face_normals(tv)), bits(orc.face_normals(tv, use_ref=True)))


def _box_mesh(lo, hi):
    """Axis-aligned box, 12 outward-facing triangles."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    v = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], np.float32)
    f = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4],
                  [2, 3, 7], [2, 7, 6], [1, 2, 6], [1, 6, 5], [0, 4, 7], [0, 7, 3]], np.int32)
    return v, f


def test_box_mesh_normals_closed_form(orc, gv):
    v, f = _box_mesh((0.2, 0.25, 0.3), (0.7, 0.8, 0.75))
    soup = gv.flatten(v, f)
    bmin, bmax, mms = orc.bbox(v)
    g = orc.grid_setup(bmin, bmax, mms, 16, 4)
    r = orc.voxelize(soup, g, normals=True)
    tris = soup.reshape(-1, 9)
    fn = orc.face_normals(tris)
    # every face normal is an exact signed unit axis
    for t in range(len(tris)):
        assert np.count_nonzero(fn[t]) == 1 and abs(fn[t]).max() == 1.0
    seen_face = seen_edge = 0
    for b, vox in enumerate(r["bidx"]):
        tl = r["list_tri"][r["list_off"][vox]:r["list_off"][vox + 1]]
        dirs = {tuple(fn[t]) for t in tl}
        got = r["l1n"][b]
        if len(dirs) == 1:
            assert np.array_equal(got, np.array(next(iter(dirs)), np.float32))
            seen_face += 1
        elif len(dirs) == 2:
            a, c = (np.array(d, np.float32) for d in dirs)
            cnt = {d: sum(1 for t in tl if tuple(fn[t]) == d) for d in dirs}
            if len(set(cnt.values())) == 1:   # equal weights: (a + c)/|a + c|
                s = (a + c) / np.float32(np.sqrt(np.float32(2.0)))
                assert np.allclose(got, s, atol=0, rtol=2 ** -22), (got, s)
                assert np.all(np.sign(got) == np.sign(a + c))
                seen_edge += 1
    assert seen_face > 50 and seen_edge > 4


def test_single_triangle_normals(orc, gv):
    tri = np.array([[0.1, 0.2, 0.3], [0.9, 0.35, 0.5], [0.4, 0.8, 0.7]], np.float32)
    soup = tri.reshape(1, 9)
    bmin, bmax, mms = orc.bbox(tri)
    g = orc.grid_setup(bmin, bmax, mms, 24, 4)
    r = orc.voxelize(soup, g, normals=True)
    want = orc.normalize(orc.face_normals(soup))[0]   # average of one, normalized again
    assert r["nb"] > 20 and np.array_equal(bits(r["l1n"]), bits(np.tile(want, (r["nb"], 1))))
    hit = r["l2"] == 2
    assert np.array_equal(bits(r["l2n"][hit]), bits(np.tile(want, (int(hit.sum()), 1))))
    assert not np.any(r["l2n"][~hit])


def test_quantize_normal_bytes(orc):
    # 256/3 = 85.333336 in float: 0 -> 127, 1 -> 212.33 -> 212, -1 -> 41.67 -> 41,
    # 0.5 -> 169.67 -> 169, -0.5 -> 84.33 -> 84
    assert orc.quantize_normals([0.0, 1.0, -1.0, 0.5, -0.5]).tolist() == [127, 212, 41, 169, 84]
