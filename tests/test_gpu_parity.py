"""HIP path vs the oracle, bit for bit, through the C ABI.

Every case compares the four products of the path -- Level1 states
(Level1InOut.raw semantics), the exclusive boundary prefix sum, the boundary
index list and the Level2 states -- plus the reference's counts
(src/Object.cpp:3353-3378).  Inputs are the seeded generators and the
reference's own sample mesh (files/cessna.obj, copied as an input fixture).
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, compare_results, say

pytestmark = pytest.mark.gpu


def run_case(gv, orc, ctx, name, verts, faces, vc, vc2, level2=True, mode=None, initial_capacity=0,
             threads=0):
    soup = gv.flatten(verts, faces)
    bmin, bmax, mms = gv.mesh_bbox(verts)
    g = orc.grid_setup(bmin, bmax, mms, vc, vc2 if level2 else 0)
    want = orc.voxelize(soup, g, orc.ORC_FAST if mode is None else mode, threads)
    say(f"{name}: T={len(soup)} grid={list(g.n)} n2={g.n2}")
    grid = ctx.voxelize(soup, bmin, bmax, mms, vc, vc2, level2, initial_capacity)
    try:
        got = grid.download()
        info = grid.info()
    finally:
        grid.release()
    compare_results(name, got, want, g.n2)
    for k in ("l1_inside", "l1_boundary", "l2_inside", "l2_boundary", "tests_l1", "tests_l2", "pairs",
              "max_tris_per_voxel"):
        assert info[k] == want[k], f"{name}: {k} {info[k]} != {want[k]}"
    assert info["nb"] == want["nb"]
    return info, want


def test_c1_icosphere_exact(gv, orc, ctx):
    """Config 1: icosphere L4 (5,120 triangles), Level1 64, Level2 4, against
    the literal all-triangle parity loops."""
    v, f = gv.icosphere(4, 0.4, (0.5, 0.5, 0.5), 0.02, seed=1)
    info, want = run_case(gv, orc, ctx, "c1-exact", v, f, 64, 4, mode=orc.ORC_EXACT)
    assert info["n"] == [64, 64, 64] and want["nb"] > 10000 and want["l1_inside"] > 50000


def test_cessna_reference_defaults(gv, orc, ctx):
    """files/cessna.obj at the reference defaults voxelCount=8, voxelCount2=4
    (src/GLParameters.cpp:66-67)."""
    v, f, _ = gv.read_obj(os.path.join(GOLDEN, "cessna.obj"))
    run_case(gv, orc, ctx, "cessna-8-4", v, f, 8, 4, mode=orc.ORC_EXACT)


@pytest.mark.parametrize("vc,vc2", [(64, 8), (128, 16)])
def test_cessna_fine(gv, orc, ctx, vc, vc2):
    v, f, _ = gv.read_obj(os.path.join(GOLDEN, "cessna.obj"))
    run_case(gv, orc, ctx, f"cessna-{vc}-{vc2}", v, f, vc, vc2)


@pytest.mark.parametrize("kind,target,vc,vc2,seed", [
    (0, 20000, 64, 8, 10), (1, 30000, 128, 4, 11), (2, 20000, 64, 8, 12), (2, 60000, 128, 16, 13),
    (1, 3000, 256, 4, 14), (0, 5000, 100, 5, 15), (2, 4000, 48, 3, 16), (0, 3000, 40, 1, 17),
    (1, 5000, 72, 2, 18), (0, 20000, 64, 12, 19), (2, 15000, 56, 9, 20),
])
def test_mixed_kinds(gv, orc, ctx, kind, target, vc, vc2, seed):
    v, f = gv.mixed(kind, target, seed)
    run_case(gv, orc, ctx, f"mixed-k{kind}-{target}", v, f, vc, vc2)


def test_stress_snapped_slivers(gv, orc, ctx):
    """Lattice-snapped vertices (p=0.3), 5% face-coplanar triangles and 10%
    slivers on a drilled box."""
    v, f = gv.mixed(2, 20000, 21)
    vs, fs = gv.stress(v, f, 64, 8, p_snap=0.3, f_sliver=0.1, f_coplanar=0.05, seed=22)
    run_case(gv, orc, ctx, "stress-box", vs, fs, 64, 8)


def test_stress_icosphere_exact(gv, orc, ctx):
    v, f = gv.icosphere(3, 0.4, (0.5, 0.5, 0.5), 0.02, seed=3)
    vs, fs = gv.stress(v, f, 32, 4, p_snap=0.5, f_sliver=0.2, f_coplanar=0.1, seed=23)
    run_case(gv, orc, ctx, "stress-ico-exact", vs, fs, 32, 4, mode=orc.ORC_EXACT)


def test_level2_off(gv, orc, ctx):
    v, f = gv.icosphere(3)
    run_case(gv, orc, ctx, "level2-off", v, f, 48, 4, level2=False)


def test_single_triangle(gv, orc, ctx):
    v = np.float32([[0.1, 0.2, 0.3], [0.8, 0.3, 0.5], [0.4, 0.9, 0.7]])
    f = np.int32([[0, 1, 2]])
    run_case(gv, orc, ctx, "one-triangle", v, f, 16, 4)


def test_no_triangles(gv, orc, ctx):
    """Vertices but no faces: every voxel outside, no boundary."""
    v, _, _ = gv.read_obj(os.path.join(GOLDEN, "cessna.obj"))
    f = np.zeros((0, 3), np.int32)
    info, _ = run_case(gv, orc, ctx, "no-triangles", v, f, 16, 4)
    assert info["nb"] == 0 and info["l1_inside"] == 0


def test_capacity_regrow(gv, orc, ctx):
    """initial_capacity=1 forces the reference's rerun-with-larger-buffer path
    (src/Object.cpp:3198-3214); the result must not change."""
    v, f = gv.icosphere(4)
    info, _ = run_case(gv, orc, ctx, "regrow", v, f, 64, 4, initial_capacity=1)
    assert info["regrows"] >= 1


def test_staged_entry_points_match(gv, orc, ctx):
    """The stage entry points in PerformVoxelization's order give the same grid."""
    v, f = gv.mixed(1, 8000, 31)
    soup = gv.flatten(v, f)
    bmin, bmax, mms = gv.mesh_bbox(v)
    say("staged: torus 8k")
    grid = ctx.create(soup, bmin, bmax, mms, 64, 8)
    try:
        grid.classify_inout()
        while grid.classify_tessellation():
            pass
        grid.compact_boundary()
        grid.classify_inout_level2()
        grid.classify_tessellation_level2()
        grid.finish()
        got = grid.download()
    finally:
        grid.release()
    g = orc.grid_setup(bmin, bmax, mms, 64, 8)
    compare_results("staged", got, orc.voxelize(soup, g), 8)


def test_repeat_is_deterministic(gv, ctx):
    v, f = gv.mixed(2, 30000, 41)
    soup = gv.flatten(v, f)
    bmin, bmax, mms = gv.mesh_bbox(v)
    outs = []
    for i in range(3):
        say(f"repeat {i}")
        grid = ctx.voxelize(soup, bmin, bmax, mms, 96, 8)
        outs.append(grid.download())
        grid.release()
    for o in outs[1:]:
        for k in outs[0]:
            assert np.array_equal(o[k], outs[0][k])


def test_save_raw_bytes(gv, orc, ctx, tmp_path):
    """SaveVoxelization files (src/Object.cpp:2934-3075) byte-identical to the
    oracle's writer."""
    v, f = gv.icosphere(4)
    soup = gv.flatten(v, f)
    bmin, bmax, mms = gv.mesh_bbox(v)
    g = orc.grid_setup(bmin, bmax, mms, 64, 4)
    want_dir = tmp_path / "want"
    got_dir = tmp_path / "got"
    want_dir.mkdir()
    got_dir.mkdir()
    import ctypes
    r = orc._Result()
    orc.lib().orc_voxelize(np.ascontiguousarray(soup.ravel()), len(soup), ctypes.byref(g), orc.ORC_FAST, 0,
                           ctypes.byref(r))
    orc.lib().orc_save_raw(ctypes.byref(g), ctypes.byref(r), -1, str(want_dir).encode())
    orc.lib().orc_result_free(ctypes.byref(r))
    say("save_raw icosphere")
    grid = ctx.voxelize(soup, bmin, bmax, mms, 64, 4)
    grid.save_raw(-1, str(got_dir))
    grid.release()
    names = sorted(os.listdir(want_dir))
    assert names == sorted(os.listdir(got_dir)) and len(names) == 4
    for n in names:
        assert (want_dir / n).read_bytes() == (got_dir / n).read_bytes(), n


@pytest.mark.parametrize("sat_w,par_w,stage", [(1, 1, 128), (4, 2, 256), (2, 4, 128)])
def test_launch_shapes_single_mesh(gv, orc, sat_w, par_w, stage):
    """Every launch shape the tuner can pick (Level2 SAT and parity waves per
    block, triangles per LDS batch of the expansions) gives the oracle's
    grids on the single-mesh kernels (n2 = 8 and 4 specialisations, both
    Level2 SAT forms)."""
    with gv.Context(0) as c:
        for below in (0, 1 << 20):
            c.set_tune(l2_sat_waves=sat_w, l2_parity_waves=par_w, expand_stage_tris=stage, l2_group_below=below)
            v, f = gv.mixed(1, 20000, 71)
            run_case(gv, orc, c, f"shapes {sat_w}/{par_w}/{stage}/{below} torus 64/8", v, f, 64, 8)
            v, f = gv.mixed(2, 9000, 72)
            run_case(gv, orc, c, f"shapes {sat_w}/{par_w}/{stage}/{below} box 48/4", v, f, 48, 4)
