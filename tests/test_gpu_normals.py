"""Normals (Level1Normal.raw / Level2Normal.raw) of the HIP path against the
oracle, bit for bit on the float normals and byte for byte on the files.

The oracle's normals are pinned by tests/test_oracle_normals.py (reference
FloatVector.h golden vectors and closed forms).
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, compare_results, say

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def normals_case(gv, orc, ctx, name, verts, faces, vc, vc2, level2=True):
    soup = gv.flatten(verts, faces)
    bmin, bmax, mms = gv.mesh_bbox(verts)
    g = orc.grid_setup(bmin, bmax, mms, vc, vc2 if level2 else 0)
    want = orc.voxelize(soup, g, orc.ORC_FAST, normals=True)
    say(f"normals {name}: T={len(soup)} grid={list(g.n)} n2={g.n2}")
    grid = ctx.voxelize(soup, bmin, bmax, mms, vc, vc2, level2, normals=True)
    try:
        got = grid.download()
        l1n, l2n = grid.download_normals()
        info = grid.info()
    finally:
        grid.release()
    compare_results(name, got, want, g.n2)
    assert info["has_normals"] == 1
    bad = np.nonzero(np.any(bits(l1n) != bits(want["l1n"]), axis=1))[0]
    assert not len(bad), f"{name}: {len(bad)} Level1 normals differ, first b={bad[0]}: {l1n[bad[0]]} vs {want['l1n'][bad[0]]}"
    if g.n2:
        bad = np.nonzero(np.any(bits(l2n) != bits(want["l2n"]), axis=1))[0]
        assert not len(bad), f"{name}: {len(bad)} Level2 normals differ, first {bad[0]}"
    return want


def test_normals_icosphere(gv, orc, ctx):
    v, f = gv.icosphere(4, 0.4, (0.5, 0.5, 0.5), 0.02, seed=1)
    normals_case(gv, orc, ctx, "ico-64-4", v, f, 64, 4)


def test_normals_cessna_reference_defaults(gv, orc, ctx):
    v, f, _ = gv.read_obj(os.path.join(GOLDEN, "cessna.obj"))
    normals_case(gv, orc, ctx, "cessna-8-4", v, f, 8, 4)


@pytest.mark.parametrize("kind,target,vc,vc2,seed", [(2, 6000, 32, 8, 61), (1, 4000, 48, 16, 62)])
def test_normals_mixed(gv, orc, ctx, kind, target, vc, vc2, seed):
    v, f = gv.mixed(kind, target, seed)
    normals_case(gv, orc, ctx, f"mixed-k{kind}", v, f, vc, vc2)


def test_normals_stress_level2_off(gv, orc, ctx):
    v, f = gv.mixed(0, 3000, 63)
    vs, fs = gv.stress(v, f, 32, 4, p_snap=0.3, f_sliver=0.1, f_coplanar=0.05, seed=64)
    normals_case(gv, orc, ctx, "stress-l1-only", vs, fs, 32, 4, level2=False)


def test_normals_in_batch_and_files(gv, orc, ctx, tmp_path):
    """Batch members with and without normals; the .raw normal files equal
    the oracle's floats quantised by the reference's byte rule."""
    from gpview_amd import workloads
    ms = workloads.mixed_stream(gv, 5, seed=71, tri_range=(1000, 8000), level1=(16, 32), level2=(4, 8))
    for i, m in enumerate(ms):
        m["normals"] = i % 2 == 0
    say("normals batch of 5")
    grids = ctx.voxelize_batch(ms)
    try:
        for i, (m, gr) in enumerate(zip(ms, grids)):
            g = orc.grid_setup(m["bmin"], m["bmax"], m["mms"], m["voxel_count"], m["voxel_count2"])
            want = orc.voxelize(m["soup"], g, orc.ORC_FAST, normals=True)
            compare_results(f"batch mesh {i}", gr.download(), want, g.n2)
            if not m["normals"]:
                with pytest.raises(gv.GpvError):
                    gr.download_normals()
                continue
            l1n, l2n = gr.download_normals()
            assert np.array_equal(bits(l1n), bits(want["l1n"])), f"batch mesh {i} Level1 normals"
            assert np.array_equal(bits(l2n), bits(want["l2n"])), f"batch mesh {i} Level2 normals"
            d = tmp_path / f"m{i}"
            d.mkdir()
            gr.save_raw(3, str(d))
            n1 = want["l1"].size
            l1b = np.full((n1, 3), 127, np.uint8)
            l1b[want["bidx"]] = orc.quantize_normals(want["l1n"]).reshape(-1, 3)
            assert (d / "Obj3Level1Normal.raw").read_bytes() == l1b.tobytes()
            assert (d / "Obj3Level2Normal.raw").read_bytes() == orc.quantize_normals(want["l2n"]).tobytes()
    finally:
        for gr in grids:
            gr.release()
