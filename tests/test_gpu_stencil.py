"""The GL-stencil-equivalent Level1 fill (gpv_params::l1_stencil;
ClassifyInOut2x, src/Object.cpp:1446-1601) on the device, through the C
ABI, against the oracle's restatement bit for bit (orc_stencil_inside: its
rule is pinned by tests/test_stencil_oracle.py), up to the headline mesh."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, compare_results, say
from test_stencil_oracle import UNIT, prism, square

pytestmark = pytest.mark.gpu


def stencil_case(gv, orc, ctx, name, soup, bmin, bmax, mms, vc, vc2):
    g = orc.grid_setup(bmin, bmax, mms, vc, vc2, l1_stencil=True)
    want = orc.voxelize(soup, g, orc.ORC_FAST)
    say(f"{name}: T={len(soup)} grid={list(g.n)} n2={g.n2} (stencil)")
    grid = ctx.voxelize(soup, bmin, bmax, mms, vc, vc2, vc2 > 0, l1_stencil=True)
    try:
        got = grid.download()
        info = grid.info()
    finally:
        grid.release()
    compare_results(name, got, want, g.n2)
    for k in ("l1_inside", "l1_boundary", "l2_inside", "l2_boundary"):
        assert info[k] == want[k], f"{name}: {k} {info[k]} != {want[k]}"
    return got, want


def test_stencil_tie_rules_on_sample_centers(gv, orc, ctx):
    """Squares with corners and a shared edge on sample centers, and prisms
    covering 3 and 2 of a voxel's samples (the cases the oracle's rule is
    pinned on), in one mesh over the unit box."""
    soup = np.concatenate([square(8.5, 8.5, 24.5, 24.5, 0.3), square(8.5, 8.5, 24.5, 24.5, 0.31, ccw=False),
                           square(4.5, 4.5, 12.5, 20.5, 0.6), square(12.5, 4.5, 27.5, 20.5, 0.6, diagonal=1),
                           prism(((10, 10), (12.4, 10), (10, 12.4)), 0.2, 0.8),
                           prism(((20, 20), (23, 20), (20, 21.3)), 0.1, 0.7)]).reshape(-1, 9)
    got, want = stencil_case(gv, orc, ctx, "ties-16", soup, *UNIT, 16, 0)
    stencil_case(gv, orc, ctx, "ties-16-4", soup, *UNIT, 16, 4)
    # the lone squares are open surfaces: the two fills differ
    ray = orc.voxelize(soup, orc.grid_setup(*UNIT, 16, 0), orc.ORC_FAST)
    assert not np.array_equal(want["l1"], ray["l1"])


def test_stencil_cessna(gv, orc, ctx):
    """files/cessna.obj, the reference's sample mesh."""
    v, f, _ = gv.read_obj(os.path.join(GOLDEN, "cessna.obj"))
    soup = gv.flatten(v, f)
    bmin, bmax, mms = gv.mesh_bbox(v)
    stencil_case(gv, orc, ctx, "cessna-64-4", soup, bmin, bmax, mms, 64, 4)
    stencil_case(gv, orc, ctx, "cessna-128-0", soup, bmin, bmax, mms, 128, 0)


@pytest.mark.parametrize("vc,vc2", [(48, 4), (100, 8)])
def test_stencil_generated_meshes(gv, orc, ctx, vc, vc2):
    cases = [("icosphere", gv.icosphere(4, 0.4, (0.5, 0.5, 0.5), 0.02, seed=1)),
             ("torus", gv.mixed(1, 20000, 5)),
             ("drilled", gv.mixed(2, 12000, 6)),
             ("sliver-knot", gv.torus_knot(p=2, q=5, segments=400, ring=12, tube=0.05, jitter=0.02, extent=0.9,
                                           seed=9))]
    for name, (v, f) in cases:
        soup = gv.flatten(v, f)
        bmin, bmax, mms = gv.mesh_bbox(v)
        stencil_case(gv, orc, ctx, f"{name}-{vc}-{vc2}", soup, bmin, bmax, mms, vc, vc2)


def test_stencil_batch_mixed_modes(gv, orc, ctx):
    """A batch whose meshes alternate between the two Level1 fills."""
    from gpview_amd import workloads
    meshes = workloads.mixed_stream(gv, 10, tri_range=(1000, 40000), level1=(64, 128), level2=(4, 8))
    for i, m in enumerate(meshes):
        m["l1_stencil"] = i % 2 == 0
    say(f"stencil batch: {len(meshes)} meshes")
    grids = ctx.voxelize_batch(meshes)
    try:
        for i, (m, gr) in enumerate(zip(meshes, grids)):
            g = orc.grid_setup(m["bmin"], m["bmax"], m["mms"], m["voxel_count"], m["voxel_count2"],
                               l1_stencil=m["l1_stencil"])
            compare_results(f"batch member {i}", gr.download(), orc.voxelize(m["soup"], g, orc.ORC_FAST), g.n2)
    finally:
        for gr in grids:
            gr.release()


def test_stencil_staged_entry_points(gv, orc, ctx):
    v, f = gv.mixed(0, 6000, 21)
    soup = gv.flatten(v, f)
    bmin, bmax, mms = gv.mesh_bbox(v)
    say("stencil staged")
    grid = ctx.create(soup, bmin, bmax, mms, 64, 4, l1_stencil=True)
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
    g = orc.grid_setup(bmin, bmax, mms, 64, 4, l1_stencil=True)
    compare_results("stencil staged", got, orc.voxelize(soup, g, orc.ORC_FAST), 4)


def test_stencil_headline_torus_knot(gv, orc, ctx):
    """The (3,7) torus knot at 256/8 (1,048,576 triangles).  Its ray-parity
    grid differs in thousands of columns: a fifth of its triangles have
    0 < |det| < 1e-6 for the +z ray and are invisible to the reference's
    Moller-Trumbore test (src/TriRayIntersection.cpp:96); the stencil counts
    them."""
    from gpview_amd import workloads
    m = workloads.plain_large(gv)
    got, want = stencil_case(gv, orc, ctx, "knot-256-8", m["soup"], m["bmin"], m["bmax"], m["mms"], 256, 8)
    assert want["l1_inside"] > 800000


def test_stencil_batch_edge_cases(gv, orc, ctx):
    """The stencil fill on the batch edge cases: the eleven-mesh mix (Level2
    off for one, a single triangle, an empty soup), a mesh of vertical
    triangles only (no sample covered) and a lone horizontal square (open)."""
    from test_gpu_batch import small_mix
    meshes = small_mix(gv)
    vert = np.array([[0.2, 0.2, 0.1, 0.7, 0.6, 0.1, 0.2, 0.2, 0.9],
                     [0.7, 0.6, 0.1, 0.7, 0.6, 0.9, 0.2, 0.2, 0.9]], np.float32)
    sq = square(2.3, 2.3, 29.7, 29.7, 0.5)
    for soup in (vert, sq):
        meshes.append({"soup": soup, "bmin": UNIT[0], "bmax": UNIT[1], "mms": UNIT[2], "voxel_count": 16,
                       "voxel_count2": 4, "level2": True, "initial_capacity": 0})
    for m in meshes:
        m["l1_stencil"] = True
    say(f"stencil edge-case batch: {len(meshes)} meshes")
    grids = ctx.voxelize_batch(meshes)
    try:
        for i, (m, gr) in enumerate(zip(meshes, grids)):
            g = orc.grid_setup(m["bmin"], m["bmax"], m["mms"], m["voxel_count"],
                               m["voxel_count2"] if m["level2"] else 0, l1_stencil=True)
            want = orc.voxelize(m["soup"], g, orc.ORC_FAST)
            compare_results(f"edge member {i}", gr.download(), want, g.n2)
            assert gr.info()["l1_inside"] == want["l1_inside"]
    finally:
        for gr in grids:
            gr.release()
