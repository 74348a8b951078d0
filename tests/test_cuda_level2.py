"""The CUDA-path Level2 mode (gpv_params.cuda_level2, orc_grid.cuda_l2): the
Level2 arithmetic of the reference's CUDA kernels
(cuda/CUDAClassifyTessellation.cu:403-504), kept for callers comparing with
the reference's GPU build.

Parity against that build is unpinned: it needs nvcc and an NVIDIA GPU, and
nvcc may contract (2p+1)*ext2 + mid into a fused multiply-add, which this
mode does not.  What is pinned here:
- closed form: on a grid whose sizes are powers of two and whose minimum is
  0, every Level2 center and ray origin is exactly representable, so the
  CUDA float expression and the CPU composition's double expressions give
  the same values (a dropped term, wrong sign or swapped extent fails);
- elsewhere the two center tables agree to a few ulps, and the CUDA tables
  are non-decreasing in the sub-index;
- the oracle's literal Level2 parity loop over the whole column (every bin,
  deduplicated) equals its column-factorised form without the zmax rule;
- Level1 results do not depend on the mode.
"""
import numpy as np
import pytest

from conftest import compare_results


def _pow2_grid(orc, cuda):
    # bbox [0, 1]^3 before the pad is applied by the caller: use a grid set up
    # directly with power-of-two extents
    g = orc.grid_setup([0.0, 0.0, 0.0], [1.0, 2.0, 4.0], 4.0, 16, 4, cuda_l2=cuda)
    return g


def test_tables_exact_on_power_of_two_grid(orc):
    gc = _pow2_grid(orc, True)
    gr = _pow2_grid(orc, False)
    # sizes are powers of two: g = extent / n exactly, g2 = g / 4, ext = g / 2
    for a in range(3):
        assert gc.g[a] == np.float32(gc.bmax[a] - gc.bmin[a]) / gc.n[a]
    tc, tr = orc.tables(gc), orc.tables(gr)
    for name in ("c2", "r2"):
        for a in range(3):
            assert np.array_equal(tc[name][a], tr[name][a]), (name, a)
    # and the value is the exact sub-center (i + (p + 0.5)/n2) * g
    n2 = gc.n2
    for a in range(3):
        i = np.repeat(np.arange(gc.n[a]), n2)
        p = np.tile(np.arange(n2), gc.n[a])
        want = (i + (p + 0.5) / n2) * np.float64(gc.g[a]) + gc.bmin[a]
        assert np.array_equal(tc["c2"][a].astype(np.float64), want)


def test_tables_close_and_monotone(orc, gv):
    v, f = gv.mixed(1, 3000, 5)
    soup = gv.flatten(v, f)
    bmin, bmax, mms = orc.bbox(v)
    gc = orc.grid_setup(bmin, bmax, mms, 48, 8, cuda_l2=True)
    gr = orc.grid_setup(bmin, bmax, mms, 48, 8)
    tc, tr = orc.tables(gc), orc.tables(gr)
    differ = 0
    for name in ("c2", "r2"):
        for a in range(3):
            c = tc[name][a].reshape(gc.n[a], gc.n2)
            r = tr[name][a].reshape(gr.n[a], gr.n2)
            assert np.all(np.diff(c, axis=1) >= 0)
            ulp = np.spacing(np.abs(r)).max()
            assert np.abs(c.astype(np.float64) - r).max() <= 8 * ulp
            differ += int(np.count_nonzero(c != r))
    assert differ > 0   # the two compositions are not the same arithmetic


def test_oracle_parity_literal_equals_fast(orc, gv):
    v, f = gv.mixed(0, 1500, 7)
    soup = gv.flatten(v, f)
    bmin, bmax, mms = orc.bbox(v)
    g = orc.grid_setup(bmin, bmax, mms, 24, 4, cuda_l2=True)
    a = orc.voxelize(soup, g, orc.ORC_EXACT)
    b = orc.voxelize(soup, g, orc.ORC_FAST)
    assert np.array_equal(a["l2"], b["l2"]) and np.array_equal(a["l1"], b["l1"])


def test_level1_independent_of_mode(orc, gv):
    v, f = gv.mixed(2, 2500, 9)
    soup = gv.flatten(v, f)
    bmin, bmax, mms = orc.bbox(v)
    a = orc.voxelize(soup, orc.grid_setup(bmin, bmax, mms, 32, 4, cuda_l2=True))
    b = orc.voxelize(soup, orc.grid_setup(bmin, bmax, mms, 32, 4))
    assert np.array_equal(a["l1"], b["l1"]) and np.array_equal(a["bidx"], b["bidx"])


@pytest.mark.gpu
@pytest.mark.parametrize("kind,target,vc,vc2,seed", [(0, 4000, 48, 8, 21), (1, 6000, 64, 4, 22), (2, 5000, 40, 16, 23)])
def test_gpu_cuda_level2_vs_oracle(gv, orc, ctx, kind, target, vc, vc2, seed):
    v, f = gv.mixed(kind, target, seed)
    soup = gv.flatten(v, f)
    bmin, bmax, mms = gv.mesh_bbox(v)
    g = orc.grid_setup(bmin, bmax, mms, vc, vc2, cuda_l2=True)
    want = orc.voxelize(soup, g)
    grid = ctx.voxelize(soup, bmin, bmax, mms, vc, vc2, True, cuda_level2=True)
    try:
        got = grid.download()
    finally:
        grid.release()
    compare_results(f"cuda-level2 k{kind}", got, want, g.n2)


@pytest.mark.gpu
def test_gpu_cuda_level2_in_batch(gv, orc, ctx):
    from gpview_amd import workloads
    ms = workloads.mixed_stream(gv, 4, seed=31, tri_range=(1000, 6000), level1=(32, 48), level2=(4, 8))
    for i, m in enumerate(ms):
        m["cuda_level2"] = i % 2 == 1
    grids = ctx.voxelize_batch(ms)
    try:
        for i, (m, gr) in enumerate(zip(ms, grids)):
            g = orc.grid_setup(m["bmin"], m["bmax"], m["mms"], m["voxel_count"], m["voxel_count2"],
                               cuda_l2=m["cuda_level2"])
            compare_results(f"batch mesh {i}", gr.download(), orc.voxelize(m["soup"], g), g.n2)
    finally:
        for gr in grids:
            gr.release()
