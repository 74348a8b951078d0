"""Pins the four center / ray-origin tables.

Every box center and ray origin of the path comes from one of four per-axis
tables.  They are restated here in numpy from the reference's expressions,
with every float/double promotion written out:

  Level1 SAT   boxMidPoint = (p + 0.5)*boxExtents*2 + bBoxMin      src/Object.cpp:2329-2331
               boxExtents  = Float3(gridSize / 2.0)                 src/Object.cpp:2266
  Level1 ray   voxelCenter = bBoxMin + (i + 0.5)*gridSize           src/Object.cpp:743-745
  Level2 SAT   midX = (i + 0.5)*gridSize + bBoxMin (float)          src/Object.cpp:2369-2371
               center = (p + 0.5)*gridSize2 + mid - boxExtentsL1    src/Object.cpp:2389-2391
               boxExtentsLevel1 = Float3(gridSize / 2.0)            src/Object.cpp:2353
  Level2 ray   minX = i*gridSize + bBoxMin (int*float: all float)   src/Object.cpp:1134-1136
               voxelCenter = minPoint + (p + 0.5)*gridSize2         src/Object.cpp:1158-1160

(p + 0.5 is a double, so each expression is evaluated in double from float
operands and rounded to float on assignment, except minPoint, which is
float arithmetic throughout.)  The oracle's tables (orc.tables) and the
tables libgpv uploads (gpv_grid_centers) must equal the restatement on
grids whose sizes are not powers of two, where rounding decides.  The test
also shows it can fail: the plausible mistakes (all-float arithmetic, the
Level2 sum reassociated, minPoint promoted to double, mid kept in double)
each change some entries of these grids.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN

f32, f64 = np.float32, np.float64


def restate(bmin, g, g2, n, n2):
    """The four tables of one axis from the reference's expressions."""
    ext1 = f32(f64(g) / 2.0)
    p = np.arange(n, dtype=f64)
    c1 = ((p + 0.5) * f64(ext1) * 2 + f64(bmin)).astype(f32)
    r1 = (f64(bmin) + (p + 0.5) * f64(g)).astype(f32)
    out = {"c1": c1, "r1": r1}
    if n2:
        i = np.repeat(np.arange(n), n2)
        q = np.tile(np.arange(n2, dtype=f64), n)
        mid = ((i + 0.5) * f64(g) + f64(bmin)).astype(f32)
        out["c2"] = ((q + 0.5) * f64(g2) + mid.astype(f64) - f64(ext1)).astype(f32)
        minp = (i.astype(f32) * f32(g)).astype(f32) + f32(bmin)
        assert minp.dtype == f32
        out["r2"] = (minp.astype(f64) + (q + 0.5) * f64(g2)).astype(f32)
    return out


def mistakes(bmin, g, g2, n, n2):
    """Plausible restatement errors, each expected to differ somewhere."""
    ext1 = f32(f64(g) / 2.0)
    p = np.arange(n, dtype=f32)
    i = np.repeat(np.arange(n), n2)
    q = np.tile(np.arange(n2, dtype=f64), n)
    mid = ((i + 0.5) * f64(g) + f64(bmin)).astype(f32)
    minp = (i.astype(f32) * f32(g)).astype(f32) + f32(bmin)
    return {
        "c1 all float": ("c1", ((p + f32(0.5)) * ext1 * f32(2) + f32(bmin)).astype(f32)),
        "r1 all float": ("r1", (f32(bmin) + (p + f32(0.5)) * f32(g)).astype(f32)),
        "c2 all float": ("c2", (((q.astype(f32) + f32(0.5)) * f32(g2) + mid) - ext1).astype(f32)),
        "c2 reassociated (mid - ext1 first, float)": ("c2", ((q + 0.5) * f64(g2) + f64(mid - ext1)).astype(f32)),
        "c2 mid kept in double": ("c2", ((q + 0.5) * f64(g2) + ((i + 0.5) * f64(g) + f64(bmin)) - f64(ext1))
                                  .astype(f32)),
        "r2 minPoint in double": ("r2", ((i * f64(g) + f64(bmin)) + (q + 0.5) * f64(g2)).astype(f32)),
        "r2 all float": ("r2", (minp + (q.astype(f32) + f32(0.5)) * f32(g2)).astype(f32)),
    }


def grids(gv):
    v, _, _ = gv.read_obj(os.path.join(GOLDEN, "cessna.obj"))
    bm = gv.mesh_bbox(v)
    out = [("cessna-8-4", bm, 8, 4), ("cessna-128-16", bm, 128, 16), ("cessna-100-5", bm, 100, 5)]
    from gpview_amd import workloads
    m = workloads.plain_large(gv)
    out.append(("torus-knot-256-8", (m["bmin"], m["bmax"], m["mms"]), 256, 8))
    out.append(("torus-knot-512-16", (m["bmin"], m["bmax"], m["mms"]), 512, 16))
    return out


def test_tables_equal_restatement(gv, orc):
    seen = {k: 0 for k in mistakes(f32(0.1), f32(0.3), f32(0.1), 4, 2)}
    for name, (bmin, bmax, mms), vc, vc2 in grids(gv):
        info = gv.grid_derive(bmin, bmax, mms, vc, vc2)
        og = orc.grid_setup(bmin, bmax, mms, vc, vc2)
        otab = orc.tables(og)
        assert list(info.n) == list(og.n) and info.n2 == og.n2
        for a in range(3):
            n, g, g2 = info.n[a], f32(info.g[a]), f32(info.g2[a])
            assert g == f32(og.g[a]) and g2 == f32(og.g2[a])
            # a grid size with a short mantissa would make every expression exact
            assert np.frexp(g)[0] * 2 ** 24 % 2 ** 8 != 0, (name, a, g)
            want = restate(f32(bmin[a]), g, g2, n, vc2)
            for w, key in enumerate(("c1", "r1", "c2", "r2")):
                lib_tab = gv.grid_centers(info, w, a)
                assert np.array_equal(lib_tab.view(np.uint32), want[key].view(np.uint32)), (name, a, key)
                assert np.array_equal(otab[key][a].view(np.uint32), want[key].view(np.uint32)), (name, a, key)
            for mname, (key, wrong) in mistakes(f32(bmin[a]), g, g2, n, vc2).items():
                seen[mname] += int(np.count_nonzero(wrong.view(np.uint32) != want[key].view(np.uint32)))
    # every plausible mistake changes entries on these grids, so the
    # equalities above could fail
    for mname, cnt in seen.items():
        assert cnt > 0, f"'{mname}' is indistinguishable on these grids"


def test_tables_cuda_level2_float_expression(gv):
    """cuda_level2: both Level2 tables are the CUDA kernels' float expression
    (2p+1)*ext2 + mid - ext1 evaluated left to right in float
    (cuda/CUDAClassifyTessellation.cu:423-425, 472-474)."""
    v, _, _ = gv.read_obj(os.path.join(GOLDEN, "cessna.obj"))
    bmin, bmax, mms = gv.mesh_bbox(v)
    info = gv.grid_derive(bmin, bmax, mms, 100, 5)
    for a in range(3):
        n, g, g2 = info.n[a], f32(info.g[a]), f32(info.g2[a])
        ext1, ext2 = f32(f64(g) / 2.0), f32(f64(g2) / 2.0)
        i = np.repeat(np.arange(n), 5)
        q = np.tile(np.arange(5), n)
        mid = ((i + 0.5) * f64(g) + f64(bmin[a])).astype(f32)
        want = (((2 * q + 1).astype(f32) * ext2).astype(f32) + mid).astype(f32) - ext1
        for w in (2, 3):
            got = gv.grid_centers(info, w, a, cuda_level2=True)
            assert np.array_equal(got.view(np.uint32), want.astype(f32).view(np.uint32)), (a, w)
        assert not np.array_equal(gv.grid_centers(info, 2, a), want.astype(f32)) or a < 0


def test_grid_centers_rejects_bad_input(gv):
    info = gv.grid_derive(np.zeros(3, f32), np.ones(3, f32), f32(1), 8, 4, level2=False)
    with pytest.raises(gv.GpvError):
        gv.grid_centers(info, 2, 0)
    with pytest.raises(gv.GpvError):
        gv.grid_centers(info, 0, 3)
