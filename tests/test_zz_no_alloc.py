"""gpv_voxelize_noalloc against gpv_voxelize, through the C ABI.

The no-allocation entry point runs the same kernels with every transient
buffer carved from the library's fixed device scratch array and the results
written to the caller's device buffers.  Its Level1/Level2 bitplanes,
boundary index list and counts must equal what gpv_voxelize produces for the
same mesh, bit for bit, on consecutive calls (the second smaller, so stale
scratch contents from the first are present), with the caller's buffers
filled with garbage beforehand.
"""
import ctypes

import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu

_hip = None


def d2h(ptr, nbytes, dtype):
    """hipMemcpy device -> host of nbytes at ptr."""
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    out = np.zeros(max(nbytes // np.dtype(dtype).itemsize, 1), dtype)
    if nbytes:
        assert _hip.hipMemcpy(out.ctypes.data, ptr, nbytes, 2) == 0
    return out[:nbytes // np.dtype(dtype).itemsize]


def garbage(ctx, nbytes):
    return ctx.upload(np.full(max(nbytes, 8), 0xA5, np.uint8))


def run_pair(gv, ctx, name, verts, faces, vc, vc2, level2=True, l1_stencil=False):
    soup = gv.flatten(verts, faces)
    bmin, bmax, mms = gv.mesh_bbox(verts)
    say(f"noalloc {name}: plain T={len(soup)} {vc}/{vc2}")
    g = ctx.voxelize(soup, bmin, bmax, mms, vc, vc2, level2, l1_stencil=l1_stencil)
    try:
        want_bits = g.download_bits()
        want = g.download()
        winfo = g.info()
    finally:
        g.release()
    nx, ny, nz = winfo["n"]
    w1 = nz * ny * ((nx + 63) // 64)
    n2 = winfo["n2"]
    w2 = (n2 ** 3 + 63) // 64 if n2 else 0
    nb_cap = winfo["nb"] + 17
    dsoup = ctx.upload(soup)
    bufs = [garbage(ctx, w1 * 8), garbage(ctx, w1 * 8), garbage(ctx, nb_cap * 4),
            garbage(ctx, nb_cap * w2 * 8), garbage(ctx, nb_cap * w2 * 8)]
    try:
        say(f"noalloc {name}: no-allocation entry point")
        info = ctx.voxelize_noalloc(dsoup, len(soup), bmin, bmax, mms, vc, vc2, level2, bufs[0], bufs[1], w1,
                                    bufs[2], bufs[3], bufs[4], nb_cap, l1_stencil=l1_stencil)
        nb = info["nb"]
        assert nb == winfo["nb"], f"{name}: nb {nb} != {winfo['nb']}"
        got = (d2h(bufs[0], w1 * 8, np.uint64), d2h(bufs[1], w1 * 8, np.uint64),
               d2h(bufs[3], nb * w2 * 8, np.uint64), d2h(bufs[4], nb * w2 * 8, np.uint64))
        bidx = d2h(bufs[2], nb * 4, np.int32)
    finally:
        for p in bufs + [dsoup]:
            ctx.free(p)
    for k, (a, b) in enumerate(zip(got, want_bits)):
        assert np.array_equal(a, b), f"{name}: bitplane {k} differs ({np.count_nonzero(a != b)} words)"
    assert np.array_equal(bidx, want["bidx"]), f"{name}: boundary index list differs"
    for k in ("n", "n2", "nb", "l1_inside", "l1_boundary", "l2_inside", "l2_boundary", "tests_l1", "tests_l2",
              "pairs", "max_tris_per_voxel"):
        assert info[k] == winfo[k], f"{name}: {k} {info[k]} != {winfo[k]}"
    return info


def test_noalloc_matches_voxelize_twice(gv, ctx):
    """Two calls in a row, the second on a smaller mesh."""
    v, f = gv.icosphere(4, 0.4, (0.5, 0.5, 0.5), 0.02, seed=1)
    a = run_pair(gv, ctx, "icosphere-5120-64-4", v, f, 64, 4)
    v, f = gv.mixed(1, 3000, 11)
    b = run_pair(gv, ctx, "torus-3000-32-8", v, f, 32, 8)
    assert b["n1"] < a["n1"] and b["nb"] < a["nb"]


def test_noalloc_level2_off_and_stress(gv, ctx):
    v, f = gv.mixed(2, 20000, 21)
    vs, fs = gv.stress(v, f, 64, 8, p_snap=0.3, f_sliver=0.1, f_coplanar=0.05, seed=22)
    run_pair(gv, ctx, "stress-box-64-8", vs, fs, 64, 8)
    v, f = gv.icosphere(3)
    run_pair(gv, ctx, "level2-off", v, f, 48, 4, level2=False)


def test_noalloc_stencil_fill(gv, ctx):
    """The GL-stencil-equivalent Level1 fill from the scratch array, on a knot
    whose stencil and ray-parity fills differ (ray-invisible triangles)."""
    v, f = gv.torus_knot(p=3, q=7, segments=1024, ring=64, tube=0.3, jitter=0.01, extent=0.9, seed=2)
    a = run_pair(gv, ctx, "knot-131k-128-4-stencil", v, f, 128, 4, l1_stencil=True)
    b = run_pair(gv, ctx, "knot-131k-128-4", v, f, 128, 4)
    assert a["l1_inside"] != b["l1_inside"]


def test_noalloc_rejects_what_does_not_fit(gv, ctx):
    """Oversized grids, too-small result buffers and normals are refused with
    a status, and the entry point still works afterwards."""
    v, f = gv.icosphere(3)
    soup = gv.flatten(v, f)
    bmin, bmax, mms = gv.mesh_bbox(v)
    dsoup = ctx.upload(soup)
    buf = garbage(ctx, 1 << 20)
    try:
        say("noalloc: grid too large for the scratch")
        with pytest.raises(gv.GpvError) as e:
            ctx.voxelize_noalloc(dsoup, len(soup), bmin, bmax, mms, 2048, 4, True, buf, buf, 1 << 17, buf, buf,
                                 buf, 16)
        assert e.value.status == gv.GPV_ERR_TOO_LARGE
        say("noalloc: Level1 planes too small")
        with pytest.raises(gv.GpvError) as e:
            ctx.voxelize_noalloc(dsoup, len(soup), bmin, bmax, mms, 64, 4, True, buf, buf, 10, buf, buf, buf, 16)
        assert e.value.status == gv.GPV_ERR_INVALID
        say("noalloc: boundary count above nb_cap")
        w1 = 64 * 64 * 1
        b1, b2 = garbage(ctx, w1 * 8), garbage(ctx, w1 * 8)
        try:
            with pytest.raises(gv.GpvError) as e:
                ctx.voxelize_noalloc(dsoup, len(soup), bmin, bmax, mms, 64, 4, True, b1, b2, w1, buf, buf, buf, 16)
            assert e.value.status == gv.GPV_ERR_TOO_LARGE
        finally:
            ctx.free(b1)
            ctx.free(b2)
    finally:
        ctx.free(buf)
        ctx.free(dsoup)
    run_pair(gv, ctx, "after-rejections", v, f, 40, 4)
