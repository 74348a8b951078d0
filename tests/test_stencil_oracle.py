"""The oracle's GL-stencil-equivalent Level1 fill (orc_stencil_counts /
orc_stencil_inside; ClassifyInOut2x, src/Object.cpp:1446-1601).

The reference's fill needs a GL context, so no run of it pins these; they
are pinned by what the rule and the geometry fix: hand-counted coverage of
sample-aligned squares (the top-left rule: half-open edges, a shared edge
counted once), the slice test at equality, the 2 x 2 majority vote
(`numInsideCount > numSamples^2 / 2.0`, src/Object.cpp:1597), and, for
closed meshes, agreement with the ray parity on every voxel no triangle
touches (both then decide the side of the surface)."""
import numpy as np
import pytest

import gpview_amd as gv
from oracle import orc

UNIT = (np.zeros(3, np.float32), np.ones(3, np.float32), np.float32(1.0))


def unit_grid(vc=16, vc2=0):
    g = orc.grid_setup(*UNIT, vc, vc2, l1_stencil=True)
    assert list(g.n) == [vc, vc, vc]
    return g


def square(x0, y0, x1, y1, z, ccw=True, diagonal=0):
    """Two triangles over [x0,x1] x [y0,y1] at height z (sample units of a
    32-sample axis over [0, 1])."""
    p = [(x0, y0), (x1, y0), (x1, y1), (x0, y1)]
    if diagonal == 0:
        tris = [(p[0], p[1], p[2]), (p[0], p[2], p[3])]
    else:
        tris = [(p[0], p[1], p[3]), (p[1], p[2], p[3])]
    if not ccw:
        tris = [(a, c, b) for a, b, c in tris]
    return np.array([[c / 32.0 for v in t for c in (v[0], v[1], z * 32.0)] for t in tris], np.float32)


@pytest.mark.parametrize("ccw", [True, False])
@pytest.mark.parametrize("diagonal", [0, 1])
def test_square_on_sample_centers_half_open(ccw, diagonal):
    # corners on sample centers 8.5 and 24.5: the left and top edges take
    # their samples, the right and bottom ones do not, the diagonal's count once
    g = unit_grid()
    c = orc.stencil_counts(square(8.5, 8.5, 24.5, 24.5, 0.3, ccw, diagonal), g, g.n[2] - 1)
    want = np.zeros((32, 32), np.int32)
    want[9:25, 8:24] = 1            # rows (y) 9..24, columns (x) 8..23
    assert np.array_equal(c, want)


def test_adjacent_squares_share_their_edge_once():
    g = unit_grid()
    a = square(4.5, 4.5, 12.5, 20.5, 0.3)
    b = square(12.5, 4.5, 27.5, 20.5, 0.3, diagonal=1)
    c = orc.stencil_counts(np.concatenate([a, b]), g, g.n[2] - 1)
    want = np.zeros((32, 32), np.int32)
    want[5:21, 4:27] = 1
    assert np.array_equal(c, want)


def test_slice_keeps_planes_at_or_below_its_center():
    g = unit_grid()
    zc = np.float32((5 + 0.5) / 16.0)   # the center of slice 5, exact in float
    assert orc.tables(g)["r1"][2][5] == zc
    sq = square(2.3, 2.3, 6.7, 6.7, float(zc))   # samples 2..6 on both axes
    assert orc.stencil_counts(sq, g, 4).sum() == 0
    assert orc.stencil_counts(sq, g, 5).sum() == 25
    assert orc.stencil_counts(sq, g, 15).sum() == 25
    # a vertical triangle covers nothing
    vert = np.array([[0.3, 0.3, 0.1, 0.6, 0.6, 0.1, 0.3, 0.3, 0.9]], np.float32)
    assert orc.stencil_counts(vert, g, 15).sum() == 0


def prism(foot, z0, z1):
    """Closed prism over a triangle footprint (sample units, 32 per axis)."""
    a, b, c = [(x / 32.0, y / 32.0) for x, y in foot]
    lo = [(p[0], p[1], z0) for p in (a, b, c)]
    hi = [(p[0], p[1], z1) for p in (a, b, c)]
    t = [(lo[0], lo[2], lo[1]), (hi[0], hi[1], hi[2])]
    for i in range(3):
        j = (i + 1) % 3
        t += [(lo[i], lo[j], hi[j]), (lo[i], hi[j], hi[i])]
    return np.array([[c for v in tri for c in v] for tri in t], np.float32)


@pytest.mark.parametrize("foot,odd", [
    (((10, 10), (12.4, 10), (10, 12.4)), 3),   # samples (10.5|11.5, 10.5), (10.5, 11.5) of voxel (5, 5)
    (((10, 10), (13, 10), (10, 11.3)), 2),     # (10.5, 10.5), (11.5, 10.5)
])
def test_majority_of_four_samples(foot, odd):
    g = unit_grid()
    m = prism(foot, 0.2, 0.8)
    counts = orc.stencil_counts(m, g, 8)
    assert (counts % 2).sum() == odd and (counts[10:12, 10:12] % 2).sum() == odd
    ins = orc.stencil_inside(m, g)
    want = np.zeros((16, 16, 16), np.uint8)
    if odd > 2:
        want[3:13, 5, 5] = 1            # slice centers 0.219 .. 0.781 lie between the caps
    assert np.array_equal(ins, want)


def closed_meshes():
    v, f = gv.icosphere(3, 0.37, (0.5, 0.5, 0.5), 0.03, seed=3)
    yield "icosphere", v, f
    v, f = gv.torus_knot(p=2, q=3, segments=192, ring=24, tube=0.25, jitter=0.01, extent=0.9, seed=4)
    yield "torus_knot", v, f
    for kind in range(3):
        v, f = gv.mixed(kind, 3000, 11 + kind)
        yield f"mixed{kind}", v, f


@pytest.mark.parametrize("vc", [24, 48, 72])
def test_closed_meshes_agree_with_ray_parity(vc):
    # every voxel no triangle touches lies on one side of a closed surface,
    # and both fills decide that side; boundary voxels come from the SAT
    for name, v, f in closed_meshes():
        soup = gv.flatten(v, f)
        bmin, bmax, mms = gv.mesh_bbox(v)
        a = orc.voxelize(soup, orc.grid_setup(bmin, bmax, mms, vc, 0), orc.ORC_FAST, bins=False)
        b = orc.voxelize(soup, orc.grid_setup(bmin, bmax, mms, vc, 0, l1_stencil=True), orc.ORC_FAST, bins=False)
        assert np.array_equal(a["l1"], b["l1"]), name
        assert a["l1_inside"] > 0, name


def test_open_mesh_differs_from_ray_parity():
    # a lone square is no closed surface: the stencil counts it below each
    # slice above it, the +z rays above it; voxels under and over it swap
    g = orc.grid_setup(*UNIT, 16, 0, l1_stencil=True)
    g0 = orc.grid_setup(*UNIT, 16, 0)
    sq = square(2.3, 2.3, 29.7, 29.7, 0.5)
    a = orc.voxelize(sq, g0, orc.ORC_FAST, bins=False)["l1"].reshape(16, 16, 16)
    b = orc.voxelize(sq, g, orc.ORC_FAST, bins=False)["l1"].reshape(16, 16, 16)
    assert a[2, 4, 8] == 1 and b[2, 4, 8] == 0
    assert a[12, 4, 8] == 0 and b[12, 4, 8] == 1
    # on the square's diagonal the +z rays meet both triangles' shared edge
    # and count it twice (u + v <= 1 is inclusive, src/TriRayIntersection.cpp:
    # 113-116); the top-left rule counts its samples once
    assert a[2, 8, 8] == 0 and b[12, 8, 8] == 1


def exact_votes(soup, g, cols, margin=1.0 / 128):
    """Per column (i, j): [nz] votes of the 2 x 2 samples' exact crossing
    parities (double geometry on the float vertices, crossings at or below
    each slice's ray-origin z), or -1 where a sample lies within `margin`
    samples of an edge of a triangle near it (snapping to 1/256 moves a
    vertex by at most 1/512 per axis, which could decide those)."""
    v = soup.reshape(-1, 3, 3).astype(np.float64)
    W, H, nz = 2 * g.n[0], 2 * g.n[1], g.n[2]
    l, r, b, t = float(g.bmin[0]), float(g.bmax[0]), float(g.bmin[1]), float(g.bmax[1])
    px = (v[:, :, 0] - l) / (r - l) * W          # window coordinates in samples
    py = (v[:, :, 1] - b) / (t - b) * H
    rz = orc.tables(g)["r1"][2].astype(np.float64)
    n = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    lo_x, hi_x, lo_y, hi_y = px.min(1), px.max(1), py.min(1), py.max(1)
    out = {}
    for i, j in cols:
        near = np.nonzero((lo_x <= 2 * i + 2) & (hi_x >= 2 * i) & (lo_y <= 2 * j + 2) & (hi_y >= 2 * j))[0]
        votes = np.zeros(nz, np.int64)
        amb = False
        for sx, sy in ((2 * i, 2 * j), (2 * i + 1, 2 * j), (2 * i, 2 * j + 1), (2 * i + 1, 2 * j + 1)):
            qx, qy = sx + 0.5, sy + 0.5
            X, Y = px[near], py[near]
            area = (X[:, 1] - X[:, 0]) * (Y[:, 2] - Y[:, 0]) - (Y[:, 1] - Y[:, 0]) * (X[:, 2] - X[:, 0])
            sgn = np.sign(area)
            d = []
            for a, c in ((0, 1), (1, 2), (2, 0)):
                ex, ey = X[:, c] - X[:, a], Y[:, c] - Y[:, a]
                ln = np.hypot(ex, ey)
                d.append(sgn * (ex * (qy - Y[:, a]) - ey * (qx - X[:, a])) / np.where(ln > 0, ln, 1))
            d = np.stack(d, 1)
            cover = (d > margin).all(1) & (area != 0)
            miss = (d < -margin).any(1) | (area == 0)
            if not (cover | miss).all():
                amb = True
                break
            k = near[cover]
            x = l + qx * ((r - l) / W)
            y = b + qy * ((t - b) / H)
            zc = v[k, 0, 2] - (n[k, 0] * (x - v[k, 0, 0]) + n[k, 1] * (y - v[k, 0, 1])) / n[k, 2]
            votes += (np.searchsorted(np.sort(zc), rz, side="right") % 2)
        out[(i, j)] = np.full(nz, -1) if amb else (votes > 2).astype(np.int64)
    return out


def test_fine_mesh_matches_exact_crossings():
    """The headline mesh (the (3,7) torus knot, 1,048,576 triangles, Level1
    256): on sampled columns the fill equals the vote of exact crossing
    parities.  Among them are columns where the ray-parity fill misses a
    crossing: a fifth of this mesh's triangles have 0 < |det| < 1e-6 for the
    +z ray, which the reference's Moller-Trumbore test rejects
    (src/TriRayIntersection.cpp:96); the stencil counts every triangle whose
    projection has area."""
    from gpview_amd import workloads
    m = workloads.plain_large(gv)
    soup = m["soup"].reshape(-1, 9)
    g = orc.grid_setup(m["bmin"], m["bmax"], m["mms"], 256, 0, l1_stencil=True)
    nx, ny, nz = g.n[0], g.n[1], g.n[2]
    st = orc.stencil_inside(soup, g)
    ray = orc.voxelize(soup, orc.grid_setup(m["bmin"], m["bmax"], m["mms"], 256, 0), orc.ORC_FAST,
                       bins=False)["l1"].reshape(nz, ny, nx)
    free = ray != 2
    diff = np.argwhere(((st != ray) & free).any(0))
    same = np.argwhere((ray == 1).any(0) & ~((st != ray) & free).any(0))
    assert len(diff) > 1000            # columns where the ray parity misses a crossing
    rng = np.random.default_rng(5)
    pick = [tuple(int(a) for a in diff[q][::-1]) for q in rng.choice(len(diff), 48, replace=False)]
    pick += [tuple(int(a) for a in same[q][::-1]) for q in rng.choice(len(same), 48, replace=False)]
    ex = exact_votes(soup, g, pick)
    checked = [0, 0]
    for q, ((i, j), e) in enumerate(ex.items()):
        if e[0] < 0:
            continue
        assert np.array_equal(st[:, j, i], e), (i, j)
        checked[q >= 48] += 1
    assert min(checked) >= 10, checked
