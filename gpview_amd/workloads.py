"""Seeded synthetic workloads of the voxelizer benchmarks.

Every mesh comes from the library's own C++ generators (include/gpv_mesh.h),
so a seed fixes the triangles bit for bit on any host.

- plain_small:  closed icosphere, subdivision 4 (5,120 triangles), radius 0.4
                at the cube center, +-2% radial jitter; Level1 64, Level2 4.
- plain_large:  (3,7) torus-knot tube, 2,048 segments x 256 ring vertices
                (1,048,576 triangles), extent 0.9; Level1 256, Level2 8.
- mixed_stream: meshes with triangle counts log-uniform in [1k, 200k] of
                three kinds (jittered icospheres; tori with major/minor ratio
                U[2,5]; rotated boxes with 1-6 drilled cylindrical holes),
                Level1 drawn from {64, 128, 256}, Level2 from {4, 8}.
- stress:       lattice-snapped (p = 0.3), 10% sliver, 5% face-coplanar
                versions of the mixed kinds, 100k-2M triangles, at Level1
                512, Level2 16.
- served:       Poisson request streams over a device-resident library made
                of the plain_large mesh and mixed-stream members of 1k-2M
                triangles, each request at Level1 {64,128,256,512} and
                Level2 {4,8,16}.

Member i of a stream depends only on (seed, i): *_spec(i) gives its kind,
size and resolution without building it, *_member(i) builds it.
"""
import numpy as np

MIXED_SEED, STRESS_SEED = 2026, 77


def mesh_entry(gv, verts, faces, vc, vc2, level2=True):
    soup = gv.flatten(verts, faces)
    bmin, bmax, mms = gv.mesh_bbox(verts)
    return {"soup": soup, "bmin": bmin, "bmax": bmax, "mms": mms, "voxel_count": int(vc),
            "voxel_count2": int(vc2), "level2": level2, "initial_capacity": 0}


def plain_small(gv, seed=1):
    v, f = gv.icosphere(4, 0.4, (0.5, 0.5, 0.5), 0.02, seed=seed)
    return mesh_entry(gv, v, f, 64, 4)


def plain_large(gv, seed=2, vc=256, vc2=8):
    v, f = gv.torus_knot(p=3, q=7, segments=2048, ring=256, tube=0.3, jitter=0.01, extent=0.9, seed=seed)
    return mesh_entry(gv, v, f, vc, vc2)


def mixed_spec(i, seed=MIXED_SEED, tri_range=(1000, 200000), level1=(64, 128, 256), level2=(4, 8)):
    """(kind, target triangles, Level1, Level2) of mixed-stream member i."""
    lo, hi = np.log(tri_range[0]), np.log(tri_range[1])
    r = np.random.default_rng([seed, i])
    kind = int(r.integers(0, 3))
    target = int(np.exp(r.uniform(lo, hi)))
    vc = int(r.choice(level1))
    vc2 = int(r.choice(level2))
    return kind, target, vc, vc2


def mixed_member(gv, i, seed=MIXED_SEED, **kw):
    kind, target, vc, vc2 = mixed_spec(i, seed, **kw)
    v, f = gv.mixed(kind, target, seed * 1000003 + i)
    e = mesh_entry(gv, v, f, vc, vc2)
    e["kind"] = kind
    e["index"] = i
    return e


def mixed_stream(gv, n, seed=MIXED_SEED, tri_range=(1000, 200000), level1=(64, 128, 256), level2=(4, 8)):
    """n meshes of the mixed-size stream; mesh i depends only on (seed, i)."""
    return [mixed_member(gv, i, seed, tri_range=tri_range, level1=level1, level2=level2) for i in range(n)]


def stress_spec(i, seed=STRESS_SEED, tri_range=(100000, 2000000)):
    """(kind, target triangles) of stress member i."""
    lo, hi = np.log(tri_range[0]), np.log(tri_range[1])
    r = np.random.default_rng([seed, i])
    kind = int(r.integers(0, 3))
    target = int(np.exp(r.uniform(lo, hi)))
    return kind, target


def stress_member(gv, i, seed=STRESS_SEED, tri_range=(100000, 2000000), vc=512, vc2=16):
    """Exactness stress mesh i: lattice-snapped vertices (p = 0.3), about 5%
    of triangles coplanar with Level1 voxel faces, 10% slivers."""
    kind, target = stress_spec(i, seed, tri_range)
    v, f = gv.mixed(kind, target, seed * 1000003 + i)
    v, f = gv.stress(v, f, vc, vc2, p_snap=0.3, f_sliver=0.1, f_coplanar=0.05, seed=seed * 7919 + i)
    e = mesh_entry(gv, v, f, vc, vc2)
    e["kind"] = kind
    e["index"] = i
    return e


def stress(gv, n, seed=STRESS_SEED, tri_range=(100000, 2000000), vc=512, vc2=16):
    return [stress_member(gv, i, seed, tri_range, vc, vc2) for i in range(n)]


SERVED_L1 = ((64, 128, 256, 512), (0.3, 0.3, 0.3, 0.1))
SERVED_L2 = ((4, 8, 16), (0.45, 0.45, 0.1))


def served_library(gv, n_library=48, seed=5):
    """Library of the served stream: the plain_large torus knot (entry 0) and
    n_library - 1 mixed-stream members of 1k-2M triangles."""
    lib = [dict(plain_large(gv), kind=3, index=0)]
    for i in range(1, n_library):
        lib.append(mixed_member(gv, i, seed, tri_range=(1000, 2000000)))
    return lib


def served_requests(n_requests, n_library=48, rate=0.0, seed=5):
    """Request sequence: (library index, Level1, Level2, inter-arrival gap
    in seconds) arrays; rate 0 means all at once."""
    rng = np.random.default_rng(seed)
    pick = rng.integers(0, n_library, n_requests)
    l1 = rng.choice(SERVED_L1[0], n_requests, p=SERVED_L1[1])
    l2 = rng.choice(SERVED_L2[0], n_requests, p=SERVED_L2[1])
    gaps = rng.exponential(1.0 / rate, n_requests) if rate > 0 else np.zeros(n_requests)
    return pick, l1, l2, gaps
