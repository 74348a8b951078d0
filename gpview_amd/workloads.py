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
                versions of the mixed kinds at Level1 512, Level2 16.
"""
import numpy as np


def mesh_entry(gv, verts, faces, vc, vc2, level2=True):
    soup = gv.flatten(verts, faces)
    bmin, bmax, mms = gv.mesh_bbox(verts)
    return {"soup": soup, "bmin": bmin, "bmax": bmax, "mms": mms, "voxel_count": int(vc),
            "voxel_count2": int(vc2), "level2": level2, "initial_capacity": 0}


def plain_small(gv, seed=1):
    v, f = gv.icosphere(4, 0.4, (0.5, 0.5, 0.5), 0.02, seed=seed)
    return mesh_entry(gv, v, f, 64, 4)


def plain_large(gv, seed=2):
    v, f = gv.torus_knot(p=3, q=7, segments=2048, ring=256, tube=0.3, jitter=0.01, extent=0.9, seed=seed)
    return mesh_entry(gv, v, f, 256, 8)


def mixed_stream(gv, n, seed=2026, tri_range=(1000, 200000), level1=(64, 128, 256), level2=(4, 8)):
    """n meshes of the mixed-size stream; mesh i depends only on (seed, i)."""
    out = []
    lo, hi = np.log(tri_range[0]), np.log(tri_range[1])
    for i in range(n):
        r = np.random.default_rng([seed, i])
        kind = int(r.integers(0, 3))
        target = int(np.exp(r.uniform(lo, hi)))
        vc = int(r.choice(level1))
        vc2 = int(r.choice(level2))
        v, f = gv.mixed(kind, target, seed * 1000003 + i)
        e = mesh_entry(gv, v, f, vc, vc2)
        e["kind"] = kind
        out.append(e)
    return out


def stress(gv, n, seed=77, tri_range=(100000, 2000000), vc=512, vc2=16):
    """Exactness stress meshes: lattice-snapped vertices (p = 0.3), about 5%
    of triangles coplanar with Level1 voxel faces, 10% slivers."""
    out = []
    lo, hi = np.log(tri_range[0]), np.log(tri_range[1])
    for i in range(n):
        r = np.random.default_rng([seed, i])
        kind = int(r.integers(0, 3))
        target = int(np.exp(r.uniform(lo, hi)))
        v, f = gv.mixed(kind, target, seed * 1000003 + i)
        v, f = gv.stress(v, f, vc, vc2, p_snap=0.3, f_sliver=0.1, f_coplanar=0.05, seed=seed * 7919 + i)
        e = mesh_entry(gv, v, f, vc, vc2)
        e["kind"] = kind
        out.append(e)
    return out
