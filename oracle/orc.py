"""ctypes view of the CPU oracle (oracle/liborc.so) and of the reference's own
predicates compiled into oracle/_ref/libgpvref.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the
checker, and by bench.py's cpu_baseline leg.  The product never imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORC_EXACT, ORC_FAST, ORC_PLAN = 0, 1, 2

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


class Grid(ctypes.Structure):
    """Mirror of orc_grid (gpv_oracle.h)."""
    _fields_ = [
        ("bmin", ctypes.c_float * 3),
        ("bmax", ctypes.c_float * 3),
        ("max_model_size", ctypes.c_float),
        ("n", ctypes.c_int * 3),
        ("g", ctypes.c_float * 3),
        ("ext1", ctypes.c_float * 3),
        ("n2", ctypes.c_int),
        ("g2", ctypes.c_float * 3),
        ("ext2", ctypes.c_float * 3),
        ("cuda_l2", ctypes.c_int),
        ("l1_stencil", ctypes.c_int),
    ]


class _Result(ctypes.Structure):
    _fields_ = [
        ("n1", ctypes.c_int64),
        ("l1", ctypes.POINTER(ctypes.c_uint8)),
        ("prefix", ctypes.POINTER(ctypes.c_int32)),
        ("nb", ctypes.c_int64),
        ("bidx", ctypes.POINTER(ctypes.c_int32)),
        ("n2", ctypes.c_int),
        ("l2", ctypes.POINTER(ctypes.c_uint8)),
        ("list_off", ctypes.POINTER(ctypes.c_int64)),
        ("list_tri", ctypes.POINTER(ctypes.c_int32)),
        ("pairs", ctypes.c_int64),
        ("tests_l1", ctypes.c_int64),
        ("tests_l2", ctypes.c_int64),
        ("l1_inside", ctypes.c_int64),
        ("l1_boundary", ctypes.c_int64),
        ("l2_inside", ctypes.c_int64),
        ("l2_boundary", ctypes.c_int64),
        ("max_tris_per_voxel", ctypes.c_int),
        ("t_stage", ctypes.c_double * 5),
    ]


_lib = None
_ref = None


def build():
    """Compile the oracle (and _ref when the reference tree is mounted)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liborc.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_tri_box_batch.argtypes = [ctypes.c_int64, _f32p, _f32p, _f32p, _u8p]
        L.orc_tri_box_factored_batch.argtypes = [ctypes.c_int64, _f32p, _f32p, _f32p, _u8p]
        L.orc_tri_ray_batch.argtypes = [ctypes.c_int64, _f32p, _f32p, _f32p, _u8p, _f32p]
        L.orc_tri_zray_split_batch.argtypes = [ctypes.c_int64, _f32p, _f32p, _u8p]
        L.orc_bbox.argtypes = [_f32p, ctypes.c_int64, _f32p, _f32p, ctypes.POINTER(ctypes.c_float)]
        L.orc_grid_setup.argtypes = [_f32p, _f32p, ctypes.c_float, ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(Grid)]
        for fn in ("orc_l1_centers", "orc_l1_ray_origins", "orc_l2_centers", "orc_l2_ray_origins"):
            getattr(L, fn).argtypes = [ctypes.POINTER(Grid), ctypes.c_int, _f32p]
        L.orc_vertex_index.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_int]
        L.orc_vertex_index.restype = ctypes.c_int
        L.orc_ray_column_bounds.argtypes = [_f32p, _f32p, _f32p, ctypes.c_double,
                                            np.ctypeslib.ndpointer(dtype=np.float64),
                                            np.ctypeslib.ndpointer(dtype=np.float64)]
        L.orc_ray_column_bounds.restype = ctypes.c_int
        L.orc_voxelize.argtypes = [_f32p, ctypes.c_int64, ctypes.POINTER(Grid), ctypes.c_int,
                                   ctypes.c_int, ctypes.POINTER(_Result)]
        L.orc_voxelize.restype = ctypes.c_int
        L.orc_result_free.argtypes = [ctypes.POINTER(_Result)]
        L.orc_stencil_counts.argtypes = [_f32p, ctypes.c_int64, ctypes.POINTER(Grid), ctypes.c_int,
                                         np.ctypeslib.ndpointer(dtype=np.int32)]
        L.orc_stencil_counts.restype = ctypes.c_int
        L.orc_stencil_inside.argtypes = [_f32p, ctypes.c_int64, ctypes.POINTER(Grid), _u8p]
        L.orc_stencil_inside.restype = ctypes.c_int
        L.orc_face_normal_batch.argtypes = [ctypes.c_int64, _f32p, _f32p]
        L.orc_normalize_batch.argtypes = [ctypes.c_int64, _f32p, _f32p]
        L.orc_average_normals.argtypes = [ctypes.c_int64, np.ctypeslib.ndpointer(np.int64), _f32p, _f32p]
        L.orc_normals.argtypes = [_f32p, ctypes.POINTER(Grid), ctypes.POINTER(_Result), ctypes.c_int,
                                  _f32p, ctypes.c_void_p]
        L.orc_quantize_normals.argtypes = [ctypes.c_int64, _f32p, _u8p]
        L.orc_save_raw.argtypes = [ctypes.POINTER(Grid), ctypes.POINTER(_Result), ctypes.c_int,
                                   ctypes.c_char_p]
        L.orc_save_raw_normals.argtypes = [ctypes.POINTER(Grid), ctypes.POINTER(_Result), _f32p, ctypes.c_void_p,
                                           ctypes.c_int, ctypes.c_char_p]
        _lib = L
    return _lib


def ref():
    """The reference's own TriBoxOverlap / triangle_ray_intersection, or None
    when oracle/_ref was not built (no reference tree on this machine)."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libgpvref.so")
        if not os.path.exists(path):
            return None
        R = ctypes.CDLL(path)
        R.ref_tri_box_batch.argtypes = [ctypes.c_int64, _f32p, _f32p, _f32p, _u8p]
        R.ref_tri_ray_batch.argtypes = [ctypes.c_int64, _f32p, _f32p, _f32p, _u8p, _f32p]
        R.ref_face_normal_batch.argtypes = [ctypes.c_int64, _f32p, _f32p]
        R.ref_normalize_batch.argtypes = [ctypes.c_int64, _f32p, _f32p]
        R.ref_average_normals.argtypes = [ctypes.c_int64, np.ctypeslib.ndpointer(np.int64), _f32p, _f32p]
        _ref = R
    return _ref


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def tri_box(c, h, tv, use_ref=False, factored=False):
    c, h, tv = _f32(c).reshape(-1, 3), _f32(h).reshape(-1, 3), _f32(tv).reshape(-1, 9)
    out = np.zeros(len(c), np.uint8)
    fn = (ref().ref_tri_box_batch if use_ref else
          lib().orc_tri_box_factored_batch if factored else lib().orc_tri_box_batch)
    fn(len(c), c.ravel(), h.ravel(), tv.ravel(), out)
    return out


def tri_ray(tv, o, d, use_ref=False):
    tv, o, d = _f32(tv).reshape(-1, 9), _f32(o).reshape(-1, 3), _f32(d).reshape(-1, 3)
    hit = np.zeros(len(tv), np.uint8)
    t = np.zeros(len(tv), np.float32)
    fn = ref().ref_tri_ray_batch if use_ref else lib().orc_tri_ray_batch
    fn(len(tv), tv.ravel(), o.ravel(), d.ravel(), hit, t)
    return hit, t


def tri_zray_split(tv, o):
    tv, o = _f32(tv).reshape(-1, 9), _f32(o).reshape(-1, 3)
    hit = np.zeros(len(tv), np.uint8)
    lib().orc_tri_zray_split_batch(len(tv), tv.ravel(), o.ravel(), hit)
    return hit


def bbox(verts):
    v = _f32(verts).reshape(-1, 3)
    bmin = np.zeros(3, np.float32)
    bmax = np.zeros(3, np.float32)
    mms = ctypes.c_float()
    lib().orc_bbox(v.ravel(), len(v), bmin, bmax, ctypes.byref(mms))
    return bmin, bmax, np.float32(mms.value)


def grid_setup(bmin, bmax, mms, voxel_count, voxel_count2, cuda_l2=False, l1_stencil=False):
    """orc_grid_setup; cuda_l2 selects the Level2 arithmetic of the
    reference's CUDA kernels (orc_grid.cuda_l2), l1_stencil the
    GL-stencil-equivalent Level1 fill (orc_grid.l1_stencil)."""
    g = Grid()
    lib().orc_grid_setup(_f32(bmin), _f32(bmax), float(mms), int(voxel_count), int(voxel_count2),
                         ctypes.byref(g))
    g.cuda_l2 = 1 if cuda_l2 else 0
    g.l1_stencil = 1 if l1_stencil else 0
    return g


def stencil_counts(tris, g, slice_):
    """orc_stencil_counts: fragment count per sample [2ny, 2nx] of one slice
    of the GL-stencil-equivalent fill."""
    tris = _f32(tris).reshape(-1)
    out = np.zeros(4 * g.n[0] * g.n[1], np.int32)
    assert lib().orc_stencil_counts(tris, len(tris) // 9, ctypes.byref(g), int(slice_), out) == 0
    return out.reshape(2 * g.n[1], 2 * g.n[0])


def stencil_inside(tris, g):
    """orc_stencil_inside: Level1 inside (1) / outside (0) [nz, ny, nx] by
    the GL-stencil-equivalent fill."""
    tris = _f32(tris).reshape(-1)
    out = np.zeros(g.n[0] * g.n[1] * g.n[2], np.uint8)
    assert lib().orc_stencil_inside(tris, len(tris) // 9, ctypes.byref(g), out) == 0
    return out.reshape(g.n[2], g.n[1], g.n[0])


def grid_for(tris, voxel_count, voxel_count2):
    """bbox over the soup's vertices (every vertex is referenced) + grid."""
    bmin, bmax, mms = bbox(np.asarray(tris, np.float32).reshape(-1, 3))
    return grid_setup(bmin, bmax, mms, voxel_count, voxel_count2)


def tables(g):
    out = {}
    for name, fn, per in (("c1", "orc_l1_centers", 1), ("r1", "orc_l1_ray_origins", 1),
                          ("c2", "orc_l2_centers", max(g.n2, 1)),
                          ("r2", "orc_l2_ray_origins", max(g.n2, 1))):
        if per > 1 or name in ("c1", "r1") or g.n2 > 0:
            arrs = []
            for a in range(3):
                buf = np.zeros(g.n[a] * per, np.float32)
                getattr(lib(), fn)(ctypes.byref(g), a, buf)
                arrs.append(buf)
            out[name] = arrs
    return out


def column_bounds(tv, bmin, bmax, safety=1.0):
    lo = np.zeros(2, np.float64)
    hi = np.zeros(2, np.float64)
    ok = lib().orc_ray_column_bounds(_f32(tv).ravel(), _f32(bmin), _f32(bmax), float(safety), lo, hi)
    return bool(ok), lo, hi


def face_normals(tv, use_ref=False):
    tv = _f32(tv).reshape(-1, 9)
    out = np.zeros((len(tv), 3), np.float32)
    (ref().ref_face_normal_batch if use_ref else lib().orc_face_normal_batch)(len(tv), tv.ravel(), out.ravel())
    return out


def normalize(v, use_ref=False):
    v = _f32(v).reshape(-1, 3)
    out = np.zeros_like(v)
    (ref().ref_normalize_batch if use_ref else lib().orc_normalize_batch)(len(v), v.ravel(), out.ravel())
    return out


def average_normals(off, fn, use_ref=False):
    off = np.ascontiguousarray(off, np.int64)
    fn = _f32(fn).reshape(-1, 3)
    out = np.zeros((len(off) - 1, 3), np.float32)
    (ref().ref_average_normals if use_ref else lib().orc_average_normals)(len(off) - 1, off, fn.ravel(),
                                                                          out.ravel())
    return out


def quantize_normals(v):
    v = _f32(v).ravel()
    out = np.zeros(len(v), np.uint8)
    lib().orc_quantize_normals(len(v), v, out)
    return out


def voxelize(tris, g, mode=ORC_FAST, threads=0, normals=False, bins=True):
    """Run the oracle pipeline; returns a dict of numpy arrays and counters
    (with normals=True also l1n [nb,3] and l2n [nb*n2^3,3]; with
    bins=False the Level1 bin lists list_off/list_tri are not copied out,
    which saves 8 bytes per Level1 voxel on large grids)."""
    tris = _f32(tris).reshape(-1)
    T = len(tris) // 9
    r = _Result()
    rc = lib().orc_voxelize(tris, T, ctypes.byref(g), int(mode), int(threads), ctypes.byref(r))
    if rc != 0:
        raise ValueError(f"orc_voxelize failed ({rc})")
    n23 = g.n2 ** 3 if g.n2 > 0 else 0
    try:
        out = {
            "l1": np.ctypeslib.as_array(r.l1, (r.n1,)).copy(),
            "prefix": np.ctypeslib.as_array(r.prefix, (r.n1,)).copy(),
            "bidx": np.ctypeslib.as_array(r.bidx, (r.nb,)).copy() if r.nb else np.zeros(0, np.int32),
            "l2": (np.ctypeslib.as_array(r.l2, (r.nb * n23,)).copy()
                   if (r.nb and n23) else np.zeros(0, np.uint8)),
            "list_off": np.ctypeslib.as_array(r.list_off, (r.n1 + 1,)).copy() if bins else None,
            "list_tri": ((np.ctypeslib.as_array(r.list_tri, (r.pairs,)).copy()
                          if r.pairs else np.zeros(0, np.int32)) if bins else None),
            "nb": r.nb, "pairs": r.pairs, "tests_l1": r.tests_l1, "tests_l2": r.tests_l2,
            "l1_inside": r.l1_inside, "l1_boundary": r.l1_boundary,
            "l2_inside": r.l2_inside, "l2_boundary": r.l2_boundary,
            "max_tris_per_voxel": r.max_tris_per_voxel,
            "t_stage": list(r.t_stage),
        }
        if normals:
            l1n = np.zeros((max(r.nb, 1), 3), np.float32)
            l2n = np.zeros((max(r.nb * n23, 1), 3), np.float32)
            lib().orc_normals(tris, ctypes.byref(g), ctypes.byref(r), int(threads), l1n.ravel(),
                              l2n.ctypes.data if n23 else None)
            out["l1n"] = l1n[:r.nb]
            out["l2n"] = l2n[:r.nb * n23]
    finally:
        lib().orc_result_free(ctypes.byref(r))
    return out


def save_raw(tris, g, obj_id, directory, normals=True, threads=0):
    """SaveVoxelization's files for the oracle's voxelization of tris on
    grid g: VoxelConfig.txt, Level1InOut.raw and, with Level2,
    Level1BoundaryPrefixSum.raw and Level2InOut.raw; with normals also
    Level1Normal.raw and (with Level2) Level2Normal.raw."""
    tris = _f32(tris).reshape(-1)
    T = len(tris) // 9
    r = _Result()
    L = lib()
    L.orc_voxelize(tris, T, ctypes.byref(g), ORC_FAST, int(threads), ctypes.byref(r))
    try:
        d = os.fsencode(directory)
        assert L.orc_save_raw(ctypes.byref(g), ctypes.byref(r), int(obj_id), d) == 0
        if normals:
            n23 = g.n2 ** 3 if g.n2 > 0 else 0
            l1n = np.zeros(max(r.nb, 1) * 3, np.float32)
            l2n = np.zeros(max(r.nb * n23, 1) * 3, np.float32)
            L.orc_normals(tris, ctypes.byref(g), ctypes.byref(r), int(threads), l1n,
                          l2n.ctypes.data if n23 else None)
            assert L.orc_save_raw_normals(ctypes.byref(g), ctypes.byref(r), l1n, l2n.ctypes.data, int(obj_id),
                                          d) == 0
    finally:
        L.orc_result_free(ctypes.byref(r))
