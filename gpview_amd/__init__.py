"""gpview_amd -- Python view of libgpv, the gfx950 two-level hybrid voxelizer.

The product is the C ABI in include/gpv.h (libgpv.so, built in-tree by
gpview_amd/build.py).  This module binds it with ctypes and mirrors the
reference's Object-level flow (ReadObject -> CreateFlatTriangleData ->
PerformVoxelization -> SaveVoxelization, src/Object.cpp) for tests, the
benchmark and Python callers.  There is no CPU fallback: every call goes to
the HIP library, and a missing library is an ImportError-class failure.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# the in-tree build; GPV_LIBRARY names another build of the same sources
LIB_PATH = os.environ.get("GPV_LIBRARY") or os.path.join(HERE, "libgpv.so")

GPV_OK, GPV_ERR_INVALID, GPV_ERR_OOM, GPV_ERR_CAPACITY, GPV_ERR_HIP, GPV_ERR_NO_DEVICE, GPV_ERR_IO, \
    GPV_ERR_STATE, GPV_ERR_TOO_LARGE = range(9)
NOALLOC_SCRATCH_BYTES = 256 << 20   # GPV_NOALLOC_SCRATCH_BYTES


class GpvError(RuntimeError):
    def __init__(self, status, where, detail):
        super().__init__(f"{where}: status {status} ({detail})")
        self.status = status


class Params(ctypes.Structure):
    """gpv_params (include/gpv.h)."""
    _fields_ = [("voxel_count", ctypes.c_int), ("voxel_count2", ctypes.c_int), ("level2", ctypes.c_int),
                ("normals", ctypes.c_int), ("initial_capacity", ctypes.c_int64), ("cuda_level2", ctypes.c_int),
                ("l1_stencil", ctypes.c_int)]


class GridInfo(ctypes.Structure):
    """gpv_grid_info (include/gpv.h)."""
    _fields_ = [
        ("bmin", ctypes.c_float * 3), ("bmax", ctypes.c_float * 3), ("max_model_size", ctypes.c_float),
        ("n", ctypes.c_int * 3), ("g", ctypes.c_float * 3), ("n2", ctypes.c_int), ("g2", ctypes.c_float * 3),
        ("n1", ctypes.c_int64), ("nb", ctypes.c_int64),
        ("l1_inside", ctypes.c_int64), ("l1_boundary", ctypes.c_int64),
        ("l2_inside", ctypes.c_int64), ("l2_boundary", ctypes.c_int64),
        ("tests_l1", ctypes.c_int64), ("tests_l2", ctypes.c_int64), ("pairs", ctypes.c_int64),
        ("max_tris_per_voxel", ctypes.c_int32), ("regrows", ctypes.c_int32),
        ("ms_stage", ctypes.c_float * 6), ("ms_total", ctypes.c_float),
        ("has_normals", ctypes.c_int32), ("ms_normals", ctypes.c_float),
    ]

    def as_dict(self):
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            d[name] = list(v) if hasattr(v, "__len__") else v
        return d


class PoolStats(ctypes.Structure):
    """gpv_pool_stats (include/gpv.h)."""
    _fields_ = [("held_bytes", ctypes.c_int64), ("cached_bytes", ctypes.c_int64), ("cap_bytes", ctypes.c_int64),
                ("allocations", ctypes.c_int64), ("reuses", ctypes.c_int64), ("evictions", ctypes.c_int64),
                ("live_bytes", ctypes.c_int64), ("peak_live_bytes", ctypes.c_int64), ("limit_bytes", ctypes.c_int64),
                ("requested_bytes", ctypes.c_int64), ("peak_requested_bytes", ctypes.c_int64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Tune(ctypes.Structure):
    """gpv_tune (include/gpv.h)."""
    _fields_ = [("expand_blocks_per_cu", ctypes.c_int), ("l2_group_below", ctypes.c_int),
                ("sort_items_per_block", ctypes.c_int), ("l2_overlap", ctypes.c_int),
                ("l2_sat_waves", ctypes.c_int), ("l2_parity_waves", ctypes.c_int),
                ("expand_stage_tris", ctypes.c_int)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class MeshDesc(ctypes.Structure):
    """gpv_mesh_desc (include/gpv.h)."""
    _fields_ = [("tris", ctypes.c_void_p), ("ntris", ctypes.c_int64), ("bmin", ctypes.c_float * 3),
                ("bmax", ctypes.c_float * 3), ("max_model_size", ctypes.c_float), ("params", Params)]


def _desc_dtype():
    """numpy mirror of gpv_mesh_desc (offsets from the ctypes layout), so a
    batch's descriptors are filled a field at a time."""
    P = Params
    names = ["tris", "ntris", "bmin", "bmax", "max_model_size"] + ["p_" + f for f, _ in P._fields_]
    fmts = ["<u8", "<i8", ("<f4", 3), ("<f4", 3), "<f4"]
    offs = [getattr(MeshDesc, f).offset for f in ("tris", "ntris", "bmin", "bmax", "max_model_size")]
    base = MeshDesc.params.offset
    for f, t in P._fields_:
        fmts.append("<i8" if t is ctypes.c_int64 else "<i4")
        offs.append(base + getattr(P, f).offset)
    return np.dtype({"names": names, "formats": fmts, "offsets": offs, "itemsize": ctypes.sizeof(MeshDesc)})


_DESC_DTYPE = None


class ServerOpts(ctypes.Structure):
    """gpv_server_opts (include/gpv_serve.h)."""
    _fields_ = [("device", ctypes.c_int), ("lanes", ctypes.c_int), ("hbm_budget", ctypes.c_int64),
                ("max_batch_meshes", ctypes.c_int), ("max_batch_tris", ctypes.c_int64),
                ("lane_mem_limit", ctypes.c_int64)]


class RequestStats(ctypes.Structure):
    """gpv_request_stats (include/gpv_serve.h)."""
    _fields_ = [("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double),
                ("predicted_bytes", ctypes.c_int64), ("resident_bytes", ctypes.c_int64),
                ("batch_id", ctypes.c_int32), ("batch_size", ctypes.c_int32), ("lane", ctypes.c_int32),
                ("status", ctypes.c_int32), ("retries", ctypes.c_int32),
                ("batch_predicted_bytes", ctypes.c_int64), ("batch_peak_bytes", ctypes.c_int64)]


class Mesh(ctypes.Structure):
    """gpv_mesh (include/gpv_mesh.h)."""
    _fields_ = [("nverts", ctypes.c_int64), ("ntris", ctypes.c_int64),
                ("verts", ctypes.POINTER(ctypes.c_float)), ("faces", ctypes.POINTER(ctypes.c_int32)),
                ("bbox_from_faces", ctypes.c_int)]


_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_vp = ctypes.c_void_p
_lib = None

# every symbol include/gpv.h and include/gpv_mesh.h declare
EXPORTED = [
    "gpv_ctx_create", "gpv_ctx_destroy", "gpv_last_error", "gpv_status_string",
    "gpv_device_upload", "gpv_device_free", "gpv_host_alloc", "gpv_host_free", "gpv_ctx_pool_stats", "gpv_ctx_set_mem_limit", "gpv_ctx_get_tune",
    "gpv_ctx_set_tune",
    "gpv_ctx_autotune", "gpv_valu_probe",
    "gpv_mesh_bbox", "gpv_grid_derive", "gpv_grid_centers", "gpv_voxelize", "gpv_voxelize_noalloc",
    "gpv_voxelize_batch",
    "gpv_grid_create",
    "gpv_classify_tessellation", "gpv_classify_inout", "gpv_compact_boundary",
    "gpv_classify_inout_level2", "gpv_classify_tessellation_level2", "gpv_grid_finish",
    "gpv_grid_get_info", "gpv_grid_download", "gpv_grid_download_normals", "gpv_grid_device_views", "gpv_grid_download_bits",
    "gpv_save_raw",
    "gpv_grid_release",
    "gpv_mesh_free", "gpv_mesh_flatten", "gpv_mesh_bounds", "gpv_mesh_icosphere",
    "gpv_mesh_torus_knot", "gpv_mesh_torus", "gpv_mesh_drilled_box", "gpv_mesh_mixed",
    "gpv_mesh_stress", "gpv_mesh_read_obj", "gpv_mesh_read_off",
    "gpv_predict_footprint", "gpv_server_create", "gpv_server_destroy", "gpv_server_submit",
    "gpv_server_wait", "gpv_server_release", "gpv_server_counters", "gpv_server_readmitted",
    "gpv_server_pool_stats",
]


def lib():
    """Load the in-tree libgpv.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run `python -m gpview_amd.build` (hipcc, gfx950)")
    L = ctypes.CDLL(LIB_PATH)
    c_int, c_i64, c_float = ctypes.c_int, ctypes.c_int64, ctypes.c_float
    P = ctypes.POINTER
    L.gpv_last_error.restype = ctypes.c_char_p
    L.gpv_status_string.restype = ctypes.c_char_p
    L.gpv_status_string.argtypes = [c_int]
    L.gpv_ctx_create.argtypes = [c_int, ctypes.c_uint, P(_vp)]
    L.gpv_ctx_destroy.argtypes = [_vp]
    L.gpv_device_upload.argtypes = [_vp, _vp, c_i64, P(_vp)]
    L.gpv_device_free.argtypes = [_vp, _vp]
    L.gpv_host_alloc.argtypes = [c_i64, P(_vp)]
    L.gpv_host_free.argtypes = [_vp]
    L.gpv_ctx_pool_stats.argtypes = [_vp, P(PoolStats)]
    L.gpv_ctx_set_mem_limit.argtypes = [_vp, c_i64]
    L.gpv_valu_probe.argtypes = [_vp, c_int, c_int, P(c_float), P(ctypes.c_double)]
    L.gpv_server_readmitted.argtypes = [_vp, P(c_i64)]
    L.gpv_ctx_get_tune.argtypes = [_vp, P(Tune)]
    L.gpv_ctx_set_tune.argtypes = [_vp, P(Tune)]
    L.gpv_ctx_autotune.argtypes = [_vp, P(MeshDesc), c_int, c_int, c_int, P(Tune), P(c_float)]
    L.gpv_server_pool_stats.argtypes = [_vp, P(PoolStats)]
    L.gpv_mesh_bbox.argtypes = [_f32p, c_i64, _f32p, _f32p, P(c_float)]
    L.gpv_grid_derive.argtypes = [_f32p, _f32p, c_float, P(Params), P(GridInfo)]
    L.gpv_grid_centers.argtypes = [P(GridInfo), c_int, c_int, c_int, _f32p]
    for fn in ("gpv_voxelize", "gpv_grid_create"):
        getattr(L, fn).argtypes = [_vp, _vp, c_i64, c_int, _f32p, _f32p, c_float, P(Params), P(_vp)]
    L.gpv_voxelize_noalloc.argtypes = [_vp, _vp, c_i64, _f32p, _f32p, c_float, P(Params), _vp, _vp, c_i64, _vp,
                                       _vp, _vp, c_i64, P(GridInfo)]
    L.gpv_voxelize_batch.argtypes = [_vp, P(MeshDesc), c_int, c_int, P(_vp)]
    L.gpv_predict_footprint.argtypes = [P(MeshDesc), P(c_i64)]
    L.gpv_server_create.argtypes = [P(ServerOpts), P(_vp)]
    L.gpv_server_destroy.argtypes = [_vp]
    L.gpv_server_submit.argtypes = [_vp, P(MeshDesc), c_int, P(ctypes.c_uint64)]
    L.gpv_server_wait.argtypes = [_vp, ctypes.c_uint64, ctypes.c_double, P(_vp), P(_vp), P(RequestStats)]
    L.gpv_server_release.argtypes = [_vp, ctypes.c_uint64]
    L.gpv_server_counters.argtypes = [_vp, P(c_i64), P(c_i64), P(c_i64), P(c_i64)]
    L.gpv_classify_tessellation.argtypes = [_vp, _vp, P(c_i64)]
    for fn in ("gpv_classify_inout", "gpv_compact_boundary", "gpv_classify_inout_level2",
               "gpv_classify_tessellation_level2", "gpv_grid_finish", "gpv_grid_release"):
        getattr(L, fn).argtypes = [_vp, _vp]
    L.gpv_grid_get_info.argtypes = [_vp, P(GridInfo)]
    L.gpv_grid_download.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp]
    L.gpv_grid_download_normals.argtypes = [_vp, _vp, _vp, _vp]
    L.gpv_grid_device_views.argtypes = [_vp, P(_vp), P(_vp), P(_vp), P(_vp), P(_vp)]
    L.gpv_grid_download_bits.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp]
    L.gpv_save_raw.argtypes = [_vp, _vp, c_int, ctypes.c_char_p]
    L.gpv_mesh_free.argtypes = [P(Mesh)]
    L.gpv_mesh_flatten.argtypes = [P(Mesh), _f32p]
    L.gpv_mesh_bounds.argtypes = [P(Mesh), _f32p, _f32p, P(c_float)]
    L.gpv_mesh_icosphere.argtypes = [c_int, c_float, _f32p, c_float, ctypes.c_uint64, P(Mesh)]
    L.gpv_mesh_torus_knot.argtypes = [c_int, c_int, c_int, c_int, c_float, c_float, c_float, ctypes.c_uint64,
                                      P(Mesh)]
    L.gpv_mesh_torus.argtypes = [c_int, c_int, c_float, c_float, ctypes.c_uint64, P(Mesh)]
    L.gpv_mesh_drilled_box.argtypes = [c_i64, c_float, ctypes.c_uint64, P(Mesh)]
    L.gpv_mesh_mixed.argtypes = [c_int, c_i64, ctypes.c_uint64, P(Mesh)]
    L.gpv_mesh_stress.argtypes = [P(Mesh), c_int, c_int, c_float, c_float, c_float, ctypes.c_uint64]
    L.gpv_mesh_read_obj.argtypes = [ctypes.c_char_p, P(Mesh)]
    L.gpv_mesh_read_off.argtypes = [ctypes.c_char_p, P(Mesh)]
    _lib = L
    return L


def _check(st, where):
    if st != GPV_OK:
        L = lib()
        raise GpvError(st, where, f"{L.gpv_status_string(st).decode()}: {L.gpv_last_error().decode()}")


# ---------------------------------------------------------------------------
# meshes (host)
# ---------------------------------------------------------------------------
def _take_mesh(m):
    verts = np.ctypeslib.as_array(m.verts, (m.nverts * 3,)).reshape(-1, 3).copy() if m.nverts else \
        np.zeros((0, 3), np.float32)
    faces = np.ctypeslib.as_array(m.faces, (m.ntris * 3,)).reshape(-1, 3).copy() if m.ntris else \
        np.zeros((0, 3), np.int32)
    bff = bool(m.bbox_from_faces)
    lib().gpv_mesh_free(ctypes.byref(m))
    return verts, faces, bff


def _gen(fn, *args):
    m = Mesh()
    _check(getattr(lib(), fn)(*args, ctypes.byref(m)), fn)
    v, f, _ = _take_mesh(m)
    return v, f


def icosphere(level=4, radius=0.4, center=(0.5, 0.5, 0.5), jitter=0.02, seed=1):
    return _gen("gpv_mesh_icosphere", level, radius, np.asarray(center, np.float32), jitter, seed)


def torus_knot(p=3, q=7, segments=2048, ring=256, tube=0.3, jitter=0.01, extent=0.9, seed=2):
    return _gen("gpv_mesh_torus_knot", p, q, segments, ring, tube, jitter, extent, seed)


def torus(nu, nv, ratio, extent=0.9, seed=3):
    return _gen("gpv_mesh_torus", nu, nv, ratio, extent, seed)


def drilled_box(target_tris, extent=0.9, seed=4):
    return _gen("gpv_mesh_drilled_box", int(target_tris), extent, seed)


def mixed(kind, target_tris, seed):
    return _gen("gpv_mesh_mixed", int(kind), int(target_tris), seed)


def stress(verts, faces, voxel_count, voxel_count2, p_snap=0.3, f_sliver=0.1, f_coplanar=0.0, seed=5):
    """Lattice snapping, face-coplanar triangles and slivers applied to a
    copy of (verts, faces)."""
    L = lib()
    v = np.ascontiguousarray(verts, np.float32)
    f = np.ascontiguousarray(faces, np.int32)
    m = Mesh()
    # the library owns (and may reallocate) the arrays: hand it malloc'd copies
    libc = ctypes.CDLL(None)
    libc.malloc.restype = ctypes.c_void_p
    libc.malloc.argtypes = [ctypes.c_size_t]
    pv = libc.malloc(max(v.nbytes, 4))
    pf = libc.malloc(max(f.nbytes, 4))
    ctypes.memmove(pv, v.ctypes.data, v.nbytes)
    ctypes.memmove(pf, f.ctypes.data, f.nbytes)
    m.nverts, m.ntris = len(v), len(f)
    m.verts = ctypes.cast(pv, ctypes.POINTER(ctypes.c_float))
    m.faces = ctypes.cast(pf, ctypes.POINTER(ctypes.c_int32))
    m.bbox_from_faces = 0
    st = L.gpv_mesh_stress(ctypes.byref(m), voxel_count, voxel_count2, p_snap, f_sliver, f_coplanar, seed)
    if st != GPV_OK:
        L.gpv_mesh_free(ctypes.byref(m))
        _check(st, "gpv_mesh_stress")
    v2, f2, _ = _take_mesh(m)
    return v2, f2


def read_obj(path):
    m = Mesh()
    _check(lib().gpv_mesh_read_obj(os.fsencode(path), ctypes.byref(m)), "gpv_mesh_read_obj")
    return _take_mesh(m)


def read_off(path):
    m = Mesh()
    _check(lib().gpv_mesh_read_off(os.fsencode(path), ctypes.byref(m)), "gpv_mesh_read_off")
    return _take_mesh(m)


def flatten(verts, faces):
    """Object::CreateFlatTriangleData (src/Object.cpp:3496-3527)."""
    return np.ascontiguousarray(np.asarray(verts, np.float32)[np.asarray(faces)].reshape(-1, 9))


def mesh_bbox(verts):
    """bbox + 0.1% pad + maxModelSize (src/Object.cpp:441-450, 572-583)."""
    v = np.ascontiguousarray(verts, np.float32).reshape(-1)
    bmin = np.zeros(3, np.float32)
    bmax = np.zeros(3, np.float32)
    mms = ctypes.c_float()
    _check(lib().gpv_mesh_bbox(v, len(v) // 3, bmin, bmax, ctypes.byref(mms)), "gpv_mesh_bbox")
    return bmin, bmax, np.float32(mms.value)


def grid_derive(bmin, bmax, mms, voxel_count, voxel_count2, level2=True):
    info = GridInfo()
    p = Params(voxel_count=voxel_count, voxel_count2=voxel_count2, level2=1 if level2 else 0)
    _check(lib().gpv_grid_derive(np.asarray(bmin, np.float32), np.asarray(bmax, np.float32), float(mms),
                                 ctypes.byref(p), ctypes.byref(info)), "gpv_grid_derive")
    return info


def grid_centers(info, which, axis, cuda_level2=False):
    """gpv_grid_centers: table `which` (0 Level1 SAT centers, 1 Level1 ray
    origins, 2 Level2 SAT centers, 3 Level2 ray origins) along `axis` for a
    GridInfo from grid_derive."""
    n = info.n[axis] * (info.n2 if which >= 2 else 1) if 0 <= axis < 3 else 0
    out = np.zeros(max(n, 1), np.float32)
    _check(lib().gpv_grid_centers(ctypes.byref(info), 1 if cuda_level2 else 0, int(which), int(axis), out),
           "gpv_grid_centers")
    return out[:n]


# ---------------------------------------------------------------------------
# device path
# ---------------------------------------------------------------------------
class Context:
    """gpv_ctx: device, stream, memory pool."""

    def __init__(self, device=-1):
        self.h = _vp()
        _check(lib().gpv_ctx_create(device, 0, ctypes.byref(self.h)), "gpv_ctx_create")

    def close(self):
        if self.h:
            lib().gpv_ctx_destroy(self.h)
            self.h = _vp()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload(self, array):
        """Device copy of a host array (context pool); returns the pointer."""
        a = np.ascontiguousarray(array)
        d = _vp()
        _check(lib().gpv_device_upload(self.h, a.ctypes.data, a.nbytes, ctypes.byref(d)), "gpv_device_upload")
        return d.value

    def free(self, ptr):
        _check(lib().gpv_device_free(self.h, ptr), "gpv_device_free")

    def pool_stats(self):
        """gpv_ctx_pool_stats as a dict."""
        s = PoolStats()
        _check(lib().gpv_ctx_pool_stats(self.h, ctypes.byref(s)), "gpv_ctx_pool_stats")
        return s.as_dict()

    def set_mem_limit(self, nbytes):
        """gpv_ctx_set_mem_limit: cap on the pool's bytes in use (0: none)."""
        _check(lib().gpv_ctx_set_mem_limit(self.h, int(nbytes)), "gpv_ctx_set_mem_limit")

    def valu_probe(self, blocks, iters):
        """gpv_valu_probe: (device ms, VALU wave instructions) of the probe."""
        ms, wi = ctypes.c_float(), ctypes.c_double()
        _check(lib().gpv_valu_probe(self.h, int(blocks), int(iters), ctypes.byref(ms), ctypes.byref(wi)),
               "gpv_valu_probe")
        return ms.value, wi.value

    def get_tune(self):
        """gpv_ctx_get_tune as a dict."""
        t = Tune()
        _check(lib().gpv_ctx_get_tune(self.h, ctypes.byref(t)), "gpv_ctx_get_tune")
        return t.as_dict()

    def set_tune(self, **fields):
        """gpv_ctx_set_tune: the current shapes with `fields` replaced."""
        t = Tune(**dict(self.get_tune(), **fields))
        _check(lib().gpv_ctx_set_tune(self.h, ctypes.byref(t)), "gpv_ctx_set_tune")

    def autotune(self, meshes, device_ptrs=None, repeats=3):
        """gpv_ctx_autotune over a batch of mesh dicts; returns (shapes, device ms)."""
        n = len(meshes)
        descs = (MeshDesc * n)()
        keep = []
        for i, m in enumerate(meshes):
            d, soup = mesh_desc(m, device_ptrs[i] if device_ptrs is not None else None)
            descs[i] = d
            keep.append(soup)
        t = Tune()
        ms = ctypes.c_float()
        _check(lib().gpv_ctx_autotune(self.h, descs, n, 1 if device_ptrs is not None else 0, int(repeats),
                                      ctypes.byref(t), ctypes.byref(ms)), "gpv_ctx_autotune")
        return t.as_dict(), ms.value

    def voxelize(self, soup, bmin, bmax, mms, voxel_count, voxel_count2=4, level2=True, initial_capacity=0,
                 device_ptr=None, normals=False, cuda_level2=False, l1_stencil=False):
        """Object::PerformVoxelization on a flat soup; returns a Grid.
        cuda_level2: the Level2 arithmetic of the reference's CUDA kernels;
        l1_stencil: Level1 inside/outside by the GL-stencil-equivalent fill
        (ClassifyInOut2x) instead of ray parity."""
        soup = np.ascontiguousarray(soup, np.float32).reshape(-1, 9)
        p = Params(voxel_count=int(voxel_count), voxel_count2=int(voxel_count2), level2=1 if level2 else 0,
                   normals=1 if normals else 0, initial_capacity=int(initial_capacity),
                   cuda_level2=1 if cuda_level2 else 0, l1_stencil=1 if l1_stencil else 0)
        g = _vp()
        ptr = device_ptr if device_ptr is not None else soup.ctypes.data
        _check(lib().gpv_voxelize(self.h, ptr, len(soup), 1 if device_ptr is not None else 0,
                                  np.asarray(bmin, np.float32), np.asarray(bmax, np.float32), float(mms),
                                  ctypes.byref(p), ctypes.byref(g)), "gpv_voxelize")
        return Grid(self, g)

    def voxelize_noalloc(self, device_ptr, ntris, bmin, bmax, mms, voxel_count, voxel_count2, level2, l1b, l1i,
                         l1_words, bidx, l2b, l2i, nb_cap, l1_stencil=False):
        """gpv_voxelize_noalloc: device soup in, results into the caller's
        device buffers (pointers); returns the grid info dict."""
        p = Params(voxel_count=int(voxel_count), voxel_count2=int(voxel_count2), level2=1 if level2 else 0,
                   l1_stencil=1 if l1_stencil else 0)
        info = GridInfo()
        _check(lib().gpv_voxelize_noalloc(self.h, device_ptr, int(ntris), np.asarray(bmin, np.float32),
                                          np.asarray(bmax, np.float32), float(mms), ctypes.byref(p), l1b, l1i,
                                          int(l1_words), bidx, l2b, l2i, int(nb_cap), ctypes.byref(info)),
               "gpv_voxelize_noalloc")
        return info.as_dict()

    def voxelize_batch(self, meshes, device_ptrs=None):
        """gpv_voxelize_batch.  meshes: sequence of dicts with soup, bmin, bmax,
        mms, voxel_count and optionally voxel_count2, level2,
        initial_capacity.  device_ptrs: per-mesh device soups (then the
        soups are only used for their triangle counts).  Returns one Grid per
        mesh."""
        n = len(meshes)
        descs, keep = mesh_descs(meshes, device_ptrs)
        out = (_vp * n)()
        _check(lib().gpv_voxelize_batch(self.h, descs.ctypes.data_as(ctypes.POINTER(MeshDesc)), n,
                                        1 if device_ptrs is not None else 0, out), "gpv_voxelize_batch")
        return [Grid(self, _vp(h)) for h in out]

    def create(self, soup, bmin, bmax, mms, voxel_count, voxel_count2=4, level2=True, initial_capacity=0,
               normals=False, l1_stencil=False):
        """gpv_grid_create: the staged form (call the stage methods on the Grid)."""
        soup = np.ascontiguousarray(soup, np.float32).reshape(-1, 9)
        p = Params(voxel_count=int(voxel_count), voxel_count2=int(voxel_count2), level2=1 if level2 else 0,
                   normals=1 if normals else 0, initial_capacity=int(initial_capacity),
                   l1_stencil=1 if l1_stencil else 0)
        g = _vp()
        _check(lib().gpv_grid_create(self.h, soup.ctypes.data, len(soup), 0, np.asarray(bmin, np.float32),
                                     np.asarray(bmax, np.float32), float(mms), ctypes.byref(p),
                                     ctypes.byref(g)), "gpv_grid_create")
        return Grid(self, g)


class PinnedArray:
    """A numpy array in page-locked host memory (gpv_host_alloc); .array is
    the view, close() returns the memory."""

    def __init__(self, shape, dtype=np.float32):
        self.dtype = np.dtype(dtype)
        n = int(np.prod(shape)) * self.dtype.itemsize
        self.h = _vp()
        _check(lib().gpv_host_alloc(n, ctypes.byref(self.h)), "gpv_host_alloc")
        buf = (ctypes.c_char * max(n, 1)).from_address(self.h.value)
        self.array = np.frombuffer(buf, self.dtype, count=int(np.prod(shape))).reshape(shape)

    def close(self):
        if self.h:
            self.array = None
            lib().gpv_host_free(self.h)
            self.h = _vp()


def pinned_copy(a):
    """PinnedArray holding a copy of numpy array a."""
    p = PinnedArray(a.shape, a.dtype)
    p.array[...] = a
    return p


def _data_ptr(a):
    """Address of a contiguous array's data (through the buffer protocol,
    which is cheaper than __array_interface__; that one for read-only and
    empty arrays)."""
    try:
        return ctypes.addressof(ctypes.c_char.from_buffer(a))
    except (TypeError, ValueError, BufferError):
        return a.__array_interface__["data"][0]


def mesh_descs(meshes, device_ptrs=None):
    """gpv_mesh_desc array of a sequence of mesh dicts, filled field by field
    (numpy mirror of the struct); returns (array, soups) -- keep the soups
    alive while the array is in use."""
    global _DESC_DTYPE
    if _DESC_DTYPE is None:
        _DESC_DTYPE = _desc_dtype()
    n = len(meshes)
    a = np.zeros(n, _DESC_DTYPE)
    soups = []
    for m in meshes:
        s = m["soup"]
        if not (type(s) is np.ndarray and s.dtype == np.float32 and s.flags.c_contiguous):
            s = np.ascontiguousarray(s, np.float32)
        if s.size % 9:
            raise ValueError(f"a soup of {s.size} floats is not a whole number of triangles")
        soups.append(s)
    if device_ptrs is not None:
        a["tris"] = [int(p) for p in device_ptrs]
    else:
        a["tris"] = [_data_ptr(s) for s in soups]
    a["ntris"] = [s.size // 9 for s in soups]
    a["bmin"] = [m["bmin"] for m in meshes]
    a["bmax"] = [m["bmax"] for m in meshes]
    a["max_model_size"] = [m["mms"] for m in meshes]
    a["p_voxel_count"] = [m["voxel_count"] for m in meshes]
    a["p_voxel_count2"] = [m.get("voxel_count2", 4) for m in meshes]
    a["p_level2"] = [1 if m.get("level2", True) else 0 for m in meshes]
    a["p_normals"] = [1 if m.get("normals", False) else 0 for m in meshes]
    a["p_initial_capacity"] = [m.get("initial_capacity", 0) for m in meshes]
    a["p_cuda_level2"] = [1 if m.get("cuda_level2", False) else 0 for m in meshes]
    a["p_l1_stencil"] = [1 if m.get("l1_stencil", False) else 0 for m in meshes]
    return a, soups


def mesh_desc(m, device_ptr=None):
    """gpv_mesh_desc of a mesh dict (soup, bmin, bmax, mms, voxel_count, ...);
    returns (desc, soup) -- keep the soup alive while the desc is in use."""
    soup = np.ascontiguousarray(m["soup"], np.float32).reshape(-1, 9)
    d = MeshDesc()
    d.tris = device_ptr if device_ptr is not None else soup.ctypes.data
    d.ntris = len(soup)
    d.bmin[:] = [float(x) for x in m["bmin"]]
    d.bmax[:] = [float(x) for x in m["bmax"]]
    d.max_model_size = float(m["mms"])
    d.params = Params(voxel_count=int(m["voxel_count"]), voxel_count2=int(m.get("voxel_count2", 4)),
                      level2=1 if m.get("level2", True) else 0, normals=1 if m.get("normals", False) else 0,
                      initial_capacity=int(m.get("initial_capacity", 0)),
                      cuda_level2=1 if m.get("cuda_level2", False) else 0,
                      l1_stencil=1 if m.get("l1_stencil", False) else 0)
    return d, soup


def predict_footprint(m):
    d, _ = mesh_desc(m)
    b = ctypes.c_int64()
    _check(lib().gpv_predict_footprint(ctypes.byref(d), ctypes.byref(b)), "gpv_predict_footprint")
    return b.value


class _CtxRef:
    """A context owned elsewhere (a server lane): usable, never destroyed."""

    def __init__(self, h):
        self.h = h


class Server:
    """gpv_server: continuous-batching request loop (include/gpv_serve.h)."""

    def __init__(self, lanes=2, hbm_budget=0, max_batch_meshes=0, max_batch_tris=0, device=-1, lane_mem_limit=0):
        self.h = _vp()
        o = ServerOpts(device, lanes, int(hbm_budget), max_batch_meshes, int(max_batch_tris), int(lane_mem_limit))
        _check(lib().gpv_server_create(ctypes.byref(o), ctypes.byref(self.h)), "gpv_server_create")
        self._keep = {}

    def close(self):
        if self.h:
            lib().gpv_server_destroy(self.h)
            self.h = _vp()
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, m, device_ptr=None):
        d, soup = mesh_desc(m, device_ptr)
        rid = ctypes.c_uint64()
        _check(lib().gpv_server_submit(self.h, ctypes.byref(d), 1 if device_ptr is not None else 0,
                                       ctypes.byref(rid)), "gpv_server_submit")
        self._keep[rid.value] = soup
        return rid.value

    def wait(self, rid, timeout=-1.0):
        """(Grid view, stats dict) of a finished request."""
        g, c, st = _vp(), _vp(), RequestStats()
        _check(lib().gpv_server_wait(self.h, rid, float(timeout), ctypes.byref(g), ctypes.byref(c),
                                     ctypes.byref(st)), "gpv_server_wait")
        stats = {k: getattr(st, k) for k, _ in RequestStats._fields_}
        return Grid(_CtxRef(c), g, owned=False), stats

    def release(self, rid):
        _check(lib().gpv_server_release(self.h, rid), "gpv_server_release")
        self._keep.pop(rid, None)

    def counters(self):
        a, b, c, d = (ctypes.c_int64() for _ in range(4))
        _check(lib().gpv_server_counters(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                         ctypes.byref(d)), "gpv_server_counters")
        r = ctypes.c_int64()
        _check(lib().gpv_server_readmitted(self.h, ctypes.byref(r)), "gpv_server_readmitted")
        return {"completed": a.value, "batches": b.value, "charged": c.value, "peak_charged": d.value,
                "readmitted": r.value}

    def pool_stats(self):
        """gpv_server_pool_stats: the lanes' pools, summed."""
        s = PoolStats()
        _check(lib().gpv_server_pool_stats(self.h, ctypes.byref(s)), "gpv_server_pool_stats")
        return s.as_dict()


class Grid:
    """A device-resident voxelization (gpv_grid)."""

    def __init__(self, ctx, h, owned=True):
        self.ctx = ctx
        self.h = h
        self.owned = owned

    # staged entry points, in PerformVoxelization's order
    def classify_tessellation(self):
        need = ctypes.c_int64()
        _check(lib().gpv_classify_tessellation(self.ctx.h, self.h, ctypes.byref(need)),
               "gpv_classify_tessellation")
        return need.value

    def classify_inout(self):
        _check(lib().gpv_classify_inout(self.ctx.h, self.h), "gpv_classify_inout")

    def compact_boundary(self):
        _check(lib().gpv_compact_boundary(self.ctx.h, self.h), "gpv_compact_boundary")

    def classify_inout_level2(self):
        _check(lib().gpv_classify_inout_level2(self.ctx.h, self.h), "gpv_classify_inout_level2")

    def classify_tessellation_level2(self):
        _check(lib().gpv_classify_tessellation_level2(self.ctx.h, self.h), "gpv_classify_tessellation_level2")

    def finish(self):
        _check(lib().gpv_grid_finish(self.ctx.h, self.h), "gpv_grid_finish")

    def info(self):
        i = GridInfo()
        _check(lib().gpv_grid_get_info(self.h, ctypes.byref(i)), "gpv_grid_get_info")
        return i.as_dict()

    def download(self):
        i = self.info()
        n1, nb, n2 = i["n1"], i["nb"], i["n2"]
        l1 = np.zeros(n1, np.uint8)
        prefix = np.zeros(n1, np.int32)
        bidx = np.zeros(max(nb, 1), np.int32)
        l2 = np.zeros(max(nb * n2 ** 3, 1), np.uint8)
        _check(lib().gpv_grid_download(self.ctx.h, self.h, l1.ctypes.data, prefix.ctypes.data,
                                       bidx.ctypes.data if nb else None,
                                       l2.ctypes.data if (nb and n2) else None), "gpv_grid_download")
        return {"l1": l1, "prefix": prefix, "bidx": bidx[:nb], "l2": l2[:nb * n2 ** 3]}

    def download_bits(self):
        """gpv_grid_download_bits: (l1 boundary, l1 inside, l2 boundary, l2
        inside) uint64 bitplanes (Level2 ones empty without Level2)."""
        i = self.info()
        nx, ny, nz = i["n"]
        w1 = nz * ny * ((nx + 63) // 64)
        n2 = i["n2"]
        w2 = i["nb"] * ((n2 ** 3 + 63) // 64) if n2 else 0
        out = [np.zeros(max(w1, 1), np.uint64), np.zeros(max(w1, 1), np.uint64),
               np.zeros(max(w2, 1), np.uint64), np.zeros(max(w2, 1), np.uint64)]
        _check(lib().gpv_grid_download_bits(self.ctx.h, self.h, *(a.ctypes.data for a in out)),
               "gpv_grid_download_bits")
        return out[0][:w1], out[1][:w1], out[2][:w2], out[3][:w2]

    def download_normals(self):
        """(l1n [nb,3], l2n [nb*n2^3,3]) float32 of a grid made with normals."""
        i = self.info()
        nb, n2 = i["nb"], i["n2"]
        l1n = np.zeros((max(nb, 1), 3), np.float32)
        l2n = np.zeros((max(nb * n2 ** 3, 1), 3), np.float32)
        _check(lib().gpv_grid_download_normals(self.ctx.h, self.h, l1n.ctypes.data,
                                               l2n.ctypes.data if n2 else None), "gpv_grid_download_normals")
        return l1n[:nb], l2n[:nb * n2 ** 3]

    def save_raw(self, obj_id, directory):
        _check(lib().gpv_save_raw(self.ctx.h, self.h, int(obj_id), os.fsencode(directory)), "gpv_save_raw")

    def release(self):
        if self.h and self.owned:
            lib().gpv_grid_release(self.ctx.h, self.h)
        self.h = _vp()
