/*
 * gpv.h -- C ABI of libgpv, the gfx950 two-level hybrid voxelizer.
 *
 * Drop-in for GPView's voxelization path.  The reference binds three
 * extern "C" CUDA wrappers (includes/CUDAUtilities.h:87-89, defined at
 * cuda/CUDAClassifyTessellation.cu:507-540) from Object's host code
 * (src/Object.cpp:2071-2654, 3077-3430); this header exposes the same stages
 * with plain pointers and sizes, device buffers owned by the library, and a
 * status code per call instead of the reference's unconditional `return 1`.
 *
 * Results follow the reference's CPU composition bit for bit: Level1 ray
 * parity (src/Object.cpp:716-779), Level1 SAT binning (:2256-2342),
 * boundary compaction (:3270-3279), Level2 ray parity (:1111-1207), Level2
 * SAT (:2344-2424), .raw output (:2934-3075).
 *
 * Threading: calls on one context are serialised by the library (a context
 * may be shared by threads) and ordered on its stream; separate contexts run
 * concurrently on their own streams.
 */
#ifndef GPV_H
#define GPV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
enum gpv_status {
    GPV_OK = 0,
    GPV_ERR_INVALID = 1,      /* bad argument (null pointer, size, grid limits, or a
                                 vertex coordinate that is not finite or exceeds 2^40
                                 in magnitude: the input domain, checked on the
                                 device by the per-triangle set-up) */
    GPV_ERR_OOM = 2,          /* device or host allocation failed */
    GPV_ERR_CAPACITY = 3,     /* bin capacity regrow did not converge */
    GPV_ERR_HIP = 4,          /* HIP runtime error (message via gpv_last_error) */
    GPV_ERR_NO_DEVICE = 5,    /* no usable gfx950 device */
    GPV_ERR_IO = 6,           /* file could not be written */
    GPV_ERR_STATE = 7,        /* stage called out of order */
    GPV_ERR_TOO_LARGE = 8     /* input exceeds the fixed scratch or the caller's result
                                 buffers of gpv_voxelize_noalloc */
};

/* ---- parameters: GLParameters fields used by the path
 *      (includes/GLParameters.h, defaults src/GLParameters.cpp:63-67) ------- */
typedef struct gpv_params {
    int voxel_count;          /* GLParameters::voxelCount  (Level1 nominal divisions) */
    int voxel_count2;         /* GLParameters::voxelCount2 (Level2 divisions per axis) */
    int level2;               /* GLParameters::level2Voxels */
    int normals;              /* also compute the Level1/Level2 normals that
                                 PerformVoxelization stores (src/Object.cpp:3219-3253,
                                 2404-2421) and SaveVoxelization writes */
    int64_t initial_capacity; /* bin-pair capacity; <= 0 picks a default.  The
                                 analogue of PerformVoxelization's bufferSize
                                 (src/Object.cpp:3088-3089); overflow regrows and
                                 reruns Level1 binning (src/Object.cpp:3198-3214) */
    int cuda_level2;          /* 0: the CPU composition (the contract).  1: the
                                 Level2 arithmetic of the reference's CUDA kernels
                                 (cuda/CUDAClassifyTessellation.cu:403-504), for
                                 comparing with output of its GPU build: Level2
                                 SAT centers and ray origins in float,
                                 (2p+1)*ext2 + mid - ext1 (no FMA contraction), and
                                 Level2 parity over the whole xy column's triangle
                                 list instead of the bins at z >= k1.  Level1 and
                                 the normals keep the CPU semantics.  Not pinned
                                 against the CUDA build (it needs nvcc and an
                                 NVIDIA GPU). */
    int l1_stencil;           /* 0: Level1 inside/outside by +z ray parity at the
                                 voxel centers (the CPU composition, the
                                 contract).  1: by the reference's active GL path,
                                 ClassifyInOut2x (src/Object.cpp:1446-1601):
                                 per slice, the fragment count of the mesh clipped
                                 to z <= the slice's center plane on a 2nx x 2ny
                                 sample grid, inside when more than 2 of a
                                 voxel's 4 samples are odd.  Sample coverage
                                 follows a fixed rule in place of the GL
                                 driver's: vertices snapped to 1/256 sample,
                                 top-left rule, slice test on the triangle's
                                 plane in double; counts do not saturate.  Not
                                 pinned against a GL run (it needs a GL context).
                                 At most 2^19 divisions per axis and 2^29
                                 columns. */
} gpv_params;

/* ---- grid description (VoxelData, includes/Utilities.h:79-154) ----------- */
typedef struct gpv_grid_info {
    float bmin[3], bmax[3];   /* padded bbox, Object::bBoxMin/Max */
    float max_model_size;     /* Object::maxModelSize */
    int n[3];                 /* numDivX/Y/Z */
    float g[3];               /* gridSizeX/Y/Z */
    int n2;                   /* numDivX2 (= Y2 = Z2), 0 when Level2 is off */
    float g2[3];              /* gridSizeX2/Y2/Z2 */
    int64_t n1;               /* nx*ny*nz */
    int64_t nb;               /* boundary Level1 voxels (boundaryIndex.size()) */
    int64_t l1_inside, l1_boundary, l2_inside, l2_boundary;  /* src/Object.cpp:3353-3378 */
    int64_t tests_l1;         /* sum over triangles of index-AABB volume */
    int64_t tests_l2;         /* sum over boundary voxels of |bin| * n2^3 */
    int64_t pairs;            /* total Level1 bin entries */
    int32_t max_tris_per_voxel; /* "Max Triangles in Voxel", src/Object.cpp:2221 */
    int32_t regrows;          /* Level1 reruns caused by bin capacity */
    float ms_stage[6];        /* device time per stage: setup, L1 SAT, L1 parity,
                                 compaction+bins, L2 parity, L2 SAT */
    float ms_total;           /* device time, first to last stage */
    int32_t has_normals;      /* normals were computed (params.normals) */
    float ms_normals;         /* device time of the normals stage */
} gpv_grid_info;

typedef struct gpv_ctx gpv_ctx;     /* device, stream, memory pool */
typedef struct gpv_grid gpv_grid;   /* one voxelization, device resident */

/* ---- context -------------------------------------------------------------- */
/* Replaces InitializeCUDA (src/CUDAUtilities.cpp:51-55): binds `device`
 * (-1 = current), creates the context's stream and pool.  flags: reserved, 0. */
int gpv_ctx_create(int device, unsigned flags, gpv_ctx **out);
int gpv_ctx_destroy(gpv_ctx *ctx);
/* last error text of this thread (never NULL) */
const char *gpv_last_error(void);
const char *gpv_status_string(int status);

/* Device copy of a host buffer, from the context's pool: the cudaMalloc +
 * cudaMemcpy of the soup in Object::ClassifyTessellationCUDA
 * (src/Object.cpp:2096-2107), for callers that keep inputs resident and pass
 * them to gpv_voxelize with tris_on_device = 1. */
int gpv_device_upload(gpv_ctx *ctx, const void *host, int64_t bytes, void **dev);
int gpv_device_free(gpv_ctx *ctx, void *dev);
/* Page-locked host memory (hipHostMalloc) for soups that are uploaded
 * often: copies from it run at the link's full rate and overlap the
 * device's work.  No counterpart in the reference. */
int gpv_host_alloc(int64_t bytes, void **host);
int gpv_host_free(void *host);

/* The context's device memory pool.  No counterpart in the reference, which
 * cudaMallocs and frees around every call; serving loops read it to see that
 * steady state allocates nothing.  A request takes the smallest cached block
 * of one to two times its size class; before each new device allocation the
 * largest cached blocks are freed until cached bytes plus the new block fit
 * under cap_bytes (1/8 of the device's memory). */
typedef struct gpv_pool_stats {
    int64_t held_bytes;     /* device memory the pool owns (in use + cached) */
    int64_t cached_bytes;   /* free blocks kept for reuse */
    int64_t cap_bytes;
    int64_t allocations;    /* hipMalloc calls so far */
    int64_t reuses;         /* requests served from cached blocks */
    int64_t evictions;      /* cached blocks freed (cap, out of memory, destroy) */
    int64_t live_bytes;     /* block bytes in use now (size classes, results included) */
    int64_t peak_live_bytes;/* the most block bytes in use at once so far */
    int64_t limit_bytes;    /* gpv_ctx_set_mem_limit, 0 = none */
    int64_t requested_bytes;      /* bytes the blocks in use were requested for */
    int64_t peak_requested_bytes; /* the most requested bytes in use at once so far */
} gpv_pool_stats;
int gpv_ctx_pool_stats(gpv_ctx *ctx, gpv_pool_stats *out);
/* Caps the device bytes a context's pool may have in use at once (0 = no
 * cap): an allocation beyond it fails with GPV_ERR_OOM as if the device were
 * that size, so contexts can share a GPU with other tenants.  No counterpart
 * in the reference. */
int gpv_ctx_set_mem_limit(gpv_ctx *ctx, int64_t bytes);

/* Launch shapes of a context, picked by gpv_ctx_autotune or set by hand.  No
 * counterpart in the reference (its launch shapes are fixed,
 * cuda/CUDAHelper.h:23-38).  None of them changes a result. */
typedef struct gpv_tune {
    int expand_blocks_per_cu;  /* persistent blocks per CU of the Level1 SAT and ray expansion (16) */
    int l2_group_below;        /* the Level2 SAT shares a wave among consecutive boundary voxels
                                  when the bins average fewer entries than this per voxel (8);
                                  0 never, a large value always (except where a group could
                                  hold only one voxel: Level2 16 in every mesh of the run) */
    int sort_items_per_block;  /* bin entries per block in the counting sorts (2048) */
    int l2_overlap;            /* run independent stages on the context's second stream: the ray
                                  half of the set-up and the ray expansion beside the SAT half and
                                  the Level1 SAT, the parity fill and the column-entry sort beside
                                  the bins, the Level2 parity beside the Level2 SAT (1) */
    int l2_sat_waves;          /* waves per block of the Level2 SAT: 1, 2 or 4 (1) */
    int l2_parity_waves;       /* waves per block of the Level2 parity and its column pass:
                                  1, 2 or 4 (2) */
    int expand_stage_tris;     /* triangles per LDS batch of the Level1 SAT and ray expansion:
                                  128 or 256 (128) */
} gpv_tune;
int gpv_ctx_get_tune(gpv_ctx *ctx, gpv_tune *out);
int gpv_ctx_set_tune(gpv_ctx *ctx, const gpv_tune *t);

/* Measurement aid, no counterpart in the reference: runs `blocks` blocks of
 * 256 threads, each thread doing iters * 32 dependent-free FP32 adds
 * (v_add_f32, no packed forms), and returns the device time of one timed
 * launch and the number of VALU wave instructions its loop issues.  Their
 * ratio x 64 lanes is the FP32 vector issue rate the VALU roofline of the
 * path's kernels is measured against. */
int gpv_valu_probe(gpv_ctx *ctx, int blocks, int iters, float *ms, double *wave_insts);

/* ---- host-side geometry set-up ----------------------------------------- */
/* bbox over vertices + 0.1% pad + maxModelSize: Object::ReadObject
 * (src/Object.cpp:441-450, 572-583).  verts: nverts*3 floats (host). */
int gpv_mesh_bbox(const float *verts, int64_t nverts, float bmin[3], float bmax[3],
                  float *max_model_size);
/* Grid derivation of PerformVoxelization (src/Object.cpp:3094-3136) into
 * info's bbox/n/g/n2/g2 fields. */
int gpv_grid_derive(const float bmin[3], const float bmax[3], float max_model_size,
                    const gpv_params *params, gpv_grid_info *info);

/* The box centers and ray origins the path tests with, along one axis, as
 * the reference computes them (the float result of its double expressions;
 * these exact tables are uploaded for every grid):
 *   which 0  Level1 SAT centers  [n[axis]]     (p + 0.5)*ext*2 + min    src/Object.cpp:2329-2331
 *         1  Level1 ray origins  [n[axis]]     min + (i + 0.5)*g        src/Object.cpp:743-745
 *         2  Level2 SAT centers  [n[axis]*n2]  (p + 0.5)*g2 + mid - ext1, mid = (i + 0.5)*g + min
 *                                                                      src/Object.cpp:2369-2371, 2389-2391
 *         3  Level2 ray origins  [n[axis]*n2]  minPoint + (p + 0.5)*g2, minPoint = i*g + min in float
 *                                                                      src/Object.cpp:1134-1136, 1158-1160
 * Entry i*n2 + p of a Level2 table belongs to Level1 index i.  info comes
 * from gpv_grid_derive; cuda_level2 selects the CUDA kernels' float Level2
 * expression as in gpv_params. */
int gpv_grid_centers(const gpv_grid_info *info, int cuda_level2, int which, int axis, float *out);

/* ---- whole path ----------------------------------------------------------- */
/* Object::PerformVoxelization (src/Object.cpp:3077-3430) on a flat triangle
 * soup (Object::CreateFlatTriangleData layout, src/Object.cpp:3496-3527:
 * tris[t*9 + k*3 + c]).  tris_on_device != 0 means `tris` is a device
 * pointer that stays valid until the grid is released (no copy).  bmin/bmax/
 * max_model_size as produced by gpv_mesh_bbox.  The grid stays on the
 * device until gpv_grid_release. */
int gpv_voxelize(gpv_ctx *ctx, const float *tris, int64_t ntris, int tris_on_device,
                 const float bmin[3], const float bmax[3], float max_model_size,
                 const gpv_params *params, gpv_grid **out);

/* Device scratch of gpv_voxelize_noalloc, per GPU, held until the process ends. */
#define GPV_NOALLOC_SCRATCH_BYTES (256ull << 20)

/* gpv_voxelize for callers that cannot allocate device memory: code under
 * HIP graph capture, serving loops and harnesses that forbid hipMalloc.  No
 * counterpart in the reference (its wrappers cudaMalloc around every call,
 * src/Object.cpp:2096-2099, 2583-2586).  It makes no hipMalloc, hipFree or
 * other allocating call: every transient buffer comes from a fixed device
 * array of GPV_NOALLOC_SCRATCH_BYTES per GPU inside the library, and the
 * results go to the caller's device buffers, in the layout of
 * gpv_grid_device_views:
 *   l1_boundary_bits, l1_inside_bits   l1_words words each; the grid needs
 *                                      nz*ny*ceil(nx/64) (gpv_grid_derive)
 *   bidx                               nb_cap boundary voxel indices
 *   l2_boundary_bits, l2_inside_bits   nb_cap*ceil(n2^3/64) words each
 *                                      (may be NULL without Level2)
 * info receives the grid, counts and stage times (info->nb boundary voxels).
 * tris is a device soup (gpv_voxelize with tris_on_device = 1); normals are
 * not computed (params->normals must be 0).  A mesh whose working set does
 * not fit the scratch, or whose boundary count exceeds nb_cap, returns
 * GPV_ERR_TOO_LARGE.  Returns when the results are complete.  Calls on one
 * GPU share its scratch array: the library serialises them (overlapping
 * calls on one GPU would race on it), and each call clears what it needs
 * cleared, so the array's previous contents never matter. */
int gpv_voxelize_noalloc(gpv_ctx *ctx, const float *tris, int64_t ntris, const float bmin[3],
                         const float bmax[3], float max_model_size, const gpv_params *params,
                         uint64_t *l1_boundary_bits, uint64_t *l1_inside_bits, int64_t l1_words,
                         int32_t *bidx, uint64_t *l2_boundary_bits, uint64_t *l2_inside_bits,
                         int64_t nb_cap, gpv_grid_info *info);

/* ---- batches ---------------------------------------------------------------- */
/* One mesh of a batch: the arguments gpv_voxelize takes for it. */
typedef struct gpv_mesh_desc {
    const float *tris;        /* [ntris*9] soup, host or device (tris_on_device) */
    int64_t ntris;
    float bmin[3], bmax[3];   /* from gpv_mesh_bbox */
    float max_model_size;
    gpv_params params;        /* per-mesh Level1 / Level2 resolution */
} gpv_mesh_desc;

/* Voxelizes nmeshes meshes of any sizes and resolutions as one run: every
 * stage of the path is one launch over all meshes (the reference voxelizes
 * one Object per PerformVoxelization call, src/GPView.cpp:1378-1380), with
 * one size read-back for the whole batch.  out[m] receives mesh m's grid;
 * each grid holds exactly what gpv_voxelize of that mesh alone would hold
 * and works with every result call below (get_info, download, save_raw,
 * device_views).  The soups are copied into the batch, so the inputs may be
 * reused on return: host soups of more than 2 MiB in total go in two runs,
 * the second's upload under the first's kernels, and soups of up to 256 KiB
 * in page-locked memory (gpv_host_alloc) are read across the bus by one
 * gather kernel per run instead of a copy each.  Per-mesh ms_stage/ms_total report the batch's stage
 * times.  The batch's device memory returns to the context's pool when the
 * last of its grids is released.  One run holds at most 2^31-1 triangles,
 * 2^32-1 padded Level1 voxels (rows padded to 64), 2^27 xy columns, 2^27
 * boundary voxels and 2^32-1 bin entries: a batch beyond the first three, or
 * predicted to need more than half the free device memory, is cut into
 * consecutive runs before it starts, and a run that fails on the others or
 * on device memory is split in halves and rerun, down to single meshes.
 * Each mesh must fit one run by itself; on failure no grid is returned. */
int gpv_voxelize_batch(gpv_ctx *ctx, const gpv_mesh_desc *meshes, int nmeshes, int tris_on_device,
                       gpv_grid **out);

/* Launch shapes for a workload (gpv_tune): measures the device time of
 * gpv_voxelize_batch over `meshes` (min over `repeats` runs) for candidate
 * values of each field in turn, keeps a value only when it is at least 1%
 * faster, and leaves the best shapes set on the context.  best / best_ms may
 * be NULL. */
int gpv_ctx_autotune(gpv_ctx *ctx, const gpv_mesh_desc *meshes, int nmeshes, int tris_on_device, int repeats,
                     gpv_tune *best, float *best_ms);

/* ---- stage entry points (same order PerformVoxelization calls them) ------- */
/* Allocates the grid, uploads the soup and the center tables: the set-up of
 * PerformVoxelization (src/Object.cpp:3094-3193). */
int gpv_grid_create(gpv_ctx *ctx, const float *tris, int64_t ntris, int tris_on_device,
                    const float bmin[3], const float bmax[3], float max_model_size,
                    const gpv_params *params, gpv_grid **out);
/* Level1 SAT binning: CUDAClassifyTessellation (cuda/CUDAClassifyTessellation.cu:507)
 * as driven by Object::ClassifyTessellationCUDA (src/Object.cpp:2071-2253).
 * *needed receives 0, or the capacity that would have been required when the
 * bins overflowed (then call again: the reference's retry, src/Object.cpp:3198-3214). */
int gpv_classify_tessellation(gpv_ctx *ctx, gpv_grid *grid, int64_t *needed);
/* Level1 inside/outside by +z ray parity over all triangles:
 * Object::ClassifyInOutCPU (src/Object.cpp:716-779). */
int gpv_classify_inout(gpv_ctx *ctx, gpv_grid *grid);
/* Boundary compaction + exclusive prefix (src/Object.cpp:3270-3279) and the
 * per-boundary-voxel bins (the CSR of src/Object.cpp:2137-2219). */
int gpv_compact_boundary(gpv_ctx *ctx, gpv_grid *grid);
/* Level2 ray parity: CUDAClassifyInOutLevel2 (cuda/CUDAClassifyTessellation.cu:531)
 * with the CPU candidate rule of Object::ClassifyInOutLevel2CPU (src/Object.cpp:1111-1207). */
int gpv_classify_inout_level2(gpv_ctx *ctx, gpv_grid *grid);
/* Level2 SAT: CUDAClassifyTessellationLevel2 (cuda/CUDAClassifyTessellation.cu:518)
 * with the CPU centers of Object::ClassifyTessellationLevel2 (src/Object.cpp:2344-2424). */
int gpv_classify_tessellation_level2(gpv_ctx *ctx, gpv_grid *grid);
/* Counts (src/Object.cpp:3353-3378); waits for the grid's work. */
int gpv_grid_finish(gpv_ctx *ctx, gpv_grid *grid);

/* ---- results ---------------------------------------------------------------- */
int gpv_grid_get_info(const gpv_grid *grid, gpv_grid_info *info);
/* Copies results to host buffers (any may be NULL):
 *   l1_state[n1]        0 outside / 1 inside / 2 boundary, [z][y][x]
 *   prefix[n1]          exclusive boundary prefix sum (Level1BoundaryPrefixSum.raw)
 *   bidx[nb]            boundary Level1 linear indices, ascending
 *   l2_state[nb*n2^3]   0/1/2 per Level2 voxel, [b][r][q][p] */
int gpv_grid_download(gpv_ctx *ctx, const gpv_grid *grid, uint8_t *l1_state, int32_t *prefix,
                      int32_t *bidx, uint8_t *l2_state);
/* Normals of a grid made with params.normals (GPV_ERR_STATE otherwise), as
 * the reference stores them before quantising (any may be NULL):
 *   l1_normals[nb*3]       per boundary voxel in bidx order: the average of the
 *                          face normals of its Level1 bin, normalized
 *                          (src/Object.cpp:3219-3253; other voxels have none)
 *   l2_normals[nb*n2^3*3]  per Level2 voxel, [b][r][q][p]: the average of the
 *                          face normals of the bin triangles that hit it,
 *                          normalized, or 0 (src/Object.cpp:2404-2421, 2705-2720) */
int gpv_grid_download_normals(gpv_ctx *ctx, const gpv_grid *grid, float *l1_normals, float *l2_normals);
/* Raw device views (valid until release), for zero-copy consumers.
 * Level1 bitplanes: word (z*ny + y)*ceil(nx/64) + x/64, bit x%64 (rows padded
 * to whole 64-bit words).  Level2 bitplanes: ceil(n2^3/64) words per boundary
 * voxel b, bit (r*n2 + q)*n2 + p.  A voxel is boundary if its boundary bit is
 * set, otherwise inside if its inside bit is set. */
int gpv_grid_device_views(const gpv_grid *grid, const uint64_t **l1_boundary_bits,
                          const uint64_t **l1_inside_bits, const uint64_t **l2_boundary_bits,
                          const uint64_t **l2_inside_bits, const int32_t **bidx);
/* The same bitplanes copied to host memory (1 bit per voxel instead of the
 * state bytes of gpv_grid_download): l1 buffers hold nz*ny*ceil(nx/64) words,
 * l2 buffers nb*ceil(n2^3/64) words; any pointer may be NULL.  No
 * counterpart in the reference, which copies float grids. */
int gpv_grid_download_bits(gpv_ctx *ctx, const gpv_grid *grid, uint64_t *l1_boundary_bits,
                           uint64_t *l1_inside_bits, uint64_t *l2_boundary_bits, uint64_t *l2_inside_bits);
/* Object::SaveVoxelization (src/Object.cpp:2934-3075): writes
 * Obj<obj_id>VoxelConfig.txt, Obj<obj_id>Level1InOut.raw and, with Level2,
 * Obj<obj_id>Level1BoundaryPrefixSum.raw and Obj<obj_id>Level2InOut.raw
 * under dir (NULL = current directory); for a grid made with params.normals
 * also Obj<obj_id>Level1Normal.raw and (with Level2) Obj<obj_id>Level2Normal.raw,
 * uint8(n * 256/3 + 127) per component. */
int gpv_save_raw(gpv_ctx *ctx, const gpv_grid *grid, int obj_id, const char *dir);
int gpv_grid_release(gpv_ctx *ctx, gpv_grid *grid);

#ifdef __cplusplus
}
#endif
#endif /* GPV_H */
