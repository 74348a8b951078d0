/*
 * gpv_oracle.h -- CPU restatement of GPView's two-level hybrid voxelizer.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against; the product (libgpv) never links, loads or calls it.
 *
 * Every function restates the reference's CPU composition (the path
 * PerformVoxelization takes with the CUDA classifier compiled out and ray
 * parity in place of the GL stencil fill):
 *
 *   bbox + offset         src/Object.cpp:441-450, 572-583 (OFF: 257-296)
 *   grid derivation       src/Object.cpp:3094-3136, includes/Utilities.h:313-316
 *   Level1 ray parity     src/Object.cpp:716-779  + src/TriRayIntersection.cpp:78-131
 *   Level1 stencil fill   src/Object.cpp:1446-1601 (opt-in, orc_grid.l1_stencil)
 *   Level1 SAT + bins     src/Object.cpp:2256-2342, 2681-2748 + src/TriBoxIntersection.cpp:64-210
 *   compaction / prefix   src/Object.cpp:3258-3284
 *   Level2 ray parity     src/Object.cpp:1111-1207
 *   Level2 SAT            src/Object.cpp:2344-2424
 *   counts                src/Object.cpp:3353-3378
 *   normals               src/Object.cpp:3219-3253 (Level1), 2372-2421 + 2705-2720 (Level2)
 *   .raw writer           src/Object.cpp:2934-3075
 *
 * Built with gcc -O2 -ffp-contract=off (x86-64 SSE2 float semantics, no
 * contraction), matching the reference's Release flags.
 */
#ifndef GPV_ORACLE_H
#define GPV_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_grid {
    float bmin[3];        /* padded object bbox (Object::bBoxMin) */
    float bmax[3];        /* padded object bbox (Object::bBoxMax) */
    float max_model_size; /* Object::maxModelSize */
    int n[3];             /* Level1 divisions numDivX/Y/Z */
    float g[3];           /* Level1 grid size gridSizeX/Y/Z */
    float ext1[3];        /* Level1 half extents (float)(g/2.0) */
    int n2;               /* Level2 divisions per axis, 0 = Level2 off */
    float g2[3];          /* Level2 grid size gridSizeX2/Y2/Z2 */
    float ext2[3];        /* Level2 half extents (float)(g2/2.0) */
    int cuda_l2;          /* 1: Level2 in the arithmetic of the reference's CUDA
                             kernels (cuda/CUDAClassifyTessellation.cu:403-504):
                             float centers/origins (2p+1)*ext2 + mid - ext1 and
                             parity over the whole xy column's triangle list */
    int l1_stencil;       /* 1: Level1 inside/outside by the GL-stencil-equivalent
                             fill (orc_stencil_counts) instead of ray parity */
} orc_grid;

typedef struct orc_result {
    int64_t n1;           /* nx*ny*nz */
    uint8_t *l1;          /* [n1] state 0 outside, 1 inside, 2 boundary; [z][y][x] */
    int32_t *prefix;      /* [n1] exclusive boundary prefix sum */
    int64_t nb;           /* number of boundary Level1 voxels */
    int32_t *bidx;        /* [nb] boundary linear indices, ascending */
    int n2;
    uint8_t *l2;          /* [nb * n2^3] state per Level2 voxel, [b][r][q][p] */
    int64_t *list_off;    /* [n1+1] CSR offsets of the Level1 SAT bins */
    int32_t *list_tri;    /* [pairs] triangle ids per bin, ascending */
    int64_t pairs;
    int64_t tests_l1;     /* sum over triangles of clipped index-AABB volume */
    int64_t tests_l2;     /* sum over boundary voxels of |bin| * n2^3 */
    int64_t l1_inside, l1_boundary, l2_inside, l2_boundary;
    int max_tris_per_voxel;
    double t_stage[5];    /* seconds: L1 parity, L1 SAT, compaction, L2 parity, L2 SAT */
} orc_result;

enum { ORC_EXACT = 0, ORC_FAST = 1, ORC_PLAN = 2 };

/* Reference geometric predicates, restated. */
int orc_tri_box_overlap(const float center[3], const float half[3], const float tv[9]);
int orc_tri_ray(const float v1[3], const float v2[3], const float v3[3],
                const float o[3], const float d[3], float *t_out);

/* Batched forms for tests: n independent cases. */
void orc_tri_box_batch(int64_t n, const float *c, const float *h, const float *tv, uint8_t *res);
/* The SAT screened by its per-axis and per-plane predicate groups before the
 * full test (the form ORC_FAST sweeps with); must equal orc_tri_box_overlap. */
int orc_tri_box_factored(const float center[3], const float half[3], const float tv[9]);
void orc_tri_box_factored_batch(int64_t n, const float *c, const float *h, const float *tv, uint8_t *res);
void orc_tri_ray_batch(int64_t n, const float *tv, const float *o, const float *d,
                       uint8_t *hit, float *t);
/* Column-factorised +z ray test (det/u/v on O.x,O.y; t on O.z), batched. */
void orc_tri_zray_split_batch(int64_t n, const float *tv, const float *o, uint8_t *hit);

/* bbox over a vertex array (stride 3 floats) with the 0.1% pad. */
void orc_bbox(const float *verts, int64_t nverts, float bmin[3], float bmax[3],
              float *max_model_size);
/* Grid derivation; voxel_count2 <= 0 switches Level2 off. */
void orc_grid_setup(const float bmin[3], const float bmax[3], float max_model_size,
                    int voxel_count, int voxel_count2, orc_grid *g);

/* Per-axis center tables. */
void orc_l1_centers(const orc_grid *g, int axis, float *out);      /* [n]  SAT centers   */
void orc_l1_ray_origins(const orc_grid *g, int axis, float *out);  /* [n]  ray origins   */
void orc_l2_centers(const orc_grid *g, int axis, float *out);      /* [n*n2] SAT centers */
void orc_l2_ray_origins(const orc_grid *g, int axis, float *out);  /* [n*n2] ray origins */
float orc_l2_cuda_center(const orc_grid *g, int axis, float mid, int p);

/* Integer index range of one vertex coordinate (src/Object.cpp:2272-2300). */
int orc_vertex_index(float v, float mn, float mx, int n);

/* Conservative column range for the Level1 parity fill: every (ix, iy)
 * column whose ray the Moller-Trumbore test can accept lies inside
 * [lo[0],hi[0]] x [lo[1],hi[1]].  Returns 0 when the triangle is rejected by
 * the det test for every ray (then no column can count it). */
int orc_ray_column_bounds(const float tv[9], const float bmin[3], const float bmax[3],
                          double safety, double lo[2], double hi[2]);

/* The GL stencil fill of ClassifyInOut2x (src/Object.cpp:1446-1601) with a
 * fixed rasterization rule.  The reference draws the mesh orthographically
 * over the bbox into a 2nx x 2ny stencil, clipped to z <= zs for slice
 * s (zs = (s + 0.5) g + min), counting every fragment, and calls a voxel
 * inside when more than 2 of its 4 samples hold an odd count.  Which samples
 * a triangle covers is up to the GL driver; this rule fixes it:
 *   window x = ((double)x - l) / ((double)r - l) * W (W = 2 nx; y likewise
 *   with H = 2 ny), snapped to 1/256 pixel (rint(x * 256)); sample (px, py)
 *   sits at (256 px + 128, 256 py + 128); a triangle with zero snapped area
 *   covers nothing; otherwise a sample is covered when every edge function
 *   of the counter-clockwise-ordered snapped triangle is positive, or zero
 *   on a top or left edge (dy < 0, or dy == 0 and dx < 0), so samples on
 *   shared edges count once;
 *   a covered sample counts in slice s when z_T <= zs(s), z_T the height of
 *   the triangle's plane (float vertices, double arithmetic, no fused ops)
 *   at the sample's world point l + (px + 0.5) ((double)r - l) / W, and zs(s)
 *   the Level1 ray-origin table; triangles whose plane is vertical (normal
 *   z exactly 0) count nowhere;
 *   counts are exact (GL's 8-bit stencil saturates at 255).
 * counts[py * W + px] receives the fragment count of slice s; returns 0, or
 * -1 when the grid is too large for the rule (W or H above 2^20). */
int orc_stencil_counts(const float *tris, int64_t ntris, const orc_grid *g, int slice, int32_t *counts);
/* Level1 inside (1) / outside (0) per voxel by that fill: [z][y][x]. */
int orc_stencil_inside(const float *tris, int64_t ntris, const orc_grid *g, uint8_t *inside);

/* Full pipeline.  mode ORC_EXACT runs the literal loops (all-triangle
 * parity, every box of every index AABB and bin through the full SAT);
 * ORC_FAST runs the result-identical forms: the column-factorised parity
 * (with monotone layer runs at Level2) and the SAT sweeps screened by
 * orc_tri_box_factored's predicate groups; ORC_PLAN is the CPU baseline's
 * "fast-identical" composition: the column-factorised parity with a t test
 * per layer, and the literal SAT loops.  threads<=0: all. */
int orc_voxelize(const float *tris, int64_t ntris, const orc_grid *g, int mode,
                 int threads, orc_result *out);
void orc_result_free(orc_result *r);

/* Normals: VectorNormalize, the flat-soup face normal, Level1 normals per
 * boundary voxel [nb*3] and Level2 normals per Level2 voxel [nb*n2^3*3]
 * (l2n may be NULL), and the .raw normal byte. */
void orc_normalize(float v[3]);
void orc_face_normal(const float tv[9], float n[3]);
void orc_face_normal_batch(int64_t n, const float *tv, float *out);
void orc_normalize_batch(int64_t n, const float *v, float *out);
/* Level1 average of nl lists of face normals fn (CSR off[nl+1]), in order. */
void orc_average_normals(int64_t nl, const int64_t *off, const float *fn, float *out);
int orc_normals(const float *tris, const orc_grid *g, const orc_result *r, int threads, float *l1n,
                float *l2n);
uint8_t orc_quantize_normal(float v);
void orc_quantize_normals(int64_t n, const float *v, uint8_t *out);

/* Writes Obj<obj_id>{VoxelConfig.txt,Level1InOut.raw,Level1BoundaryPrefixSum.raw,
 * Level2InOut.raw} under dir.  Returns 0 on success. */
int orc_save_raw(const orc_grid *g, const orc_result *r, int obj_id, const char *dir);
/* Obj<obj_id>Level1Normal.raw and (with Level2) Obj<obj_id>Level2Normal.raw
 * from the normals of orc_normals.  Returns 0 on success. */
int orc_save_raw_normals(const orc_grid *g, const orc_result *r, const float *l1n, const float *l2n, int obj_id,
                         const char *dir);

#ifdef __cplusplus
}
#endif
#endif
