/*
 * gpv_mesh.h -- host-side mesh input for libgpv: deterministic synthetic
 * generators for the benchmark workloads and the reference's file readers.
 *
 * A gpv_mesh is an indexed triangle mesh (float xyz per vertex, three int32
 * vertex indices per triangle).  gpv_mesh_flatten produces the flat soup of
 * Object::CreateFlatTriangleData (src/Object.cpp:3496-3527).
 *
 * Generators are seeded (splitmix64/xoshiro256**, independent of the C++
 * library's distributions) so the same call gives bit-identical vertices on
 * every machine.
 */
#ifndef GPV_MESH_H
#define GPV_MESH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gpv_mesh {
    int64_t nverts;
    int64_t ntris;
    float *verts;      /* nverts * 3 */
    int32_t *faces;    /* ntris * 3, 0-based */
    int bbox_from_faces; /* 1: bbox over referenced vertices only (OFF reader) */
} gpv_mesh;

void gpv_mesh_free(gpv_mesh *m);

/* Flat soup: soup[t*9 + k*3 + c] = verts[faces[t*3+k]*3 + c]. */
int gpv_mesh_flatten(const gpv_mesh *m, float *soup);
/* bbox + pad of the mesh, honouring bbox_from_faces (gpv.h gpv_mesh_bbox). */
int gpv_mesh_bounds(const gpv_mesh *m, float bmin[3], float bmax[3], float *max_model_size);

/* Icosahedron subdivided `level` times with midpoint projection, radius r
 * around center, per-vertex radius scaled by 1 + U[-jitter, jitter].
 * level 4: 5,120 triangles, 2,562 vertices. */
int gpv_mesh_icosphere(int level, float radius, const float center[3], float jitter,
                       uint64_t seed, gpv_mesh *out);

/* (p,q) torus-knot tube: centerline ((2+cos q t) cos p t, (2+cos q t) sin p t,
 * -sin q t), `segments` rings of `ring` vertices at tube radius `tube`
 * (knot units), ring radii scaled by 1 + U[-jitter, jitter], then scaled so
 * the bbox max extent is `extent` and centered at (0.5,0.5,0.5).
 * 2*segments*ring triangles. */
int gpv_mesh_torus_knot(int p, int q, int segments, int ring, float tube, float jitter,
                        float extent, uint64_t seed, gpv_mesh *out);

/* Torus with major/minor radius ratio `ratio`, nu x nv quads, randomly
 * rotated (seeded), scaled to `extent`, centered at (0.5,0.5,0.5). */
int gpv_mesh_torus(int nu, int nv, float ratio, float extent, uint64_t seed, gpv_mesh *out);

/* Box with 1..6 cylindrical through-holes (radius U[0.03,0.12] box-relative,
 * 64 facets), faces tessellated so the total is close to `target_tris`,
 * randomly rotated, scaled to `extent`, centered at (0.5,0.5,0.5).
 * Watertight (shared lattice vertices along every seam). */
int gpv_mesh_drilled_box(int64_t target_tris, float extent, uint64_t seed, gpv_mesh *out);

/* Mixed workload member: kind 0 icosphere, 1 torus, 2 drilled box, sized to
 * about target_tris triangles (icospheres come in 20 * 4^level triangles,
 * levels 2 to 8, the nearest level in log scale that does not overshoot
 * targets above 163,840 by more than half). */
int gpv_mesh_mixed(int kind, int64_t target_tris, uint64_t seed, gpv_mesh *out);

/* Exactness stressors applied in place (used for the stress workload):
 *  - vertices snapped onto the Level1/Level2 lattice planes of the grid the
 *    mesh's own bbox derives (voxel_count, voxel_count2) with probability
 *    p_snap (bbox-defining vertices are left alone so the grid is unchanged);
 *  - a fraction f_coplanar of triangles made coplanar with a Level1 voxel
 *    face: their three vertices moved onto the lattice plane nearest the
 *    centroid along a random axis (shared vertices move, so the mesh stays
 *    watertight; later moves may undo earlier ones, so the fraction is
 *    approximate);
 *  - a fraction f_sliver of triangles split into slivers (aspect > 1e3) by
 *    inserting a vertex at 1e-4 of an edge, both neighbours split (watertight). */
int gpv_mesh_stress(gpv_mesh *m, int voxel_count, int voxel_count2, float p_snap, float f_sliver,
                    float f_coplanar, uint64_t seed);

/* Readers with the reference's semantics: Object::ReadObject
 * (src/Object.cpp:395-584; bbox over every `v` line) and
 * Object::ReadOFFObject (src/Object.cpp:171-317; bbox over the triangles). */
int gpv_mesh_read_obj(const char *path, gpv_mesh *out);
int gpv_mesh_read_off(const char *path, gpv_mesh *out);

#ifdef __cplusplus
}
#endif
#endif /* GPV_MESH_H */
