// Thin extern "C" entry points over the reference's own predicates so that
// the oracle tests and the golden-vector script can call them through ctypes.
// Linked only into oracle/_ref/libgpvref.so together with
// src/TriBoxIntersection.cpp and src/TriRayIntersection.cpp (test tooling).
#include <stdint.h>

int TriBoxOverlap(float boxcenter[3], float boxhalfsize[3], float triverts[3][3]);
int triangle_ray_intersection(const float V1[3], const float V2[3], const float V3[3],
                              const float O[3], const float D[3], float* out);

extern "C" {

// n boxes/triangles: c[n*3], h[n*3], tv[n*9] -> res[n] (0/1)
void ref_tri_box_batch(int64_t n, const float* c, const float* h, const float* tv, uint8_t* res)
{
    for (int64_t i = 0; i < n; i++) {
        float cc[3] = {c[3 * i], c[3 * i + 1], c[3 * i + 2]};
        float hh[3] = {h[3 * i], h[3 * i + 1], h[3 * i + 2]};
        float t[3][3];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) t[a][b] = tv[9 * i + 3 * a + b];
        res[i] = (uint8_t)TriBoxOverlap(cc, hh, t);
    }
}

// n rays/triangles: tv[n*9], o[n*3], d[n*3] -> hit[n], t[n] (t only valid on hit)
void ref_tri_ray_batch(int64_t n, const float* tv, const float* o, const float* d,
                       uint8_t* hit, float* t)
{
    for (int64_t i = 0; i < n; i++) {
        float out = 0.0f;
        hit[i] = (uint8_t)triangle_ray_intersection(tv + 9 * i, tv + 9 * i + 3, tv + 9 * i + 6,
                                                    o + 3 * i, d + 3 * i, &out);
        t[i] = hit[i] ? out : 0.0f;
    }
}

}
