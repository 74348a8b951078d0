// extern "C" entry points over the reference's own vector arithmetic
// (includes/FloatVector.h: Float3, operator-, operator+=, operator/,
// VectorCrossProduct, VectorNormalize) for the normals checks.  The header is
// included from where it lies in the reference tree; linked only into
// oracle/_ref/libgpvref.so (test tooling).
#include <math.h>
#include <stdint.h>

#include "FloatVector.h"

extern "C" {

// face normal of tv[n*9] exactly as src/Object.cpp:3233-3240 writes it
void ref_face_normal_batch(int64_t n, const float* tv, float* out)
{
    for (int64_t i = 0; i < n; i++) {
        const float* t = tv + 9 * i;
        Float3 v0 = Float3(t[0], t[1], t[2]);
        Float3 v1 = Float3(t[3], t[4], t[5]);
        Float3 v2 = Float3(t[6], t[7], t[8]);
        Float3 side1 = v1 - v0;
        Float3 side2 = v2 - v0;
        Float3 fn = VectorCrossProduct(side1, side2);
        VectorNormalize(fn);
        out[3 * i] = fn[0];
        out[3 * i + 1] = fn[1];
        out[3 * i + 2] = fn[2];
    }
}

void ref_normalize_batch(int64_t n, const float* v, float* out)
{
    for (int64_t i = 0; i < n; i++) {
        Float3 a = Float3(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
        VectorNormalize(a);
        out[3 * i] = a[0];
        out[3 * i + 1] = a[1];
        out[3 * i + 2] = a[2];
    }
}

// nl lists of unit normals fn (CSR off[nl+1]) -> the Level1 average of
// src/Object.cpp:3226-3250: sum with +=, divide by the int count, normalize
void ref_average_normals(int64_t nl, const int64_t* off, const float* fn, float* out)
{
    for (int64_t l = 0; l < nl; l++) {
        int numTri = (int)(off[l + 1] - off[l]);
        Float3 sumNorm = Float3(0, 0, 0);
        Float3 avgNorm = Float3(0, 0, 0);
        for (int i = 0; i < numTri; i++) {
            const float* f = fn + 3 * (off[l] + i);
            Float3 faceNormal = Float3(f[0], f[1], f[2]);
            sumNorm += faceNormal;
        }
        if (numTri != 0) {
            avgNorm = sumNorm / numTri;
            VectorNormalize(avgNorm);
        }
        out[3 * l] = avgNorm[0];
        out[3 * l + 1] = avgNorm[1];
        out[3 * l + 2] = avgNorm[2];
    }
}

}
