// gpv_mesh.cpp -- deterministic mesh generators and the reference's mesh
// readers (include/gpv_mesh.h).  Host only.
#include "gpv.h"
#include "gpv_mesh.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

// splitmix64 seeding + xoshiro256**: fixed algorithms, identical streams on
// every machine.
struct Rng {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t& x)
    {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    explicit Rng(uint64_t seed)
    {
        uint64_t x = seed;
        for (auto& v : s) v = splitmix(x);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double u01() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double uni(double a, double b) { return a + (b - a) * u01(); }
    int64_t randint(int64_t a, int64_t b) { return a + (int64_t)(next() % (uint64_t)(b - a + 1)); }
};

struct Builder {
    std::vector<double> v;    // xyz
    std::vector<int32_t> f;   // abc
    int32_t add(double x, double y, double z)
    {
        v.push_back(x);
        v.push_back(y);
        v.push_back(z);
        return (int32_t)(v.size() / 3 - 1);
    }
    void tri(int32_t a, int32_t b, int32_t c)
    {
        f.push_back(a);
        f.push_back(b);
        f.push_back(c);
    }
    int64_t nv() const { return (int64_t)v.size() / 3; }
};

// scale so the bbox max extent is `extent`, center the bbox at 0.5
void fit(Builder& B, double extent)
{
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (int64_t i = 0; i < B.nv(); i++)
        for (int q = 0; q < 3; q++) {
            mn[q] = std::min(mn[q], B.v[3 * i + q]);
            mx[q] = std::max(mx[q], B.v[3 * i + q]);
        }
    const double e = std::max(mx[0] - mn[0], std::max(mx[1] - mn[1], mx[2] - mn[2]));
    const double s = e > 0 ? extent / e : 1.0;
    for (int64_t i = 0; i < B.nv(); i++)
        for (int q = 0; q < 3; q++) B.v[3 * i + q] = (B.v[3 * i + q] - 0.5 * (mn[q] + mx[q])) * s + 0.5;
}

// uniform random rotation (unit quaternion from three uniforms)
void rotate(Builder& B, Rng& R)
{
    const double u1 = R.u01(), u2 = R.u01(), u3 = R.u01();
    const double a = std::sqrt(1 - u1), b = std::sqrt(u1);
    const double qx = a * std::sin(2 * M_PI * u2), qy = a * std::cos(2 * M_PI * u2);
    const double qz = b * std::sin(2 * M_PI * u3), qw = b * std::cos(2 * M_PI * u3);
    const double m[9] = {1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw),
                         2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw),
                         2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)};
    for (int64_t i = 0; i < B.nv(); i++) {
        const double x = B.v[3 * i], y = B.v[3 * i + 1], z = B.v[3 * i + 2];
        B.v[3 * i] = m[0] * x + m[1] * y + m[2] * z;
        B.v[3 * i + 1] = m[3] * x + m[4] * y + m[5] * z;
        B.v[3 * i + 2] = m[6] * x + m[7] * y + m[8] * z;
    }
}

int emit(const Builder& B, gpv_mesh* out)
{
    memset(out, 0, sizeof(*out));
    out->nverts = B.nv();
    out->ntris = (int64_t)B.f.size() / 3;
    out->verts = (float*)malloc(sizeof(float) * std::max<size_t>(B.v.size(), 1));
    out->faces = (int32_t*)malloc(sizeof(int32_t) * std::max<size_t>(B.f.size(), 1));
    if (!out->verts || !out->faces) {
        gpv_mesh_free(out);
        return GPV_ERR_OOM;
    }
    for (size_t i = 0; i < B.v.size(); i++) out->verts[i] = (float)B.v[i];
    std::copy(B.f.begin(), B.f.end(), out->faces);
    return GPV_OK;
}

void icosahedron(Builder& B)
{
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    const double P[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                             {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    const int F[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                          {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
                          {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    for (auto& p : P) {
        const double n = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        B.add(p[0] / n, p[1] / n, p[2] / n);
    }
    for (auto& f : F) B.tri(f[0], f[1], f[2]);
}

void subdivide(Builder& B)
{
    std::map<std::pair<int32_t, int32_t>, int32_t> mid;
    auto midpoint = [&](int32_t a, int32_t b) {
        const auto key = std::make_pair(std::min(a, b), std::max(a, b));
        auto it = mid.find(key);
        if (it != mid.end()) return it->second;
        double x = B.v[3 * a] + B.v[3 * b], y = B.v[3 * a + 1] + B.v[3 * b + 1], z = B.v[3 * a + 2] + B.v[3 * b + 2];
        const double n = std::sqrt(x * x + y * y + z * z);
        const int32_t id = B.add(x / n, y / n, z / n);
        mid.emplace(key, id);
        return id;
    };
    std::vector<int32_t> old;
    old.swap(B.f);
    for (size_t i = 0; i < old.size(); i += 3) {
        const int32_t a = old[i], b = old[i + 1], c = old[i + 2];
        const int32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
        B.tri(a, ab, ca);
        B.tri(b, bc, ab);
        B.tri(c, ca, bc);
        B.tri(ab, bc, ca);
    }
}

// --- drilled box -----------------------------------------------------------
struct BoxBuilder {
    Builder B;
    std::unordered_map<uint64_t, int32_t> lattice;
    int64_t L[3];          // lattice points per axis - 1
    double half[3];
    int32_t lat(int64_t i, int64_t j, int64_t k)
    {
        const uint64_t key = ((uint64_t)i << 42) | ((uint64_t)j << 21) | (uint64_t)k;
        auto it = lattice.find(key);
        if (it != lattice.end()) return it->second;
        const int64_t ijk[3] = {i, j, k};
        double x[3];
        for (int q = 0; q < 3; q++) x[q] = -half[q] + 2.0 * half[q] * (double)ijk[q] / (double)L[q];
        const int32_t id = B.add(x[0], x[1], x[2]);
        lattice.emplace(key, id);
        return id;
    }
};

}  // namespace

extern "C" {

void gpv_mesh_free(gpv_mesh* m)
{
    if (!m) return;
    free(m->verts);
    free(m->faces);
    memset(m, 0, sizeof(*m));
}

int gpv_mesh_flatten(const gpv_mesh* m, float* soup)
{
    if (!m || !soup) return GPV_ERR_INVALID;
    for (int64_t t = 0; t < m->ntris; t++)
        for (int k = 0; k < 3; k++) {
            const int32_t vi = m->faces[3 * t + k];
            if (vi < 0 || vi >= m->nverts) return GPV_ERR_INVALID;
            for (int c = 0; c < 3; c++) soup[t * 9 + k * 3 + c] = m->verts[3 * (int64_t)vi + c];
        }
    return GPV_OK;
}

int gpv_mesh_bounds(const gpv_mesh* m, float bmin[3], float bmax[3], float* mms)
{
    if (!m) return GPV_ERR_INVALID;
    if (!m->bbox_from_faces) return gpv_mesh_bbox(m->verts, m->nverts, bmin, bmax, mms);
    // OFF semantics: bbox over the triangles' vertices in face order
    std::vector<float> pts;
    pts.reserve((size_t)m->ntris * 9);
    for (int64_t t = 0; t < m->ntris; t++)
        for (int k = 0; k < 3; k++)
            for (int c = 0; c < 3; c++) pts.push_back(m->verts[3 * (int64_t)m->faces[3 * t + k] + c]);
    return gpv_mesh_bbox(pts.data(), (int64_t)pts.size() / 3, bmin, bmax, mms);
}

int gpv_mesh_icosphere(int level, float radius, const float center[3], float jitter, uint64_t seed,
                       gpv_mesh* out)
{
    if (!out || level < 0 || level > 9 || !center) return GPV_ERR_INVALID;
    Builder B;
    icosahedron(B);
    for (int l = 0; l < level; l++) subdivide(B);
    Rng R(seed);
    for (int64_t i = 0; i < B.nv(); i++) {
        const double s = (double)radius * (1.0 + R.uni(-jitter, jitter));
        for (int q = 0; q < 3; q++) B.v[3 * i + q] = center[q] + s * B.v[3 * i + q];
    }
    return emit(B, out);
}

int gpv_mesh_torus_knot(int p, int q, int segments, int ring, float tube, float jitter, float extent,
                        uint64_t seed, gpv_mesh* out)
{
    if (!out || segments < 3 || ring < 3 || !(tube > 0)) return GPV_ERR_INVALID;
    Builder B;
    Rng R(seed);
    for (int i = 0; i < segments; i++) {
        const double t = 2 * M_PI * i / segments;
        const double r = std::cos(q * t) + 2;
        const double c[3] = {r * std::cos(p * t), r * std::sin(p * t), -std::sin(q * t)};
        const double dr = -q * std::sin(q * t);
        double T[3] = {dr * std::cos(p * t) - p * r * std::sin(p * t), dr * std::sin(p * t) + p * r * std::cos(p * t),
                       -q * std::cos(q * t)};
        double n = std::sqrt(T[0] * T[0] + T[1] * T[1] + T[2] * T[2]);
        for (double& x : T) x /= n;
        // N = T x z (never degenerate: the tangent's xy part is >= p*r >= 3)
        double N[3] = {T[1], -T[0], 0.0};
        n = std::sqrt(N[0] * N[0] + N[1] * N[1]);
        N[0] /= n;
        N[1] /= n;
        const double Bv[3] = {T[1] * N[2] - T[2] * N[1], T[2] * N[0] - T[0] * N[2], T[0] * N[1] - T[1] * N[0]};
        for (int k = 0; k < ring; k++) {
            const double a = 2 * M_PI * k / ring;
            const double rr = (double)tube * (1.0 + R.uni(-jitter, jitter));
            B.add(c[0] + rr * (std::cos(a) * N[0] + std::sin(a) * Bv[0]),
                  c[1] + rr * (std::cos(a) * N[1] + std::sin(a) * Bv[1]),
                  c[2] + rr * (std::cos(a) * N[2] + std::sin(a) * Bv[2]));
        }
    }
    for (int i = 0; i < segments; i++) {
        const int i2 = (i + 1) % segments;
        for (int k = 0; k < ring; k++) {
            const int k2 = (k + 1) % ring;
            const int32_t a = i * ring + k, b = i2 * ring + k, c = i2 * ring + k2, d = i * ring + k2;
            B.tri(a, b, c);
            B.tri(a, c, d);
        }
    }
    fit(B, extent);
    return emit(B, out);
}

int gpv_mesh_torus(int nu, int nv, float ratio, float extent, uint64_t seed, gpv_mesh* out)
{
    if (!out || nu < 3 || nv < 3 || !(ratio > 1)) return GPV_ERR_INVALID;
    Builder B;
    const double Rr = ratio, r = 1.0;
    for (int i = 0; i < nu; i++) {
        const double u = 2 * M_PI * i / nu;
        for (int j = 0; j < nv; j++) {
            const double v = 2 * M_PI * j / nv;
            B.add((Rr + r * std::cos(v)) * std::cos(u), (Rr + r * std::cos(v)) * std::sin(u), r * std::sin(v));
        }
    }
    for (int i = 0; i < nu; i++)
        for (int j = 0; j < nv; j++) {
            const int32_t a = i * nv + j, b = ((i + 1) % nu) * nv + j, c = ((i + 1) % nu) * nv + (j + 1) % nv,
                          d = i * nv + (j + 1) % nv;
            B.tri(a, b, c);
            B.tri(a, c, d);
        }
    Rng R(seed);
    rotate(B, R);
    fit(B, extent);
    return emit(B, out);
}

int gpv_mesh_drilled_box(int64_t target, float extent, uint64_t seed, gpv_mesh* out)
{
    if (!out || target < 12) return GPV_ERR_INVALID;
    Rng R(seed);
    BoxBuilder X;
    int m[3];
    for (int q = 0; q < 3; q++) {
        X.half[q] = R.uni(0.3, 0.5);
        m[q] = std::max(2, (int)std::floor(2 * X.half[q] / 0.26));
    }
    const int axis = (int)R.randint(0, 2);
    const int ua = (axis + 1) % 3, va = (axis + 2) % 3;
    const int ncell = m[ua] * m[va];
    const int nh = (int)std::min<int64_t>(R.randint(1, 6), ncell);
    std::vector<int> cells(ncell);
    for (int i = 0; i < ncell; i++) cells[i] = i;
    for (int i = 0; i < nh; i++) std::swap(cells[i], cells[i + (int)R.randint(0, ncell - 1 - i)]);
    const double cw_u = 2 * X.half[ua] / m[ua], cw_v = 2 * X.half[va] / m[va];
    const int64_t cells_total = 2LL * (m[0] * m[1] + m[1] * m[2] + m[0] * m[2]);
    const int s = (int)std::max<double>(1.0, std::floor(std::sqrt((double)target / (2.0 * cells_total))));
    for (int q = 0; q < 3; q++) X.L[q] = (int64_t)m[q] * s;
    const int NF = 64;
    // canonical hole circles: facet f at (ua, va) angle 2*pi*f/NF, one ring per side
    std::vector<int> hole_of(ncell, -1);
    std::vector<double> hc_u(nh), hc_v(nh);
    std::vector<std::vector<int32_t>> ring[2];
    ring[0].resize(nh);
    ring[1].resize(nh);
    for (int h = 0; h < nh; h++) {
        hole_of[cells[h]] = h;
        const int cu = cells[h] / m[va], cv = cells[h] % m[va];
        hc_u[h] = -X.half[ua] + cw_u * (cu + 0.5);
        hc_v[h] = -X.half[va] + cw_v * (cv + 0.5);
        const double rad = std::min(R.uni(0.03, 0.12), 0.45 * std::min(cw_u, cw_v));
        for (int side = 0; side < 2; side++)
            for (int f = 0; f < NF; f++) {
                const double a = 2 * M_PI * f / NF;
                double x[3];
                x[axis] = side ? X.half[axis] : -X.half[axis];
                x[ua] = hc_u[h] + rad * std::cos(a);
                x[va] = hc_v[h] + rad * std::sin(a);
                ring[side][h].push_back(X.B.add(x[0], x[1], x[2]));
            }
    }
    // faces: for axis k and side, (u, v) with u x v = outward normal
    for (int k = 0; k < 3; k++) {
        for (int side = 0; side < 2; side++) {
            int u = (k + 1) % 3, v = (k + 2) % 3;
            if (side == 0) std::swap(u, v);
            const int64_t kk = side ? X.L[k] : 0;
            auto P = [&](int64_t iu, int64_t iv) {
                int64_t ijk[3];
                ijk[k] = kk;
                ijk[u] = iu;
                ijk[v] = iv;
                return X.lat(ijk[0], ijk[1], ijk[2]);
            };
            for (int cu = 0; cu < m[u]; cu++)
                for (int cv = 0; cv < m[v]; cv++) {
                    int hole = -1;
                    if (k == axis) hole = hole_of[(u == ua) ? cu * m[va] + cv : cv * m[va] + cu];
                    const int64_t u0 = (int64_t)cu * s, v0 = (int64_t)cv * s;
                    if (hole < 0) {
                        for (int a = 0; a < s; a++)
                            for (int b = 0; b < s; b++) {
                                const int32_t p00 = P(u0 + a, v0 + b), p10 = P(u0 + a + 1, v0 + b),
                                              p11 = P(u0 + a + 1, v0 + b + 1), p01 = P(u0 + a, v0 + b + 1);
                                X.B.tri(p00, p10, p11);
                                X.B.tri(p00, p11, p01);
                            }
                        continue;
                    }
                    // cell perimeter, CCW in (u, v)
                    std::vector<int32_t> per;
                    for (int a = 0; a < s; a++) per.push_back(P(u0 + a, v0));
                    for (int b = 0; b < s; b++) per.push_back(P(u0 + s, v0 + b));
                    for (int a = s; a > 0; a--) per.push_back(P(u0 + a, v0 + s));
                    for (int b = s; b > 0; b--) per.push_back(P(u0, v0 + b));
                    const double cu_x = (u == ua) ? hc_u[hole] : hc_v[hole];
                    const double cv_x = (u == ua) ? hc_v[hole] : hc_u[hole];
                    // circle in the face's CCW order, point j at face angle 2*pi*j/NF:
                    // side 1 (u = ua): canonical order; side 0 (u = va): f = (NF/4 - j) mod NF
                    std::vector<int32_t> circ(NF);
                    for (int j = 0; j < NF; j++)
                        circ[j] = (u == ua) ? ring[side][hole][j] : ring[side][hole][((NF / 4 - j) % NF + NF) % NF];
                    auto ang = [&](int32_t id) {
                        const double a = std::atan2(X.B.v[3 * id + v] - cv_x, X.B.v[3 * id + u] - cu_x);
                        return a < 0 ? a + 2 * M_PI : a;
                    };
                    int start = 0;
                    double best = 1e9;
                    for (size_t i = 0; i < per.size(); i++) {
                        const double a = ang(per[i]);
                        if (a < best) { best = a; start = (int)i; }
                    }
                    std::rotate(per.begin(), per.begin() + start, per.end());
                    const int np = (int)per.size();
                    int i = 0, j = 0;
                    while (i < np || j < NF) {
                        const double an = (i < np) ? (i + 1 < np ? ang(per[i + 1]) : 2 * M_PI) : 1e9;
                        const double cn = (j < NF) ? 2 * M_PI * (j + 1) / NF : 1e9;
                        if (an <= cn) {
                            X.B.tri(per[i % np], per[(i + 1) % np], circ[j % NF]);
                            i++;
                        } else {
                            X.B.tri(per[i % np], circ[(j + 1) % NF], circ[j % NF]);
                            j++;
                        }
                    }
                }
        }
    }
    // cylinder walls between the two rings, oriented toward the hole axis
    for (int h = 0; h < nh; h++) {
        for (int f = 0; f < NF; f++) {
            const int32_t b0 = ring[0][h][f], b1 = ring[0][h][(f + 1) % NF];
            const int32_t t1 = ring[1][h][(f + 1) % NF], t0 = ring[1][h][f];
            auto oriented = [&](int32_t a, int32_t b, int32_t c) {
                const double* A = &X.B.v[3 * a];
                const double* Bq = &X.B.v[3 * b];
                const double* C = &X.B.v[3 * c];
                const double e1[3] = {Bq[0] - A[0], Bq[1] - A[1], Bq[2] - A[2]};
                const double e2[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
                const double nrm[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                                       e1[0] * e2[1] - e1[1] * e2[0]};
                double rad[3] = {0, 0, 0};
                rad[ua] = (A[ua] + Bq[ua] + C[ua]) / 3 - hc_u[h];
                rad[va] = (A[va] + Bq[va] + C[va]) / 3 - hc_v[h];
                if (nrm[0] * rad[0] + nrm[1] * rad[1] + nrm[2] * rad[2] > 0) X.B.tri(a, c, b);
                else X.B.tri(a, b, c);
            };
            oriented(b0, b1, t1);
            oriented(b0, t1, t0);
        }
    }
    rotate(X.B, R);
    fit(X.B, extent);
    return emit(X.B, out);
}

int gpv_mesh_mixed(int kind, int64_t target, uint64_t seed, gpv_mesh* out)
{
    if (!out || target < 20) return GPV_ERR_INVALID;
    Rng R(seed ^ 0x5bd1e995ull);
    if (kind == 0) {
        // nearest subdivision level in log scale (20 * 4^level triangles);
        // levels 7 and 8 (327,680 and 1,310,720 triangles) only when they do
        // not overshoot the target by more than half
        int level = (int)std::lround(std::log((double)target / 20.0) / std::log(4.0));
        level = std::max(2, std::min(level, 8));
        if (level >= 7 && 20.0 * std::pow(4.0, level) > 1.5 * (double)target) level--;
        const double r = R.uni(0.25, 0.45);
        float c[3];
        for (int q = 0; q < 3; q++) c[q] = (float)(0.5 + R.uni(-1, 1) * (0.48 - r));
        return gpv_mesh_icosphere(level, (float)r, c, 0.02f, seed, out);
    }
    if (kind == 1) {
        const double ratio = R.uni(2.0, 5.0);
        const int nu = std::max(3, (int)std::lround(std::sqrt((double)target / 2.0 * ratio)));
        const int nv = std::max(3, (int)std::lround((double)target / 2.0 / nu));
        return gpv_mesh_torus(nu, nv, (float)ratio, (float)R.uni(0.6, 0.95), seed, out);
    }
    if (kind == 2) return gpv_mesh_drilled_box(target, (float)R.uni(0.6, 0.95), seed, out);
    return GPV_ERR_INVALID;
}

int gpv_mesh_stress(gpv_mesh* m, int voxel_count, int voxel_count2, float p_snap, float f_sliver, float f_coplanar,
                    uint64_t seed)
{
    if (!m || m->nverts <= 0 || voxel_count < 1 || voxel_count2 < 1) return GPV_ERR_INVALID;
    Rng R(seed);
    float bmin[3], bmax[3], mms;
    int r = gpv_mesh_bbox(m->verts, m->nverts, bmin, bmax, &mms);
    if (r != GPV_OK) return r;
    gpv_params prm{voxel_count, voxel_count2, 1, 0, 0};
    gpv_grid_info info;
    if ((r = gpv_grid_derive(bmin, bmax, mms, &prm, &info)) != GPV_OK) return r;
    float lo[3], hi[3];
    for (int q = 0; q < 3; q++) lo[q] = hi[q] = m->verts[q];
    for (int64_t i = 0; i < m->nverts; i++)
        for (int q = 0; q < 3; q++) {
            lo[q] = std::min(lo[q], m->verts[3 * i + q]);
            hi[q] = std::max(hi[q], m->verts[3 * i + q]);
        }
    // lattice snapping (Level1 planes bmin + i*g, Level2 planes bmin + i*g2)
    for (int64_t i = 0; i < m->nverts; i++) {
        if (R.u01() >= p_snap) continue;
        const bool l2 = R.u01() < 0.5;
        float* p = m->verts + 3 * i;
        bool extreme = false;
        for (int q = 0; q < 3; q++) extreme |= (p[q] == lo[q] || p[q] == hi[q]);
        if (extreme) continue;
        for (int q = 0; q < 3; q++) {
            const double step = l2 ? (double)info.g2[q] : (double)info.g[q];
            const double k = std::nearbyint(((double)p[q] - bmin[q]) / step);
            const float snapped = (float)((double)bmin[q] + k * step);
            if (snapped > lo[q] && snapped < hi[q]) p[q] = snapped;
        }
    }
    // face-coplanar triangles: the three vertices onto one Level1 lattice plane
    if (f_coplanar > 0 && m->ntris > 0) {
        const int64_t want = (int64_t)std::llround(f_coplanar * (double)m->ntris);
        for (int64_t s = 0, tries = 0; s < want && tries < 20 * want + 100; tries++) {
            const int64_t t = R.randint(0, m->ntris - 1);
            const int q = (int)R.randint(0, 2);
            float* p[3];
            bool extreme = false;
            for (int k = 0; k < 3; k++) {
                p[k] = m->verts + 3 * (int64_t)m->faces[3 * t + k];
                for (int a = 0; a < 3; a++) extreme |= (p[k][a] == lo[a] || p[k][a] == hi[a]);
            }
            if (extreme) continue;
            const double c = ((double)p[0][q] + p[1][q] + p[2][q]) / 3.0;
            const double step = (double)info.g[q];
            const double k = std::nearbyint((c - bmin[q]) / step);
            const float plane = (float)((double)bmin[q] + k * step);
            if (!(plane > lo[q] && plane < hi[q])) continue;
            for (int v = 0; v < 3; v++) p[v][q] = plane;
            s++;
        }
    }
    // slivers: split both triangles of an edge at 1e-4 of its length
    if (f_sliver > 0) {
        std::unordered_map<uint64_t, int64_t> edge;   // (a<<32|b) -> triangle
        for (int64_t t = 0; t < m->ntris; t++)
            for (int k = 0; k < 3; k++) {
                const uint32_t a = (uint32_t)m->faces[3 * t + k], b = (uint32_t)m->faces[3 * t + (k + 1) % 3];
                edge[((uint64_t)a << 32) | b] = t;
            }
        const int64_t splits = (int64_t)std::llround(f_sliver * (double)m->ntris / 2.0);
        std::vector<float> V(m->verts, m->verts + 3 * m->nverts);
        std::vector<int32_t> F(m->faces, m->faces + 3 * m->ntris);
        std::vector<uint8_t> used(m->ntris, 0);
        for (int64_t s = 0, tries = 0; s < splits && tries < 20 * splits + 100; tries++) {
            const int64_t t = R.randint(0, m->ntris - 1);
            const int k = (int)R.randint(0, 2);
            if (used[t]) continue;
            const int32_t a = F[3 * t + k], b = F[3 * t + (k + 1) % 3], c = F[3 * t + (k + 2) % 3];
            auto it = edge.find(((uint64_t)(uint32_t)b << 32) | (uint32_t)a);
            if (it == edge.end() || used[it->second]) continue;
            const int64_t t2 = it->second;
            int k2 = 0;
            while (!(F[3 * t2 + k2] == b && F[3 * t2 + (k2 + 1) % 3] == a)) k2++;
            const int32_t d = F[3 * t2 + (k2 + 2) % 3];
            const int32_t mid = (int32_t)(V.size() / 3);
            for (int q = 0; q < 3; q++) V.push_back((float)(V[3 * a + q] + 1e-4 * ((double)V[3 * b + q] - V[3 * a + q])));
            // t = (a, b, c) -> (a, mid, c) + (mid, b, c); t2 = (b, a, d) -> (b, mid, d) + (mid, a, d)
            F[3 * t + 0] = a; F[3 * t + 1] = mid; F[3 * t + 2] = c;
            F.push_back(mid); F.push_back(b); F.push_back(c);
            F[3 * t2 + 0] = b; F[3 * t2 + 1] = mid; F[3 * t2 + 2] = d;
            F.push_back(mid); F.push_back(a); F.push_back(d);
            used[t] = used[t2] = 1;
            s++;
        }
        float* nv = (float*)malloc(V.size() * sizeof(float));
        int32_t* nf = (int32_t*)malloc(F.size() * sizeof(int32_t));
        if (!nv || !nf) {
            free(nv);
            free(nf);
            return GPV_ERR_OOM;
        }
        std::copy(V.begin(), V.end(), nv);
        std::copy(F.begin(), F.end(), nf);
        free(m->verts);
        free(m->faces);
        m->verts = nv;
        m->faces = nf;
        m->nverts = (int64_t)V.size() / 3;
        m->ntris = (int64_t)F.size() / 3;
    }
    return GPV_OK;
}

// ---------------------------------------------------------------------------
// readers
// ---------------------------------------------------------------------------

// split of src/Utilities.cpp:989-1005: empty tokens are kept, a trailing
// separator yields a final empty token
static std::vector<std::string> ref_split(std::string str, const std::string& tok)
{
    std::vector<std::string> out;
    while (!str.empty()) {
        const size_t i = str.find(tok);
        if (i != std::string::npos) {
            out.push_back(str.substr(0, i));
            str = str.substr(i + tok.size());
            if (str.empty()) out.push_back(str);
        } else {
            out.push_back(str);
            str.clear();
        }
    }
    return out;
}

// Object::ReadObject, src/Object.cpp:395-584: a line is processed only when
// getline ended on a newline (the reference breaks on !in.good() first);
// words split on ' ' or '\t', whichever gives more tokens (ties: '\t');
// 'v' lines feed the bbox whether or not a face uses them.
int gpv_mesh_read_obj(const char* path, gpv_mesh* out)
{
    if (!path || !out) return GPV_ERR_INVALID;
    memset(out, 0, sizeof(*out));
    std::ifstream in(path);
    if (!in.good()) return GPV_ERR_IO;
    std::vector<float> V;
    std::vector<int32_t> F;
    while (in.good()) {
        std::string line;
        std::getline(in, line);
        if (!in.good()) break;
        const auto w1 = ref_split(line, " "), w2 = ref_split(line, "\t");
        const auto& w = w1.size() > w2.size() ? w1 : w2;
        if (w.empty() || w[0].empty()) continue;
        if (w[0][0] == 'v' && w[0].size() == 1) {
            float pt[3] = {0, 0, 0};
            for (size_t i = 1; i < w.size() && i <= 3; i++) pt[i - 1] = strtof(w[i].c_str(), nullptr);
            V.insert(V.end(), pt, pt + 3);
        } else if (w[0][0] == 'f' && w[0].size() == 1) {
            if (w.size() != 4) return GPV_ERR_INVALID;   // triangles only (assert at :475)
            for (int i = 1; i < 4; i++) {
                const long idx = strtol(w[i].c_str(), nullptr, 10);   // "a", "a/b", "a/b/c", "a//c"
                if (idx < 1) return GPV_ERR_INVALID;
                F.push_back((int32_t)(idx - 1));
            }
        }
    }
    for (int32_t i : F)
        if ((int64_t)i >= (int64_t)V.size() / 3) return GPV_ERR_INVALID;
    out->nverts = (int64_t)V.size() / 3;
    out->ntris = (int64_t)F.size() / 3;
    out->verts = (float*)malloc(std::max<size_t>(V.size(), 1) * sizeof(float));
    out->faces = (int32_t*)malloc(std::max<size_t>(F.size(), 1) * sizeof(int32_t));
    if (!out->verts || !out->faces) {
        gpv_mesh_free(out);
        return GPV_ERR_OOM;
    }
    std::copy(V.begin(), V.end(), out->verts);
    std::copy(F.begin(), F.end(), out->faces);
    return GPV_OK;
}

// Object::ReadOFFObject, src/Object.cpp:171-317: header, counts, vertices,
// then "n a b c" per face (0-based); the bbox is taken over the triangles.
int gpv_mesh_read_off(const char* path, gpv_mesh* out)
{
    if (!path || !out) return GPV_ERR_INVALID;
    memset(out, 0, sizeof(*out));
    std::ifstream in(path);
    if (!in.good()) return GPV_ERR_IO;
    std::string header;
    long nv = 0, nf = 0, ne = 0;
    in >> header >> nv >> nf >> ne;
    if (!in.good() || nv <= 0 || nf < 0) return GPV_ERR_INVALID;
    std::vector<float> V((size_t)nv * 3);
    for (long i = 0; i < nv; i++) in >> V[3 * i] >> V[3 * i + 1] >> V[3 * i + 2];
    std::vector<int32_t> F((size_t)nf * 3);
    for (long i = 0; i < nf; i++) {
        int cnt;
        in >> cnt >> F[3 * i] >> F[3 * i + 1] >> F[3 * i + 2];
        for (int k = 0; k < 3; k++)
            if (F[3 * i + k] < 0 || F[3 * i + k] >= nv) return GPV_ERR_INVALID;
    }
    if (in.fail()) return GPV_ERR_INVALID;
    out->nverts = nv;
    out->ntris = nf;
    out->bbox_from_faces = 1;
    out->verts = (float*)malloc(V.size() * sizeof(float));
    out->faces = (int32_t*)malloc(std::max<size_t>(F.size(), 1) * sizeof(int32_t));
    if (!out->verts || !out->faces) {
        gpv_mesh_free(out);
        return GPV_ERR_OOM;
    }
    std::copy(V.begin(), V.end(), out->verts);
    std::copy(F.begin(), F.end(), out->faces);
    return GPV_OK;
}

}  // extern "C"
