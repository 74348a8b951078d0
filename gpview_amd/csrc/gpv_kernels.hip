// gpv_kernels.hip -- gfx950 kernels of the two-level hybrid voxelizer.
//
// Stage map (reference function -> kernels here):
//   index AABB, src/Object.cpp:2272-2318 ........... k_tri_setup
//   Level1 SAT sweep + bins, src/Object.cpp:2256-2342,
//     cuda/CUDAClassifyTessellation.cu:320-401 ...... k_sat_expand
//   Level1 ray parity, src/Object.cpp:716-779 ....... k_ray_expand, k_scatter_hits, k_l1_parity
//   compaction + prefix, src/Object.cpp:3270-3279 ... k_popc_words, scan, k_boundary_index
//   bins by boundary voxel (the CSR of :2137-2219) .. k_pair_rank, scan, k_pair_scatter
//   Level2 SAT, src/Object.cpp:2344-2424 ............ k_l2_sat
//   Level2 ray parity, src/Object.cpp:1111-1207 ..... k_scatter_cols, k_l2_parity
//   counts, src/Object.cpp:3353-3378 ................ k_count_states
//   .raw layout, src/Object.cpp:2934-3075 ........... k_export_*
//
// All launches are 64-wide-wave shaped.  No float atomics: occupancy is set
// with integer OR / ballot only, so results are deterministic.
#include "gpv_device.hpp"
#include "gpv_internal.hpp"

namespace gpv {

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src)
{
    const uint32_t lo = __shfl((unsigned)(v & 0xffffffffu), src);
    const uint32_t hi = __shfl((unsigned)(v >> 32), src);
    return ((uint64_t)hi << 32) | lo;
}

template <typename T>
__device__ __forceinline__ T shfl_up_t(T v, unsigned d);
template <>
__device__ __forceinline__ uint32_t shfl_up_t<uint32_t>(uint32_t v, unsigned d)
{
    return __shfl_up(v, d);
}
template <>
__device__ __forceinline__ uint64_t shfl_up_t<uint64_t>(uint64_t v, unsigned d)
{
    const uint32_t lo = __shfl_up((unsigned)(v & 0xffffffffu), d);
    const uint32_t hi = __shfl_up((unsigned)(v >> 32), d);
    return ((uint64_t)hi << 32) | lo;
}

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T x)
{
    const unsigned lane = __lane_id();
#pragma unroll
    for (unsigned d = 1; d < 64; d <<= 1) {
        const T y = shfl_up_t<T>(x, d);
        if (lane >= d) x += y;
    }
    return x;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t x)
{
    const uint64_t s = wave_incl_scan<uint64_t>(x);
    return shfl64(s, 63);
}

// Wave-aggregated append: one atomic per wave per call, lanes write at
// base + rank among the lanes that append.  Must be reached by every active
// lane of the wave (it may sit inside divergent code).
__device__ __forceinline__ void wave_append(const Append64& a, bool pred, uint64_t key, uint32_t val)
{
    const uint64_t m = __ballot(pred);
    if (m == 0) return;
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint64_t base = 0;
    if ((int)__lane_id() == leader) base = atomicAdd(a.count, (unsigned long long)__popcll(m));
    base = shfl64(base, leader);
    if (pred) {
        const uint64_t idx = base + (uint64_t)__popcll(m & lanes_below_mask());
        if (idx < a.cap) {
            a.key[idx] = key;
            if (a.val) a.val[idx] = val;
        }
    }
}

// largest t in [0, n) with off[t] <= x, given off[0] <= x < off[n]
__device__ __forceinline__ int64_t find_segment(const uint64_t* __restrict__ off, int64_t n, uint64_t x)
{
    int64_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (off[mid] <= x) lo = mid; else hi = mid;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// fills (used instead of memset so that the same code can sit in a graph)
// ---------------------------------------------------------------------------
__global__ void k_fill_u32(uint32_t* p, uint32_t v, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_fill_u64(uint64_t* p, uint64_t v, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

static int grid_for(int64_t n, int block, int cap = 65536)
{
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

void launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s)
{
    if (n <= 0) return;
    k_fill_u32<<<grid_for(n, 256, 8192), 256, 0, s>>>(p, v, n);
}
void launch_fill_u64(uint64_t* p, uint64_t v, int64_t n, hipStream_t s)
{
    if (n <= 0) return;
    k_fill_u64<<<grid_for(n, 256, 8192), 256, 0, s>>>(p, v, n);
}

// ---------------------------------------------------------------------------
// triangle setup: index AABB (src/Object.cpp:2272-2318) and the conservative
// ray-column range used by the all-triangle parity fill
// ---------------------------------------------------------------------------
__device__ __forceinline__ int vertex_index(float v, float mn, float mx, int n)
{
    const float num = v - mn;
    const float den = mx - mn;
    const float r = num / den;
    float s = r * (float)n;
    // identity on [-1, n+1]; only pins down vertices outside the bbox,
    // where the float->int conversion would otherwise be undefined
    s = fminf(fmaxf(s, -1.0f), (float)n + 1.0f);
    int i = (int)s;
    if (i == n && v == mx) i--;
    return i;
}

// first i in [0, n) with tab[i] >= x (n if none)
__device__ __forceinline__ int lower_index(const float* __restrict__ tab, int n, double x)
{
    int a = 0, b = n;
    while (a < b) {
        const int m = (a + b) >> 1;
        if ((double)tab[m] < x) a = m + 1; else b = m;
    }
    return a;
}

// Bound on the columns whose +z ray the float MT test can accept.  With
// U = 2^-24 and |T| bounded over the grid, the computed u, v differ from the
// exact barycentric coordinates of the ray point (w.r.t. the float edges) by
// at most du, dv (rounding of det, of the u/v numerators and of the two
// products by 1/det).  Accepted columns lie in the triangle grown to
// {a >= -du, b >= -dv, a + b <= 1 + 2U + du + dv}; its xy bbox, widened by
// a factor kSafety on du and dv, is returned.
constexpr double kSafety = 4.0;

__device__ bool ray_column_bounds(const float* tv, const MtTri& m, const GridDev& g, double& xlo,
                                  double& xhi, double& ylo, double& yhi)
{
    const double U = 5.9604644775390625e-08;
    const double ad = fabs((double)m.det);
    const double tmx = fmax(fabs((double)g.bminx - tv[0]), fabs((double)g.bmaxx - tv[0]));
    const double tmy = fmax(fabs((double)g.bminy - tv[1]), fabs((double)g.bmaxy - tv[1]));
    const double e1x = m.e1x, e1y = m.e1y, e2x = m.e2x, e2y = m.e2y;
    const double edet = U * (fabs(e1x * e2y) + fabs(e1y * e2x) + ad);
    const double eun = U * (3.0 * (fabs(e2y) * tmx + fabs(e2x) * tmy) + 2.0 * ad);
    const double evn = U * (3.0 * (fabs(e1y) * tmx + fabs(e1x) * tmy) + 2.0 * ad);
    const double den = ad - edet;
    if (den <= 0.5 * ad) {
        xlo = g.bminx; xhi = g.bmaxx; ylo = g.bminy; yhi = g.bmaxy;
        return true;
    }
    const double du = kSafety * ((eun + edet * (1 + 2 * U)) / den + 3 * U);
    const double dv = kSafety * ((evn + edet * (1 + 2 * U)) / den + 3 * U);
    const double ds = du + dv + kSafety * 2 * U;
    const double ca0 = -du, ca1 = 1 + ds + dv, ca2 = -du;
    const double cb0 = -dv, cb1 = -dv, cb2 = 1 + ds + du;
    const double x0 = tv[0] + ca0 * e1x + cb0 * e2x, x1 = tv[0] + ca1 * e1x + cb1 * e2x,
                 x2 = tv[0] + ca2 * e1x + cb2 * e2x;
    const double y0 = tv[1] + ca0 * e1y + cb0 * e2y, y1 = tv[1] + ca1 * e1y + cb1 * e2y,
                 y2 = tv[1] + ca2 * e1y + cb2 * e2y;
    xlo = fmin(x0, fmin(x1, x2)); xhi = fmax(x0, fmax(x1, x2));
    ylo = fmin(y0, fmin(y1, y2)); yhi = fmax(y0, fmax(y1, y2));
    const double sx = 1e-12 * (fabs(xlo) + fabs(xhi) + 1.0), sy = 1e-12 * (fabs(ylo) + fabs(yhi) + 1.0);
    xlo -= sx; xhi += sx; ylo -= sy; yhi += sy;
    return true;
}

__global__ __launch_bounds__(256) void k_tri_setup(const float* __restrict__ tris, int64_t T, GridDev g,
                                                   int4* __restrict__ box0, int4* __restrict__ box1,
                                                   uint32_t* __restrict__ sat_items,
                                                   int4* __restrict__ ray_rng,
                                                   uint32_t* __restrict__ ray_items,
                                                   Counters* __restrict__ ctr)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t tests = 0, est = 0;
    if (t < T) {
        float tv[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) tv[c] = tris[t * 9 + c];
        int lo[3], dd[3];
        const int n[3] = {g.nx, g.ny, g.nz};
        const float mn[3] = {g.bminx, g.bminy, g.bminz};
        const float mx[3] = {g.bmaxx, g.bmaxy, g.bmaxz};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int i0 = vertex_index(tv[a], mn[a], mx[a], n[a]);
            const int i1 = vertex_index(tv[3 + a], mn[a], mx[a], n[a]);
            const int i2 = vertex_index(tv[6 + a], mn[a], mx[a], n[a]);
            int l = min(i0, min(i1, i2));
            int h = max(i0, max(i1, i2));
            l = max(l, 0);
            h = min(h, n[a] - 1);
            lo[a] = l;
            dd[a] = h >= l ? h - l + 1 : 0;
        }
        MtTri m;
        const bool vis = mt_setup(tv, m);
        box0[t] = make_int4(lo[0], lo[1], lo[2], vis ? 1 : 0);
        box1[t] = make_int4(dd[0], dd[1], dd[2], 0);
        const bool any = dd[0] > 0 && dd[1] > 0 && dd[2] > 0;
        sat_items[t] = any ? (uint32_t)(dd[0] * dd[1]) : 0u;
        tests = any ? (uint64_t)dd[0] * dd[1] * dd[2] : 0;
        est = any ? (uint64_t)dd[0] * dd[1] + (uint64_t)dd[1] * dd[2] + (uint64_t)dd[0] * dd[2] : 0;
        uint32_t rit = 0;
        int4 rr = make_int4(0, 0, 0, 0);
        if (vis) {
            double xlo, xhi, ylo, yhi;
            ray_column_bounds(tv, m, g, xlo, xhi, ylo, yhi);
            const int ia = lower_index(g.r1x, g.nx, xlo);
            int ib = lower_index(g.r1x, g.nx, xhi);
            while (ib < g.nx && (double)g.r1x[ib] <= xhi) ++ib;
            const int ja = lower_index(g.r1y, g.ny, ylo);
            int jb = lower_index(g.r1y, g.ny, yhi);
            while (jb < g.ny && (double)g.r1y[jb] <= yhi) ++jb;
            if (ib > ia && jb > ja) {
                rr = make_int4(ia, ja, ib - ia, jb - ja);
                rit = (uint32_t)((ib - ia) * (jb - ja));
            }
        }
        ray_rng[t] = rr;
        ray_items[t] = rit;
    }
    tests = wave_sum_u64(tests);
    est = wave_sum_u64(est);
    if (__lane_id() == 0) {
        if (tests) atomicAdd(&ctr->tests_l1, (unsigned long long)tests);
        if (est) atomicAdd(&ctr->pair_est, (unsigned long long)est);
    }
}

void launch_tri_setup(const float* tris, int64_t T, const GridDev& g, int4* box0, int4* box1,
                      uint32_t* sat_items, int4* ray_rng, uint32_t* ray_items, Counters* ctr,
                      hipStream_t s)
{
    if (T <= 0) return;
    k_tri_setup<<<(unsigned)((T + 255) / 256), 256, 0, s>>>(tris, T, g, box0, box1, sat_items, ray_rng,
                                                            ray_items, ctr);
}

// ---------------------------------------------------------------------------
// exclusive scan (reduce -> scan block sums -> apply), 4096 items per block
// ---------------------------------------------------------------------------
constexpr int SCAN_NT = 1024;
constexpr int SCAN_IPT = 4;
constexpr int SCAN_TILE = SCAN_NT * SCAN_IPT;

template <typename TO>
__device__ TO block_excl_scan(TO x, TO* s_wave, TO* s_total)
{
    const int lane = __lane_id(), wid = threadIdx.x >> 6;
    const TO inc = wave_incl_scan<TO>(x);
    if (lane == 63) s_wave[wid] = inc;
    __syncthreads();
    if (wid == 0) {
        const TO w = lane < SCAN_NT / 64 ? s_wave[lane] : (TO)0;
        const TO wi = wave_incl_scan<TO>(w);
        if (lane < SCAN_NT / 64) s_wave[lane] = wi - w;
        if (lane == SCAN_NT / 64 - 1) *s_total = wi;
    }
    __syncthreads();
    const TO r = s_wave[wid] + inc - x;
    return r;
}

template <typename TI, typename TO>
__global__ __launch_bounds__(SCAN_NT) void k_scan_reduce(const TI* __restrict__ in, int64_t n,
                                                         TO* __restrict__ sums,
                                                         unsigned int* __restrict__ max_out)
{
    __shared__ TO s_wave[SCAN_NT / 64];
    __shared__ TO s_total;
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_IPT;
    TO acc = 0;
    TI mx = 0;
#pragma unroll
    for (int k = 0; k < SCAN_IPT; ++k) {
        if (base + k < n) {
            const TI v = in[base + k];
            acc += (TO)v;
            mx = v > mx ? v : mx;
        }
    }
    (void)block_excl_scan<TO>(acc, s_wave, &s_total);
    if (threadIdx.x == 0) sums[blockIdx.x] = s_total;
    if (max_out) {
        unsigned m = (unsigned)mx;
        for (unsigned d = 32; d >= 1; d >>= 1) {
            const unsigned o = __shfl_xor(m, d);
            m = o > m ? o : m;
        }
        if (__lane_id() == 0 && m) atomicMax(max_out, m);
    }
}

template <typename TI, typename TO>
__global__ __launch_bounds__(SCAN_NT) void k_scan_apply(const TI* __restrict__ in, int64_t n,
                                                        const TO* __restrict__ block_prefix,
                                                        TO* __restrict__ out)
{
    __shared__ TO s_wave[SCAN_NT / 64];
    __shared__ TO s_total;
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_IPT;
    TO v[SCAN_IPT];
    TO acc = 0;
#pragma unroll
    for (int k = 0; k < SCAN_IPT; ++k) {
        v[k] = (base + k < n) ? (TO)in[base + k] : (TO)0;
        acc += v[k];
    }
    TO run = block_excl_scan<TO>(acc, s_wave, &s_total) + block_prefix[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_IPT; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
        if (base + k == n - 1) out[n] = run;
    }
}

// scan_rec(n) uses sums[nb+1], then (nb > 1) rest[nb+1] and scan_rec(nb)'s own
size_t scan_tmp_elems(int64_t n)
{
    if (n <= 0) return 2;
    const int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nb == 1) return 2;
    return (size_t)(nb + 1) * 2 + scan_tmp_elems(nb);
}

template <typename TI, typename TO>
static void scan_rec(const TI* in, TO* out, int64_t n, TO* tmp, unsigned int* max_out, hipStream_t s)
{
    if (n <= 0) {
        if constexpr (sizeof(TO) == 8) launch_fill_u64(reinterpret_cast<uint64_t*>(out), 0, 1, s);
        else launch_fill_u32(reinterpret_cast<uint32_t*>(out), 0, 1, s);
        return;
    }
    const int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    TO* sums = tmp;            // nb + 1
    TO* rest = tmp + nb + 1;
    k_scan_reduce<TI, TO><<<(unsigned)nb, SCAN_NT, 0, s>>>(in, n, sums, max_out);
    if (nb == 1) {
        // block prefix of the single block is 0
        if constexpr (sizeof(TO) == 8) launch_fill_u64(reinterpret_cast<uint64_t*>(sums), 0, 1, s);
        else launch_fill_u32(reinterpret_cast<uint32_t*>(sums), 0, 1, s);
        k_scan_apply<TI, TO><<<1, SCAN_NT, 0, s>>>(in, n, sums, out);
        return;
    }
    // scan the block sums in place (exclusive) into rest, then apply
    scan_rec<TO, TO>(sums, rest, nb, rest + nb + 1, nullptr, s);
    k_scan_apply<TI, TO><<<(unsigned)nb, SCAN_NT, 0, s>>>(in, n, rest, out);
}

void launch_scan_u32_u64(const uint32_t* in, uint64_t* out, int64_t n, uint64_t* tmp, hipStream_t s)
{
    scan_rec<uint32_t, uint64_t>(in, out, n, tmp, nullptr, s);
}

void launch_scan_u32_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* tmp,
                         unsigned int* max_out, hipStream_t s)
{
    scan_rec<uint32_t, uint32_t>(in, out, n, tmp, max_out, s);
}

// ---------------------------------------------------------------------------
// item expansion with LDS staging: a block takes XB consecutive items of the
// flattened (triangle, sub-range) space; the triangles those items belong to
// are staged in LDS (offsets + vertices) with one coalesced copy, and each
// lane finds its triangle by a binary search in LDS.  Falls back to global
// memory when a block spans more than XCAP triangles (long runs of empty
// triangles).
// ---------------------------------------------------------------------------
constexpr int XB = 256;
constexpr int XCAP = 512;

template <typename F>
__device__ __forceinline__ void expand_items(const float* __restrict__ tris, int64_t T,
                                             const uint64_t* __restrict__ off, F&& body)
{
    __shared__ uint64_t s_off[XCAP + 1];
    __shared__ float s_tv[XCAP * 9];
    __shared__ int64_t s_t[2];
    const uint64_t total = off[T];
    const int tid = threadIdx.x;
    for (uint64_t base = (uint64_t)blockIdx.x * XB; base < total; base += (uint64_t)gridDim.x * XB) {
        const uint64_t last = (base + XB < total ? base + XB : total) - 1;
        if (tid == 0) s_t[0] = find_segment(off, T, base);
        if (tid == 64) s_t[1] = find_segment(off, T, last);
        __syncthreads();
        const int64_t t0 = s_t[0];
        const int64_t nt = s_t[1] - t0 + 1;
        const bool staged = nt <= XCAP;
        if (staged) {
            for (int k = tid; k <= nt; k += XB) s_off[k] = off[t0 + k];
            const float* src = tris + t0 * 9;
            for (int k = tid; k < nt * 9; k += XB) s_tv[k] = src[k];
        }
        __syncthreads();
        const uint64_t item = base + tid;
        if (item < total) {
            int64_t t;
            uint64_t l;
            float tv[9];
            if (staged) {
                int a = 0, b = (int)nt;
                while (b - a > 1) {
                    const int m = (a + b) >> 1;
                    if (s_off[m] <= item) a = m; else b = m;
                }
                t = t0 + a;
                l = item - s_off[a];
#pragma unroll
                for (int c = 0; c < 9; ++c) tv[c] = s_tv[a * 9 + c];
            } else {
                t = find_segment(off, T, item);
                l = item - off[t];
#pragma unroll
                for (int c = 0; c < 9; ++c) tv[c] = tris[t * 9 + c];
            }
            body(t, l, tv);
        }
        __syncthreads();
    }
}

// Level1 SAT: one item = (triangle, xy column of its index-AABB); the lane
// walks the AABB's z range.  Hits set the boundary bit, append the
// (voxel, triangle) bin pair, and the lane's highest hit z becomes the
// (column, triangle, zmax) entry used by the Level2 parity candidates.
__global__ __launch_bounds__(XB) void k_sat_expand(const float* __restrict__ tris, int64_t T,
                                                   const int4* __restrict__ box0,
                                                   const int4* __restrict__ box1,
                                                   const uint64_t* __restrict__ off, GridDev g,
                                                   unsigned long long* __restrict__ bnd, Append64 pairs,
                                                   Append64 cols)
{
    expand_items(tris, T, off, [&](int64_t t, uint64_t l, const float* tv) {
        const int4 b0 = box0[t];
        const int4 b1 = box1[t];
        const int p = b0.x + (int)(l % (uint64_t)b1.x);
        const int q = b0.y + (int)(l / (uint64_t)b1.x);
        const float cx = g.c1x[p], cy = g.c1y[q];
        const uint64_t rowword = (uint64_t)q * g.nxw + (uint64_t)(p >> 6);
        const uint64_t bit = 1ull << (p & 63);
        int zmax = -1;
        for (int z = b0.z; z < b0.z + b1.z; ++z) {
            const bool hit = tri_box_overlap(cx, cy, g.c1z[z], g.h1x, g.h1y, g.h1z, tv);
            const uint64_t w = (uint64_t)z * g.ny * g.nxw + rowword;
            if (hit) {
                zmax = z;
                atomicOr(&bnd[w], (unsigned long long)bit);
            }
            wave_append(pairs, hit, ((w * 64 + (uint64_t)(p & 63)) << 32) | (uint64_t)t, 0);
        }
        const bool emit = (zmax >= 0) && (b0.w != 0);
        wave_append(cols, emit, ((uint64_t)((uint32_t)q * g.nx + p) << 32) | (uint64_t)t, (uint32_t)zmax);
    });
}

void launch_sat_expand(const float* tris, int64_t T, const int4* box0, const int4* box1,
                       const uint64_t* off, const GridDev& g, uint64_t* bnd, Append64 pairs,
                       Append64 cols, int blocks, hipStream_t s)
{
    if (T <= 0) return;
    k_sat_expand<<<blocks, XB, 0, s>>>(tris, T, box0, box1, off, g,
                                       reinterpret_cast<unsigned long long*>(bnd), pairs, cols);
}

// Level1 ray parity, first half: one item = (triangle, column inside its
// conservative bound); the exact det/u/v test keeps the (column, triangle)
// pairs whose ray the reference's MT accepts at some height.
__global__ __launch_bounds__(XB) void k_ray_expand(const float* __restrict__ tris, int64_t T,
                                                   const int4* __restrict__ rng,
                                                   const uint64_t* __restrict__ off, GridDev g,
                                                   Append64 hits)
{
    expand_items(tris, T, off, [&](int64_t t, uint64_t l, const float* tv) {
        const int4 r = rng[t];
        const int i = r.x + (int)(l % (uint64_t)r.z);
        const int j = r.y + (int)(l / (uint64_t)r.z);
        MtTri m;
        bool ok = mt_setup(tv, m);
        float tx, ty;
        ok = ok && mt_column(m, tv[0], tv[1], g.r1x[i], g.r1y[j], tx, ty);
        wave_append(hits, ok, ((uint64_t)((uint32_t)j * g.nx + i) << 32) | (uint64_t)t, 0);
    });
}

void launch_ray_expand(const float* tris, int64_t T, const int4* ray_rng, const uint64_t* off,
                       const GridDev& g, Append64 hits, int blocks, hipStream_t s)
{
    if (T <= 0) return;
    k_ray_expand<<<blocks, XB, 0, s>>>(tris, T, ray_rng, off, g, hits);
}

// ---------------------------------------------------------------------------
// counting sorts by the high 32 bits of a key
// ---------------------------------------------------------------------------
__global__ void k_count_keys_hi(const uint64_t* __restrict__ key, uint64_t n, uint32_t* __restrict__ cnt)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        atomicAdd(&cnt[key[i] >> 32], 1u);
}

void launch_count_keys_hi(const uint64_t* key, uint64_t n,
                          uint32_t* cnt, hipStream_t s)
{
    if (!n) return;
    k_count_keys_hi<<<grid_for((int64_t)n, 256, 16384), 256, 0, s>>>(key, n, cnt);
}

__global__ void k_scatter_hits(const float* __restrict__ tris, const uint64_t* __restrict__ key, uint64_t n,
                               GridDev g, const uint32_t* __restrict__ off, uint32_t* __restrict__ cnt,
                               float4* __restrict__ rec)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t k = key[i];
        const uint32_t col = (uint32_t)(k >> 32);
        const uint32_t t = (uint32_t)(k & 0xffffffffu);
        const uint32_t pos = off[col] + atomicSub(&cnt[col], 1u) - 1u;
        float tv[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) tv[c] = tris[(uint64_t)t * 9 + c];
        MtTri m;
        mt_setup(tv, m);
        const int i1 = (int)(col % (uint32_t)g.nx), j1 = (int)(col / (uint32_t)g.nx);
        const float tx = g.r1x[i1] - tv[0];
        const float ty = g.r1y[j1] - tv[1];
        rec[3 * (uint64_t)pos + 0] = make_float4(m.e1x, m.e1y, m.e1z, m.e2x);
        rec[3 * (uint64_t)pos + 1] = make_float4(m.e2y, m.e2z, m.inv_det, tx);
        rec[3 * (uint64_t)pos + 2] = make_float4(ty, tv[2], 0.0f, 0.0f);
    }
}

void launch_scatter_hits(const float* tris, const uint64_t* key, uint64_t n,
                          const GridDev& g, const uint32_t* off, uint32_t* cnt,
                         float4* rec, hipStream_t s)
{
    if (!n) return;
    k_scatter_hits<<<grid_for((int64_t)n, 256, 16384), 256, 0, s>>>(tris, key, n, g, off, cnt, rec);
}

__global__ void k_scatter_cols(const uint64_t* __restrict__ key, const uint32_t* __restrict__ zmax,
                               uint64_t n, const uint32_t* __restrict__ off, uint32_t* __restrict__ cnt,
                               uint32_t* __restrict__ ct, uint32_t* __restrict__ cz)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t k = key[i];
        const uint32_t col = (uint32_t)(k >> 32);
        const uint32_t pos = off[col] + atomicSub(&cnt[col], 1u) - 1u;
        ct[pos] = (uint32_t)(k & 0xffffffffu);
        cz[pos] = zmax[i];
    }
}

void launch_scatter_cols(const uint64_t* key, const uint32_t* zmax, uint64_t n,
                          const uint32_t* off, uint32_t* cnt, uint32_t* ct, uint32_t* cz,
                         hipStream_t s)
{
    if (!n) return;
    k_scatter_cols<<<grid_for((int64_t)n, 256, 16384), 256, 0, s>>>(key, zmax, n, off, cnt, ct, cz);
}

// ---------------------------------------------------------------------------
// Level1 parity: one wave per inside-bitplane word (64 x-adjacent voxels of
// one row); each lane XORs t > 1e-6 over its column's accepted triangles.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_l1_parity(const float4* __restrict__ rec,
                                                   const uint32_t* __restrict__ hoff, GridDev g,
                                                   uint64_t* __restrict__ ins, int64_t W)
{
    const int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (w >= W) return;
    const int lane = __lane_id();
    const int xw = (int)(w % g.nxw);
    const int64_t row = w / g.nxw;
    const int j = (int)(row % g.ny);
    const int k = (int)(row / g.ny);
    const int x = xw * 64 + lane;
    int par = 0;
    if (x < g.nx) {
        const float oz = g.r1z[k];
        const uint32_t c = (uint32_t)j * g.nx + x;
        const uint32_t h1 = hoff[c + 1];
        for (uint32_t h = hoff[c]; h < h1; ++h) {
            const float4 a = rec[3 * (uint64_t)h + 0];
            const float4 b = rec[3 * (uint64_t)h + 1];
            const float4 d = rec[3 * (uint64_t)h + 2];
            par ^= (int)mt_t_hit(a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, d.x, d.y, oz);
        }
    }
    const uint64_t m = __ballot(par);
    if (lane == 0) ins[w] = m;
}

void launch_l1_parity(const float4* rec, const uint32_t* hoff, const GridDev& g, uint64_t* ins,
                      hipStream_t s)
{
    const int64_t W = (int64_t)g.nz * g.ny * g.nxw;
    if (W <= 0) return;
    const int64_t blocks = (W * 64 + 255) / 256;
    k_l1_parity<<<(unsigned)blocks, 256, 0, s>>>(rec, hoff, g, ins, W);
}

// ---------------------------------------------------------------------------
// boundary compaction (src/Object.cpp:3270-3279): popcount per word ->
// exclusive scan -> ascending linear indices of the boundary voxels
// ---------------------------------------------------------------------------
__global__ void k_popc_words(const uint64_t* __restrict__ bits, int64_t W, uint32_t* __restrict__ cnt)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < W;
         i += (int64_t)gridDim.x * blockDim.x)
        cnt[i] = (uint32_t)__popcll(bits[i]);
}

void launch_popc_words(const uint64_t* bits, int64_t W, uint32_t* cnt, hipStream_t s)
{
    if (W <= 0) return;
    k_popc_words<<<grid_for(W, 256, 16384), 256, 0, s>>>(bits, W, cnt);
}

__global__ void k_boundary_index(const uint64_t* __restrict__ bnd, const uint32_t* __restrict__ wprefix,
                                 GridDev g, int64_t W, int32_t* __restrict__ bidx)
{
    for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < W;
         w += (int64_t)gridDim.x * blockDim.x) {
        uint64_t bits = bnd[w];
        if (!bits) continue;
        uint32_t pos = wprefix[w];
        const int xw = (int)(w % g.nxw);
        const int64_t row = w / g.nxw;
        while (bits) {
            const int bit = __ffsll((unsigned long long)bits) - 1;
            bits &= bits - 1;
            bidx[pos++] = (int32_t)(row * g.nx + xw * 64 + bit);
        }
    }
}

void launch_boundary_index(const uint64_t* bnd, const uint32_t* wprefix, const GridDev& g,
                           int32_t* bidx, hipStream_t s)
{
    const int64_t W = (int64_t)g.nz * g.ny * g.nxw;
    if (W <= 0) return;
    k_boundary_index<<<grid_for(W, 256, 16384), 256, 0, s>>>(bnd, wprefix, g, W, bidx);
}

__device__ __forceinline__ uint32_t rank_of(const uint64_t* __restrict__ bnd,
                                            const uint32_t* __restrict__ wprefix, uint64_t bitidx)
{
    const uint64_t w = bitidx >> 6;
    const int bit = (int)(bitidx & 63);
    const uint64_t below = bit ? (~0ull >> (64 - bit)) : 0ull;
    return wprefix[w] + (uint32_t)__popcll(bnd[w] & below);
}

__global__ void k_pair_rank(const uint64_t* __restrict__ pairs, uint64_t n, const uint64_t* __restrict__ bnd,
                            const uint32_t* __restrict__ wprefix, uint32_t* __restrict__ pb,
                            uint32_t* __restrict__ cntb)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t b = rank_of(bnd, wprefix, pairs[i] >> 32);
        pb[i] = b;
        atomicAdd(&cntb[b], 1u);
    }
}

void launch_pair_rank(const uint64_t* pairs, uint64_t n,
                      const uint64_t* bnd, const uint32_t* wprefix, uint32_t* pb, uint32_t* cntb,
                      hipStream_t s)
{
    if (!n) return;
    k_pair_rank<<<grid_for((int64_t)n, 256, 16384), 256, 0, s>>>(pairs, n, bnd, wprefix, pb, cntb);
}

__global__ void k_pair_scatter(const uint64_t* __restrict__ pairs, const uint32_t* __restrict__ pb, uint64_t n,
                               const uint32_t* __restrict__ boff, uint32_t* __restrict__ cntb,
                               uint32_t* __restrict__ list)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t b = pb[i];
        const uint32_t pos = boff[b] + atomicSub(&cntb[b], 1u) - 1u;
        list[pos] = (uint32_t)(pairs[i] & 0xffffffffu);
    }
}

void launch_pair_scatter(const uint64_t* pairs, const uint32_t* pb, uint64_t n,
                          const uint32_t* boff, uint32_t* cntb, uint32_t* list,
                         hipStream_t s)
{
    if (!n) return;
    k_pair_scatter<<<grid_for((int64_t)n, 256, 16384), 256, 0, s>>>(pairs, pb, n, boff, cntb, list);
}

// ---------------------------------------------------------------------------
// Level2 SAT (src/Object.cpp:2344-2424): one wave per boundary voxel.  For
// each triangle of the parent's bin the wave first evaluates the box-axis
// predicate of the SAT on every sub-index of each axis (the same float test
// the full SAT runs, so sub-voxels failing it are misses), then spreads the
// surviving sub-box over its lanes.  Bits accumulate in LDS.
// ---------------------------------------------------------------------------
constexpr int L2_WAVES = 4;

__global__ __launch_bounds__(64 * L2_WAVES) void k_l2_sat(const float* __restrict__ tris,
                                                          const int32_t* __restrict__ bidx, int64_t B,
                                                          const uint32_t* __restrict__ boff,
                                                          const uint32_t* __restrict__ list, GridDev g,
                                                          uint64_t* __restrict__ sat2)
{
    __shared__ unsigned long long s_acc[L2_WAVES][64];
    const int wv = threadIdx.x >> 6;
    const int lane = __lane_id();
    const int64_t b = (int64_t)blockIdx.x * L2_WAVES + wv;
    if (b >= B) return;
    unsigned long long* acc = s_acc[wv];
    for (int i = lane; i < g.w2; i += 64) acc[i] = 0ull;
    __builtin_amdgcn_wave_barrier();
    const int v = bidx[b];
    const int nxy = g.nx * g.ny;
    const int k1 = v / nxy;
    const int ij = v - k1 * nxy;
    const int j1 = ij / g.nx;
    const int i1 = ij - j1 * g.nx;
    const int n2 = g.n2;
    const float* cx = g.c2x + (int64_t)i1 * n2;
    const float* cy = g.c2y + (int64_t)j1 * n2;
    const float* cz = g.c2z + (int64_t)k1 * n2;
    const float ax = lane < n2 ? cx[lane] : 0.0f;
    const float ay = lane < n2 ? cy[lane] : 0.0f;
    const float az = lane < n2 ? cz[lane] : 0.0f;
    const uint32_t e1 = boff[b + 1];
    for (uint32_t e = boff[b]; e < e1; ++e) {
        const uint64_t t = list[e];
        float tv[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) tv[c] = tris[t * 9 + c];
        const uint64_t mx = __ballot(lane < n2 && sat_axis_ok(tv[0], tv[3], tv[6], ax, g.h2x));
        const uint64_t my = __ballot(lane < n2 && sat_axis_ok(tv[1], tv[4], tv[7], ay, g.h2y));
        const uint64_t mz = __ballot(lane < n2 && sat_axis_ok(tv[2], tv[5], tv[8], az, g.h2z));
        if (!mx || !my || !mz) continue;
        const int pa = __ffsll((unsigned long long)mx) - 1, pb = 63 - __clzll((long long)mx);
        const int qa = __ffsll((unsigned long long)my) - 1, qb = 63 - __clzll((long long)my);
        const int ra = __ffsll((unsigned long long)mz) - 1, rb = 63 - __clzll((long long)mz);
        const int dp = pb - pa + 1, dq = qb - qa + 1, dr = rb - ra + 1;
        const int V = dp * dq * dr;
        for (int l = lane; l < V; l += 64) {
            const int p = pa + l % dp;
            const int rem = l / dp;
            const int q = qa + rem % dq;
            const int r = ra + rem / dq;
            const int s = (r * n2 + q) * n2 + p;
            if ((acc[s >> 6] >> (s & 63)) & 1ull) continue;   // already boundary
            if (tri_box_overlap(cx[p], cy[q], cz[r], g.h2x, g.h2y, g.h2z, tv))
                atomicOr(&acc[s >> 6], 1ull << (s & 63));
        }
    }
    __builtin_amdgcn_wave_barrier();
    for (int i = lane; i < g.w2; i += 64) sat2[b * g.w2 + i] = acc[i];
}

void launch_l2_sat(const float* tris, const int32_t* bidx, int64_t B, const uint32_t* boff,
                   const uint32_t* list, const GridDev& g, uint64_t* sat2, hipStream_t s)
{
    if (B <= 0) return;
    const int64_t blocks = (B + L2_WAVES - 1) / L2_WAVES;
    k_l2_sat<<<(unsigned)blocks, 64 * L2_WAVES, 0, s>>>(tris, bidx, B, boff, list, g, sat2);
}

// ---------------------------------------------------------------------------
// Level2 ray parity (src/Object.cpp:1111-1207): one wave per boundary voxel
// (i1, j1, k1).  Candidates are the column's (triangle, zmax) entries with
// zmax >= k1 -- the deduplicated union of the bins at z' >= k1.  Lanes own
// sub-columns (p, q): the det/u/v part runs once per sub-column, t per r.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64 * L2_WAVES) void k_l2_parity(const float* __restrict__ tris,
                                                             const int32_t* __restrict__ bidx, int64_t B,
                                                             const uint32_t* __restrict__ coff,
                                                             const uint32_t* __restrict__ ct,
                                                             const uint32_t* __restrict__ cz, GridDev g,
                                                             uint64_t* __restrict__ par2)
{
    __shared__ unsigned long long s_acc[L2_WAVES][64];
    const int wv = threadIdx.x >> 6;
    const int lane = __lane_id();
    const int64_t b = (int64_t)blockIdx.x * L2_WAVES + wv;
    if (b >= B) return;
    unsigned long long* acc = s_acc[wv];
    const int v = bidx[b];
    const int nxy = g.nx * g.ny;
    const int k1 = v / nxy;
    const int ij = v - k1 * nxy;
    const int j1 = ij / g.nx;
    const int i1 = ij - j1 * g.nx;
    const int n2 = g.n2;
    const int nsc = n2 * n2;
    const int nm = (nsc + 63) >> 6;
    const float* rx = g.r2x + (int64_t)i1 * n2;
    const float* ry = g.r2y + (int64_t)j1 * n2;
    const float* rz = g.r2z + (int64_t)k1 * n2;
    uint64_t bits = 0;   // bit (m*n2 + r) for sub-column lane + 64*m
    const uint32_t col = (uint32_t)ij;
    const uint32_t e1 = coff[col + 1];
    for (uint32_t e = coff[col]; e < e1; ++e) {
        if ((int)cz[e] < k1) continue;
        const uint64_t t = ct[e];
        float tv[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) tv[c] = tris[t * 9 + c];
        MtTri m;
        if (!mt_setup(tv, m)) continue;
        for (int mm = 0; mm < nm; ++mm) {
            const int sc = lane + 64 * mm;
            if (sc < nsc) {
                const int p = sc % n2, q = sc / n2;
                float tx, ty;
                if (mt_column(m, tv[0], tv[1], rx[p], ry[q], tx, ty)) {
                    for (int r = 0; r < n2; ++r)
                        if (mt_t_hit(m.e1x, m.e1y, m.e1z, m.e2x, m.e2y, m.e2z, m.inv_det, tx, ty, tv[2], rz[r]))
                            bits ^= 1ull << (mm * n2 + r);
                }
            }
        }
    }
    for (int i = lane; i < g.w2; i += 64) acc[i] = 0ull;
    __builtin_amdgcn_wave_barrier();
    for (int mm = 0; mm < nm; ++mm) {
        const int sc = lane + 64 * mm;
        if (sc >= nsc) break;
        for (int r = 0; r < n2; ++r) {
            if ((bits >> (mm * n2 + r)) & 1ull) {
                const int s = r * nsc + sc;
                atomicOr(&acc[s >> 6], 1ull << (s & 63));
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
    for (int i = lane; i < g.w2; i += 64) par2[b * g.w2 + i] = acc[i];
}

void launch_l2_parity(const float* tris, const int32_t* bidx, int64_t B, const uint32_t* coff,
                      const uint32_t* ct, const uint32_t* cz, const GridDev& g, uint64_t* par2,
                      hipStream_t s)
{
    if (B <= 0) return;
    const int64_t blocks = (B + L2_WAVES - 1) / L2_WAVES;
    k_l2_parity<<<(unsigned)blocks, 64 * L2_WAVES, 0, s>>>(tris, bidx, B, coff, ct, cz, g, par2);
}

// ---------------------------------------------------------------------------
// counts (src/Object.cpp:3353-3378)
// ---------------------------------------------------------------------------
__global__ void k_count_states(const uint64_t* __restrict__ bnd, const uint64_t* __restrict__ ins, int64_t W,
                               const uint64_t* __restrict__ sat2, const uint64_t* __restrict__ par2,
                               int64_t W2, Counters* __restrict__ ctr)
{
    uint64_t a = 0, b = 0, c = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < W; i += (int64_t)gridDim.x * blockDim.x)
        a += (uint64_t)__popcll(ins[i] & ~bnd[i]);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < W2; i += (int64_t)gridDim.x * blockDim.x) {
        b += (uint64_t)__popcll(par2[i] & ~sat2[i]);
        c += (uint64_t)__popcll(sat2[i]);
    }
    a = wave_sum_u64(a);
    b = wave_sum_u64(b);
    c = wave_sum_u64(c);
    if (__lane_id() == 0) {
        if (a) atomicAdd(&ctr->l1_inside, (unsigned long long)a);
        if (b) atomicAdd(&ctr->l2_inside, (unsigned long long)b);
        if (c) atomicAdd(&ctr->l2_boundary, (unsigned long long)c);
    }
}

void launch_count_states(const uint64_t* bnd, const uint64_t* ins, int64_t W, const uint64_t* sat2,
                         const uint64_t* par2, int64_t W2, Counters* ctr, hipStream_t s)
{
    const int64_t n = W > W2 ? W : W2;
    if (n <= 0) return;
    k_count_states<<<grid_for(n, 256, 4096), 256, 0, s>>>(bnd, ins, W, sat2, par2, W2, ctr);
}

// ---------------------------------------------------------------------------
// egress: one byte per voxel holding state * unit (unit 1: the 0/1/2 states;
// unit 127: the .raw bytes uint8(state * 127.0f), src/Object.cpp:3024-3063)
// and the int32 exclusive boundary prefix per Level1 voxel
// ---------------------------------------------------------------------------
__global__ void k_export_l1(const uint64_t* __restrict__ bnd, const uint64_t* __restrict__ ins, GridDev g,
                            int64_t n1, uint8_t unit, uint8_t* __restrict__ out)
{
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n1; v += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(v % g.nx);
        const int64_t row = v / g.nx;
        const int64_t w = row * g.nxw + (x >> 6);
        const uint64_t bit = 1ull << (x & 63);
        out[v] = (bnd[w] & bit) ? (uint8_t)(2 * unit) : ((ins[w] & bit) ? unit : (uint8_t)0);
    }
}

void launch_export_l1(const uint64_t* bnd, const uint64_t* ins, const GridDev& g, uint8_t unit, uint8_t* out,
                      hipStream_t s)
{
    const int64_t n1 = (int64_t)g.nx * g.ny * g.nz;
    if (n1 <= 0) return;
    k_export_l1<<<grid_for(n1, 256, 65536), 256, 0, s>>>(bnd, ins, g, n1, unit, out);
}

__global__ void k_export_prefix(const uint64_t* __restrict__ bnd, const uint32_t* __restrict__ wprefix, GridDev g,
                                int64_t n1, int32_t* __restrict__ out)
{
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n1; v += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(v % g.nx);
        const int64_t row = v / g.nx;
        const uint64_t bitidx = (uint64_t)(row * g.nxw + (x >> 6)) * 64 + (uint64_t)(x & 63);
        out[v] = (int32_t)rank_of(bnd, wprefix, bitidx);
    }
}

void launch_export_prefix(const uint64_t* bnd, const uint32_t* wprefix, const GridDev& g, int32_t* out,
                          hipStream_t s)
{
    const int64_t n1 = (int64_t)g.nx * g.ny * g.nz;
    if (n1 <= 0) return;
    k_export_prefix<<<grid_for(n1, 256, 65536), 256, 0, s>>>(bnd, wprefix, g, n1, out);
}

__global__ void k_export_l2(const uint64_t* __restrict__ sat2, const uint64_t* __restrict__ par2, int64_t B,
                            int n23, int w2, uint8_t unit, uint8_t* __restrict__ out)
{
    const int64_t n = B * n23;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = i / n23;
        const int s = (int)(i - b * n23);
        const int64_t w = b * w2 + (s >> 6);
        const uint64_t bit = 1ull << (s & 63);
        out[i] = (sat2[w] & bit) ? (uint8_t)(2 * unit) : ((par2[w] & bit) ? unit : (uint8_t)0);
    }
}

void launch_export_l2(const uint64_t* sat2, const uint64_t* par2, int64_t B, const GridDev& g, uint8_t unit,
                      uint8_t* out, hipStream_t s)
{
    const int n23 = g.n2 * g.n2 * g.n2;
    if (B <= 0 || n23 <= 0) return;
    k_export_l2<<<grid_for(B * n23, 256, 65536), 256, 0, s>>>(sat2, par2, B, n23, g.w2, unit, out);
}

}  // namespace gpv
