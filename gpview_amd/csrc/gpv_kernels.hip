// gpv_kernels.hip -- gfx950 kernels of the two-level hybrid voxelizer.
//
// Stage map (reference function -> kernels here):
//   index AABB, src/Object.cpp:2272-2318 ........... k_tri_setup
//   Level1 SAT sweep + bins, src/Object.cpp:2256-2342,
//     cuda/CUDAClassifyTessellation.cu:320-401 ...... k_sat_expand
//   Level1 ray parity, src/Object.cpp:716-779 ....... k_ray_expand, k_scatter_hits, k_l1_parity
//   compaction + prefix, src/Object.cpp:3270-3279 ... k_popc_words, scan, k_boundary_index
//   bins by boundary voxel (the CSR of :2137-2219) .. k_ts_count, k_ts_bucket, k_ts_place
//   Level2 SAT, src/Object.cpp:2344-2424 ............ k_l2_sat
//   Level2 ray parity, src/Object.cpp:1111-1207 ..... k_ts_* (column entries), k_col_pass, k_l2_parity
//   counts, src/Object.cpp:3353-3378 ................ k_count_states
//   .raw layout, src/Object.cpp:2934-3075 ........... k_export_*
//
// All launches are 64-wide-wave shaped.  No float atomics: occupancy is set
// with integer OR / ballot only, so results are deterministic.
#include "gpv_device.hpp"
#include "gpv_internal.hpp"
#include "gpv_slab.hpp"

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

// Block-wide sum folded into *dst with one atomic per block (blockDim <= 1024,
// every thread of the block must call it).  Same-address atomics serialise in
// L2, so kernels reduce per block before touching a shared counter.
__device__ __forceinline__ void block_add_u64(uint64_t x, unsigned long long* dst)
{
    __shared__ uint64_t s_part[16];
    x = wave_sum_u64(x);
    const int wid = threadIdx.x >> 6;
    __syncthreads();
    if (__lane_id() == 0) s_part[wid] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (int i = 0; i < (int)((blockDim.x + 63) >> 6); ++i) t += s_part[i];
        if (t) atomicAdd(dst, (unsigned long long)t);
    }
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
// multi-mesh lookups.  Every kernel that depends on the grid is a template on
// kM: kM = false is the single-mesh form (grid passed by value, no offsets),
// kM = true finds the mesh of its work item in the batch descriptors.
// ---------------------------------------------------------------------------
template <bool kM>
__device__ __forceinline__ const GridDev& grid_of(const GridDev& g0, const Meshes& M, int m)
{
    if constexpr (kM) return M.grids[m];
    else return g0;
}

// mesh owning batch-wide Level1 word w: largest m with off[m].word0 <= w
__device__ __forceinline__ int mesh_of_word(const MeshOff* __restrict__ off, int nm, uint64_t w)
{
    int lo = 0, hi = nm;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid].word0 <= w) lo = mid; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ int uniform_i(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Largest m in [0, nm) with key(off[m]) <= x for a wave-uniform x (given
// key(off[0]) <= x), by the whole wave: 64 meshes per round of independent
// loads, so two rounds of latency up to 4096 meshes instead of log2(nm).
template <typename K>
__device__ __forceinline__ int wave_mesh_search(const MeshOff* __restrict__ off, int nm, uint64_t x, K key)
{
    const int lane = __lane_id();
    int lo = 0, hi = nm;   // key(off[lo]) <= x; the answer lies in [lo, hi)
    while (hi - lo > 1) {
        const int step = (hi - lo + 63) / 64;
        const int idx = lo + lane * step;
        const bool le = idx < hi && (uint64_t)key(off[idx]) <= x;
        const uint64_t m = __ballot(le);
        const int L = 63 - __clzll((long long)m);
        const int nlo = lo + L * step;
        hi = nlo + step < hi ? nlo + step : hi;
        lo = nlo;
    }
    return uniform_i(lo);
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

// mesh of every triangle of the concatenated soup
__global__ void k_mesh_ids(const MeshOff* __restrict__ off, int nm, int64_t T, uint32_t* __restrict__ tri_mesh)
{
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < T; t += (int64_t)gridDim.x * blockDim.x) {
        int lo = 0, hi = nm;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if ((int64_t)off[mid].tri0 <= t) lo = mid; else hi = mid;
        }
        tri_mesh[t] = (uint32_t)lo;
    }
}

void launch_mesh_ids(const MeshOff* off, int nm, int64_t T, uint32_t* tri_mesh, hipStream_t s)
{
    if (T <= 0) return;
    k_mesh_ids<<<grid_for(T, 256, 8192), 256, 0, s>>>(off, nm, T, tri_mesh);
}

// Single-mesh form: grid-stride over all triangles, one block reduction into
// the mesh's test counter.  Multi-mesh form: one triangle per thread; blocks
// inside one mesh (tri_mesh is non-decreasing) reduce into that mesh's
// counter, the few straddling blocks add per lane.
// Triangles whose index AABB spans at least kZSlabMin Level1 layers also get a
// plane slab (gpv_slab.hpp) over the centers of that AABB: the Level1 sweep
// then visits, per column, only the layers the slab can keep.
constexpr int kZSlabMin = 3;

// kPart: 1 the SAT half (index AABB, items, slabs, test counts), 2 the ray
// half (column bound, items), 3 both.  The halves share nothing but the
// vertices, so they can run on two streams.
template <bool kM, int kPart>
__global__ __launch_bounds__(256) void k_tri_setup(const float* __restrict__ tris, int64_t T, GridDev g0, Meshes M,
                                                   int4* __restrict__ box0, int4* __restrict__ box1,
                                                   uint32_t* __restrict__ sat_items,
                                                   int4* __restrict__ ray_rng,
                                                   uint32_t* __restrict__ ray_items,
                                                   double4* __restrict__ slab1,
                                                   float4* __restrict__ slab2,
                                                   unsigned long long* __restrict__ bad_input)
{
    uint64_t tests = 0;
    const int64_t first = (int64_t)blockIdx.x * blockDim.x;
    for (int64_t t = first + threadIdx.x; t < T; t += (int64_t)gridDim.x * blockDim.x) {
        const int mesh = kM ? (int)M.tri_mesh[t] : 0;
        const GridDev& g = grid_of<kM>(g0, M, mesh);
        float tv[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) tv[c] = tris[t * 9 + c];
        MtTri m;
        const bool vis = mt_setup(tv, m);
        if constexpr ((kPart & 1) != 0) {
            // the input domain: finite coordinates within +-2^40, where every
            // intermediate of the SAT and of the split +z ray test is finite
            bool bad = false;
#pragma unroll
            for (int c = 0; c < 9; ++c) bad = bad || !(fabsf(tv[c]) <= 0x1p40f);
            if (bad) atomicOr(bad_input, 1ull);
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
            box0[t] = make_int4(lo[0], lo[1], lo[2], vis ? 1 : 0);
            box1[t] = make_int4(dd[0], dd[1], dd[2], 0);
            const bool any = dd[0] > 0 && dd[1] > 0 && dd[2] > 0;
            sat_items[t] = any ? (uint32_t)(dd[0] * dd[1]) : 0u;
            tests += any ? (uint64_t)dd[0] * dd[1] * dd[2] : 0;
            if (any && dd[2] >= kZSlabMin) {
                const double clo[3] = {g.c1x[lo[0]], g.c1y[lo[1]], g.c1z[lo[2]]};
                const double chi[3] = {g.c1x[lo[0] + dd[0] - 1], g.c1y[lo[1] + dd[1] - 1], g.c1z[lo[2] + dd[2] - 1]};
                const PlaneSlab sl = plane_slab(tv, g.h1x, g.h1y, g.h1z, clo, chi);
                slab1[t] = make_double4(sl.nx, sl.ny, sl.nz, sl.lim);
            }
            if (any && g.n2 > 0) {
                // over every Level2 center of the index AABB's Level1 voxels (the
                // tables are monotone, so the extreme entries bound them)
                const int n2 = g.n2;
                const double clo[3] = {g.c2x[(int64_t)lo[0] * n2], g.c2y[(int64_t)lo[1] * n2], g.c2z[(int64_t)lo[2] * n2]};
                const double chi[3] = {g.c2x[(int64_t)(lo[0] + dd[0]) * n2 - 1], g.c2y[(int64_t)(lo[1] + dd[1]) * n2 - 1],
                                       g.c2z[(int64_t)(lo[2] + dd[2]) * n2 - 1]};
                const FloatSlab f = float_slab(plane_slab(tv, g.h2x, g.h2y, g.h2z, clo, chi), tv, clo, chi);
                slab2[t] = make_float4(f.nx, f.ny, f.nz, f.lim);
            }
        }
        if constexpr ((kPart & 2) == 0) continue;
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
    if constexpr ((kPart & 1) == 0) return;
    if constexpr (!kM) {
        block_add_u64(tests, &M.ctr[0].tests_l1);
    } else {
        const int64_t last = (first + (int64_t)blockDim.x < T ? first + (int64_t)blockDim.x : T) - 1;
        const uint32_t ma = M.tri_mesh[first], mb = M.tri_mesh[last];
        if (ma == mb) {
            block_add_u64(tests, &M.ctr[ma].tests_l1);
        } else if (tests) {
            atomicAdd(&M.ctr[M.tri_mesh[first + threadIdx.x]].tests_l1, (unsigned long long)tests);
        }
    }
}

template <int kPart>
static void tri_setup_part(const float* tris, int64_t T, const GridDev& g0, const Meshes& M, int4* box0, int4* box1,
                           uint32_t* sat_items, int4* ray_rng, uint32_t* ray_items, double4* slab1, float4* slab2,
                           unsigned long long* bad_input, hipStream_t s)
{
    if (M.nm > 1)
        k_tri_setup<true, kPart><<<grid_for(T, 256, 1 << 30), 256, 0, s>>>(tris, T, g0, M, box0, box1, sat_items,
                                                                            ray_rng, ray_items, slab1, slab2,
                                                                            bad_input);
    else
        k_tri_setup<false, kPart><<<grid_for(T, 256, 1024), 256, 0, s>>>(tris, T, g0, M, box0, box1, sat_items,
                                                                          ray_rng, ray_items, slab1, slab2,
                                                                          bad_input);
}

void launch_tri_setup(const float* tris, int64_t T, const GridDev& g0, const Meshes& M, int4* box0, int4* box1,
                      uint32_t* sat_items, int4* ray_rng, uint32_t* ray_items, double4* slab1, float4* slab2,
                      int part, unsigned long long* bad_input, hipStream_t s)
{
    if (T <= 0) return;
    if (part == 1)
        tri_setup_part<1>(tris, T, g0, M, box0, box1, sat_items, ray_rng, ray_items, slab1, slab2, bad_input, s);
    else if (part == 2)
        tri_setup_part<2>(tris, T, g0, M, box0, box1, sat_items, ray_rng, ray_items, slab1, slab2, bad_input, s);
    else
        tri_setup_part<3>(tris, T, g0, M, box0, box1, sat_items, ray_rng, ray_items, slab1, slab2, bad_input, s);
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

// apply with the block prefix summed in place from the block sums (two
// launches per scan instead of a recursion, for up to SCAN_TILE blocks)
template <typename TI, typename TO>
__global__ __launch_bounds__(SCAN_NT) void k_scan_apply_direct(const TI* __restrict__ in, int64_t n,
                                                              const TO* __restrict__ sums, TO* __restrict__ out)
{
    __shared__ TO s_wave[SCAN_NT / 64];
    __shared__ TO s_total;
    // this block's prefix: the sum of the sums of the blocks before it
    TO pre = 0;
    for (int k = threadIdx.x; k < (int)blockIdx.x; k += SCAN_NT) pre += sums[k];
    (void)block_excl_scan<TO>(pre, s_wave, &s_total);   // leaves the block-wide sum in s_total
    const TO block_prefix = s_total;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_IPT;
    TO v[SCAN_IPT];
    TO acc = 0;
#pragma unroll
    for (int k = 0; k < SCAN_IPT; ++k) {
        v[k] = (base + k < n) ? (TO)in[base + k] : (TO)0;
        acc += v[k];
    }
    TO run = block_excl_scan<TO>(acc, s_wave, &s_total) + block_prefix;
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
    if (nb <= SCAN_TILE) {
        k_scan_apply_direct<TI, TO><<<(unsigned)nb, SCAN_NT, 0, s>>>(in, n, sums, out);
        return;
    }
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
// triangles).  Batches use blocks of 512 threads (the many short item runs
// of small meshes amortise each pass's searches and staging over twice the
// items: Level1 SAT of the mixed-size stream 11.9 -> 8.7 ms); one mesh uses
// 256 (with 512 its Level1 SAT beside the ray expansion took 0.21 -> 0.25 ms).
// ---------------------------------------------------------------------------
template <bool kM>
constexpr int expand_threads()
{
    return kM ? 512 : 256;
}

// 64-ary search by one wave: largest t in [lo, hi) with off[t] <= x, given
// off[lo] <= x and (hi == n or off[hi] > x).  ~log64 rounds of dependent
// loads instead of log2.
__device__ __forceinline__ int64_t wave_find_segment(const uint64_t* __restrict__ off, int64_t lo, int64_t hi,
                                                     uint64_t x)
{
    const int lane = __lane_id();
    while (hi - lo > 1) {
        const int64_t step = (hi - lo + 63) / 64;
        const int64_t idx = lo + (int64_t)lane * step;
        const bool le = idx < hi && off[idx] <= x;
        const uint64_t m = __ballot(le);
        const int L = 63 - __clzll((long long)m);
        const int64_t nlo = lo + (int64_t)L * step;
        hi = nlo + step < hi ? nlo + step : hi;
        lo = nlo;
    }
    return lo;
}

// Block-local append buffer in LDS: wave-aggregated LDS atomics, one global
// reservation per flush.  Entries that do not fit spill straight to the
// global buffer with a wave-aggregated global atomic.
struct LdsAppend {
    uint64_t* key;
    uint32_t* val;
    uint32_t* count;
    uint32_t cap;
};

__device__ __forceinline__ void lds_append(const LdsAppend& L, const Append64& G, bool pred, uint64_t key,
                                           uint32_t val)
{
    const uint64_t m = __ballot(pred);
    if (m == 0) return;
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if ((int)__lane_id() == leader) base = atomicAdd(L.count, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    const uint32_t idx = base + (uint32_t)__popcll(m & lanes_below_mask());
    if (pred && idx < L.cap) {
        L.key[idx] = key;
        if (L.val) L.val[idx] = val;
    }
    wave_append(G, pred && idx >= L.cap, key, val);
}

// block-synchronous: every thread of the block calls it
__device__ __forceinline__ void lds_flush(const LdsAppend& L, const Append64& G, uint64_t* s_gbase)
{
    __syncthreads();
    const uint32_t n = *L.count < L.cap ? *L.count : L.cap;
    if (threadIdx.x == 0 && n) *s_gbase = atomicAdd(G.count, (unsigned long long)n);
    __syncthreads();
    if (n) {
        const uint64_t gb = *s_gbase;
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
            const uint64_t idx = gb + i;
            if (idx < G.cap) {
                G.key[idx] = L.key[i];
                if (G.val) G.val[idx] = L.val[i];
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) *L.count = 0;
    __syncthreads();
}

struct NoAfter {
    __device__ void operator()(bool) const {}
};

// Each block owns a contiguous range of items, walked XB at a time, so the
// triangle searches of consecutive iterations start where the last ended.
// after(last) runs block-synchronously after every iteration.  XCAP: the
// triangles one pass stages in LDS (the tuned "triangles per LDS batch").
template <int XB, int XCAP, typename F, typename A>
__device__ __forceinline__ void expand_items(const float* __restrict__ tris, int64_t T,
                                             const uint64_t* __restrict__ off, F&& body, A&& after)
{
    __shared__ uint64_t s_off[XCAP + 1];
    __shared__ float s_tv[XCAP * 9];
    __shared__ int64_t s_t[2];
    const uint64_t total = off[T];
    const int tid = threadIdx.x;
    const uint64_t per = ((total + gridDim.x - 1) / gridDim.x + XB - 1) / XB * XB;
    // workgroups are dealt to the 8 XCDs round-robin and each XCD has its own
    // L2: give the blocks of one XCD one contiguous run of ranges, so the
    // atomics on neighbouring bitplane words mostly stay in one L2
    const uint32_t G = gridDim.x;
    const uint32_t lb = (G % 8u == 0u) ? (blockIdx.x % 8u) * (G / 8u) + blockIdx.x / 8u : blockIdx.x;
    const uint64_t b0 = (uint64_t)lb * per;
    const uint64_t b1 = b0 + per < total ? b0 + per : total;
    int64_t lo_hint = 0;
    for (uint64_t base = b0; base < b1; base += XB) {
        const uint64_t last = (base + XB < b1 ? base + XB : b1) - 1;
        if (tid < 64) {
            const int64_t r = wave_find_segment(off, lo_hint, T, base);
            if (tid == 0) s_t[0] = r;
        } else if (tid < 128) {
            const int64_t r = wave_find_segment(off, lo_hint, T, last);
            if (tid == 64) s_t[1] = r;
        }
        __syncthreads();
        const int64_t t0 = s_t[0];
        const int64_t nt = s_t[1] - t0 + 1;
        lo_hint = s_t[1];
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
        after(base + XB >= b1);
    }
}

// Level1 SAT: one item = (triangle, xy column of its index-AABB); the lane
// walks the AABB's z range.  Hits set the voxel's boundary byte, append the
// (voxel, triangle) bin pair -- key = batch-wide bit index, value = triangle --
// and the lane's highest hit z becomes the (column, triangle, zmax) entry used
// by the Level2 parity candidates.
template <bool kM, int XCAP>
__global__ __launch_bounds__(expand_threads<kM>()) void k_sat_expand(const float* __restrict__ tris, int64_t T,
                                                   const int4* __restrict__ box0,
                                                   const int4* __restrict__ box1,
                                                   const double4* __restrict__ slab1,
                                                   const uint64_t* __restrict__ off, GridDev g0, Meshes M,
                                                   uint8_t* __restrict__ bbytes, Append64 pairs,
                                                   Append64 cols)
{
    constexpr uint32_t PB = 1536, CB = 768;
    __shared__ uint64_t s_pk[PB];
    __shared__ uint32_t s_pv[PB];
    __shared__ uint64_t s_ck[CB];
    __shared__ uint32_t s_cv[CB];
    __shared__ uint32_t s_n[2];
    __shared__ uint64_t s_gbase;
    if (threadIdx.x < 2) s_n[threadIdx.x] = 0;
    __syncthreads();
    const LdsAppend lp{s_pk, s_pv, &s_n[0], PB};
    const LdsAppend lc{s_ck, s_cv, &s_n[1], CB};
    expand_items<expand_threads<kM>(), XCAP>(tris, T, off, [&](int64_t t, uint64_t l, const float* tv) {
        const int mesh = kM ? (int)M.tri_mesh[t] : 0;
        const GridDev& g = grid_of<kM>(g0, M, mesh);
        const uint64_t word0 = kM ? M.off[mesh].word0 : 0;
        const uint32_t col0 = kM ? M.off[mesh].col0 : 0;
        const int4 b0 = box0[t];
        const int4 b1 = box1[t];
        const int p = b0.x + (int)(l % (uint64_t)b1.x);
        const int q = b0.y + (int)(l / (uint64_t)b1.x);
        const float cx = g.c1x[p], cy = g.c1y[q];
        const float hx = g.h1x, hy = g.h1y, hz = g.h1z;
        const uint64_t plane = (uint64_t)g.ny * g.nxw;
        const uint64_t rowword = word0 + (uint64_t)q * g.nxw + (uint64_t)(p >> 6);
        const float* c1z = g.c1z;
        int za = b0.z, zb = b0.z + b1.z;
        if (b1.z >= kZSlabMin) {
            const double4 sl = slab1[t];
            slab_layers(sl.x, sl.y, sl.z, sl.w, cx, cy, tv, c1z, b0.z, b0.z + b1.z, za, zb);
        }
        // the e x z edge predicates hold or fail for the whole column
        if (!xy_edges_ok(cx, cy, hx, hy, tv)) zb = za;
        int zmax = -1;
        for (int z = za; z < zb; ++z) {
            const bool hit = tri_box_overlap<true, false>(cx, cy, c1z[z], hx, hy, hz, tv);
            const uint64_t w = (uint64_t)z * plane + rowword;
            if (hit) {
                zmax = z;
                bbytes[w * 64 + (uint64_t)(p & 63)] = 1;   // every writer stores the same byte value
            }
            lds_append(lp, pairs, hit, w * 64 + (uint64_t)(p & 63), (uint32_t)t);
        }
        const bool emit = (zmax >= 0) && (b0.w != 0);
        lds_append(lc, cols, emit, ((uint64_t)(col0 + (uint32_t)q * g.nx + p) << 32) | (uint64_t)t,
                   (uint32_t)zmax);
    }, [&](bool last) {
        // flush when half full (block-uniform: counters read after a barrier)
        __syncthreads();
        const bool fp = last || s_n[0] >= PB / 2;
        const bool fc = last || s_n[1] >= CB / 2;
        if (fp) lds_flush(lp, pairs, &s_gbase);
        if (fc) lds_flush(lc, cols, &s_gbase);
    });
}

template <int XC>
static void launch_sat_expand_x(const float* tris, int64_t T, const int4* box0, const int4* box1,
                                const double4* slab1, const uint64_t* off, const GridDev& g0, const Meshes& M,
                                uint8_t* b, Append64 pairs, Append64 cols, int blocks, hipStream_t s)
{
    if (M.nm > 1)
        k_sat_expand<true, XC><<<blocks, expand_threads<true>(), 0, s>>>(tris, T, box0, box1, slab1, off, g0, M, b,
                                                                        pairs, cols);
    else
        k_sat_expand<false, XC><<<blocks, expand_threads<false>(), 0, s>>>(tris, T, box0, box1, slab1, off, g0, M,
                                                                          b, pairs, cols);
}

void launch_sat_expand(const float* tris, int64_t T, const int4* box0, const int4* box1, const double4* slab1,
                       const uint64_t* off, const GridDev& g0, const Meshes& M, uint8_t* b, Append64 pairs,
                       Append64 cols, int blocks, int stage, hipStream_t s)
{
    if (T <= 0) return;
    if (stage <= 128)
        launch_sat_expand_x<128>(tris, T, box0, box1, slab1, off, g0, M, b, pairs, cols, blocks, s);
    else
        launch_sat_expand_x<256>(tris, T, box0, box1, slab1, off, g0, M, b, pairs, cols, blocks, s);
}

// boundary flags (one byte per padded voxel, written without atomics by the
// Level1 SAT) -> the boundary bitplane, one word per wave
__global__ void k_pack_bytes(const uint8_t* __restrict__ by, int64_t W, uint64_t* __restrict__ bits)
{
    const int64_t n = W * 64;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t m = __ballot(by[i] != 0);
        if (__lane_id() == 0) bits[i >> 6] = m;
    }
}

void launch_pack_bytes(const uint8_t* by, int64_t W, uint64_t* bits, hipStream_t s)
{
    if (W <= 0) return;
    k_pack_bytes<<<grid_for(W * 64, 256, 65536), 256, 0, s>>>(by, W, bits);
}

// Level1 ray parity, first half: one item = (triangle, column inside its
// conservative bound); the exact det/u/v test keeps the (column, triangle)
// pairs whose ray the reference's MT accepts at some height.
template <bool kM, int XCAP>
__global__ __launch_bounds__(expand_threads<kM>()) void k_ray_expand(const float* __restrict__ tris, int64_t T,
                                                   const int4* __restrict__ rng,
                                                   const uint64_t* __restrict__ off, GridDev g0, Meshes M,
                                                   Append64 hits)
{
    constexpr uint32_t HB = 1024;
    __shared__ uint64_t s_hk[HB];
    __shared__ uint32_t s_n;
    __shared__ uint64_t s_gbase;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const LdsAppend lh{s_hk, nullptr, &s_n, HB};
    expand_items<expand_threads<kM>(), XCAP>(tris, T, off, [&](int64_t t, uint64_t l, const float* tv) {
        const int mesh = kM ? (int)M.tri_mesh[t] : 0;
        const GridDev& g = grid_of<kM>(g0, M, mesh);
        const uint32_t col0 = kM ? M.off[mesh].col0 : 0;
        const int4 r = rng[t];
        const int i = r.x + (int)(l % (uint64_t)r.z);
        const int j = r.y + (int)(l / (uint64_t)r.z);
        MtTri m;
        bool ok = mt_setup(tv, m);
        float tx, ty;
        ok = ok && mt_column(m, tv[0], tv[1], g.r1x[i], g.r1y[j], tx, ty);
        lds_append(lh, hits, ok, ((uint64_t)(col0 + (uint32_t)j * g.nx + i) << 32) | (uint64_t)t, 0);
    }, [&](bool last) {
        __syncthreads();
        if (last || s_n >= HB / 2) lds_flush(lh, hits, &s_gbase);
    });
}

template <int XC>
static void launch_ray_expand_x(const float* tris, int64_t T, const int4* ray_rng, const uint64_t* off,
                                const GridDev& g0, const Meshes& M, Append64 hits, int blocks, hipStream_t s)
{
    if (M.nm > 1)
        k_ray_expand<true, XC><<<blocks, expand_threads<true>(), 0, s>>>(tris, T, ray_rng, off, g0, M, hits);
    else
        k_ray_expand<false, XC><<<blocks, expand_threads<false>(), 0, s>>>(tris, T, ray_rng, off, g0, M, hits);
}

void launch_ray_expand(const float* tris, int64_t T, const int4* ray_rng, const uint64_t* off, const GridDev& g0,
                       const Meshes& M, Append64 hits, int blocks, int stage, hipStream_t s)
{
    if (T <= 0) return;
    if (stage <= 128)
        launch_ray_expand_x<128>(tris, T, ray_rng, off, g0, M, hits, blocks, s);
    else
        launch_ray_expand_x<256>(tris, T, ray_rng, off, g0, M, hits, blocks, s);
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

template <bool kM>
__global__ void k_scatter_hits(const float* __restrict__ tris, const uint64_t* __restrict__ key, uint64_t n,
                               GridDev g0, Meshes M, const uint32_t* __restrict__ off, uint32_t* __restrict__ cnt,
                               float4* __restrict__ rec)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t k = key[i];
        const uint32_t gcol = (uint32_t)(k >> 32);
        const uint32_t t = (uint32_t)(k & 0xffffffffu);
        const uint32_t pos = off[gcol] + atomicSub(&cnt[gcol], 1u) - 1u;
        const int mesh = kM ? (int)M.tri_mesh[t] : 0;
        const GridDev& g = grid_of<kM>(g0, M, mesh);
        const uint32_t col = gcol - (kM ? M.off[mesh].col0 : 0u);
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

void launch_scatter_hits(const float* tris, const uint64_t* key, uint64_t n, const GridDev& g0, const Meshes& M,
                         const uint32_t* off, uint32_t* cnt, float4* rec, hipStream_t s)
{
    if (!n) return;
    const int blocks = grid_for((int64_t)n, 256, 16384);
    if (M.nm > 1)
        k_scatter_hits<true><<<blocks, 256, 0, s>>>(tris, key, n, g0, M, off, cnt, rec);
    else
        k_scatter_hits<false><<<blocks, 256, 0, s>>>(tris, key, n, g0, M, off, cnt, rec);
}

// column entries -> CSR by column, each with the MT record its candidates need
struct ColSrc {
    const uint64_t* key;
    const uint32_t* zmax;
    __device__ uint32_t sort_key(uint64_t i) const { return (uint32_t)(key[i] >> 32); }
    __device__ uint32_t count_key(uint64_t i) const { return sort_key(i); }
    __device__ void keep_key(uint64_t, uint32_t) const {}
    // (column, triangle, zmax)
    __device__ uint4 item(uint64_t i) const
    {
        const uint64_t k = key[i];
        return make_uint4((uint32_t)(k >> 32), (uint32_t)(k & 0xffffffffu), zmax[i], 0u);
    }
};

// place(pos, it, ok): loads for any valid item, stores only when ok
struct ColPlace {
    const float* tris;
    float4* rec;
    __device__ void operator()(uint32_t pos, const uint4& it, bool ok) const
    {
        const uint32_t t = it.y;
        float tv[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) tv[c] = tris[(uint64_t)t * 9 + c];
        MtTri m;
        mt_setup(tv, m);   // entries exist only for triangles whose det passes
        if (!ok) return;
        rec[3 * (uint64_t)pos + 0] = make_float4(m.e1x, m.e1y, m.e1z, m.e2x);
        rec[3 * (uint64_t)pos + 1] = make_float4(m.e2y, m.e2z, m.inv_det, tv[0]);
        // the last word: the entry's batch-wide column (read by the packed column pass)
        rec[3 * (uint64_t)pos + 2] = make_float4(tv[1], tv[2], __int_as_float((int)it.z), __uint_as_float(it.x));
    }
};

// ---------------------------------------------------------------------------
// Level1 parity: one wave per task = L1P_Z consecutive layers of a 64-column
// strip of one row; each lane loads its column's accepted (triangle, column)
// records once and XORs t > 1e-6 over them at each of the task's layers.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int wave_mesh_of_task(const MeshOff* __restrict__ off, int nm, uint64_t task)
{
    return wave_mesh_search(off, nm, task, [](const MeshOff& o) { return (uint64_t)o.ptask0; });
}

template <bool kM>
__global__ __launch_bounds__(256) void k_l1_parity(const float4* __restrict__ rec,
                                                   const uint32_t* __restrict__ hoff, GridDev g0, Meshes M,
                                                   uint64_t* __restrict__ ins, int64_t NT)
{
    const int64_t task = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (task >= NT) return;
    const int lane = __lane_id();
    const int mesh = kM ? wave_mesh_of_task(M.off, M.nm, (uint64_t)task) : 0;
    const GridDev& g = grid_of<kM>(g0, M, mesh);
    if (g.l1st) return;   // the stencil fill writes this mesh's inside bits
    const int64_t tl = task - (int64_t)(kM ? M.off[mesh].ptask0 : 0);
    const uint64_t word0 = kM ? M.off[mesh].word0 : 0;
    const uint32_t col0 = kM ? M.off[mesh].col0 : 0;
    const int nxw = g.nxw, ny = g.ny, nx = g.nx, nz = g.nz;
    const int xw = (int)(tl % nxw);
    const int64_t rowt = tl / nxw;
    const int j = (int)(rowt % ny);
    const int z0 = (int)(rowt / ny) * L1P_Z;
    const int nzc = nz - z0 < L1P_Z ? nz - z0 : L1P_Z;
    float oz[L1P_Z];
#pragma unroll
    for (int k = 0; k < L1P_Z; ++k) oz[k] = k < nzc ? g.r1z[z0 + k] : 0.0f;
    const int x = xw * 64 + lane;
    uint32_t par = 0;   // bit k: parity at layer z0 + k
    if (x < nx) {
        const uint32_t c = col0 + (uint32_t)j * nx + x;
        const uint32_t h1 = hoff[c + 1];
        for (uint32_t h = hoff[c]; h < h1; ++h) {
            const float4 a = rec[3 * (uint64_t)h + 0];
            const float4 b = rec[3 * (uint64_t)h + 1];
            const float4 d = rec[3 * (uint64_t)h + 2];
#pragma unroll
            for (int k = 0; k < L1P_Z; ++k)
                if (k < nzc) par ^= (uint32_t)mt_t_hit(a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, d.x, d.y, oz[k]) << k;
        }
    }
    for (int k = 0; k < nzc; ++k) {
        const uint64_t m = __ballot((par >> k) & 1u);
        if (lane == 0) ins[word0 + ((uint64_t)(z0 + k) * ny + j) * nxw + xw] = m;
    }
}

void launch_l1_parity(const float4* rec, const uint32_t* hoff, const GridDev& g0, const Meshes& M, int64_t NT,
                      uint64_t* ins, hipStream_t s)
{
    if (NT <= 0) return;
    const int64_t blocks = (NT * 64 + 255) / 256;
    if (M.nm > 1)
        k_l1_parity<true><<<(unsigned)blocks, 256, 0, s>>>(rec, hoff, g0, M, ins, NT);
    else
        k_l1_parity<false><<<(unsigned)blocks, 256, 0, s>>>(rec, hoff, g0, M, ins, NT);
}

// ---------------------------------------------------------------------------
// a batch's device soups gathered into the batch's soup in one launch (one
// grid row per mesh, each row's blocks striding over the mesh's floats)
// ---------------------------------------------------------------------------
// kHost: the sources are page-locked host memory, read with system-scope
// loads so that no line a cache kept from an earlier call is returned (the
// caller may have rewritten the buffer since)
template <bool kHost>
__global__ __launch_bounds__(256) void k_gather_soups(const SoupPiece* __restrict__ pc, int m0,
                                                      float* __restrict__ dst)
{
    const SoupPiece p = pc[m0 + blockIdx.y];
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < p.n; i += stride) {
        if constexpr (kHost)
            dst[p.at + i] = __hip_atomic_load(p.src + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        else
            dst[p.at + i] = p.src[i];
    }
}

void launch_gather_soups(const SoupPiece* pc, int nm, float* dst, int blocks, bool host, hipStream_t s)
{
    for (int m0 = 0; m0 < nm; m0 += 65535) {
        const int rows = nm - m0 < 65535 ? nm - m0 : 65535;
        if (host)
            k_gather_soups<true><<<dim3((unsigned)blocks, (unsigned)rows), 256, 0, s>>>(pc, m0, dst);
        else
            k_gather_soups<false><<<dim3((unsigned)blocks, (unsigned)rows), 256, 0, s>>>(pc, m0, dst);
    }
}

// ---------------------------------------------------------------------------
// Level1 inside/outside by the GL-stencil-equivalent fill (opt-in,
// gpv_params::l1_stencil; ClassifyInOut2x, src/Object.cpp:1446-1601).  Per
// slice s the reference draws the mesh orthographically over the bbox into a
// 2nx x 2ny stencil, clipped to z <= zs, counts the fragments per sample,
// and calls a voxel inside when more than 2 of its 4 samples hold an odd
// count.  Which samples a triangle covers is the GL driver's choice; the
// rule here is fixed: vertices snapped to 1/256 sample in window
// coordinates, sample (px, py) at (256 px + 128, 256 py + 128), edge
// functions of the counter-clockwise snapped triangle positive, or zero on a
// top or left edge (shared edges count once), no sample for a zero snapped
// area; a covered sample counts in slice s when the height of the triangle's
// plane (float vertices, double arithmetic) over the sample's world point is
// at most the slice's ray-origin z.  A covered sample toggles bit (s0, py,
// px) of a [nz][2ny][2nx] bit set at its first slice s0; the running XOR
// along s is then the stencil's parity, and the 2 x 2 vote writes the
// Level1 inside bits.  Counts are exact (GL's 8-bit stencil saturates).
// The kernels read only the soup and the grid, so the pipeline runs them
// beside the Level1 SAT.
// ---------------------------------------------------------------------------
struct StTri {
    long long X[3], Y[3];             // snapped window coordinates, counter-clockwise
    int pxlo, pxhi, pylo, pyhi;       // samples whose centers lie in the snapped bbox
    double v0x, v0y, v0z, nx, ny, nz; // first vertex and plane normal
};

__device__ __forceinline__ long long st_snap(float x, float lo, float hi, int samples)
{
    double w = ((double)x - (double)lo) / ((double)hi - (double)lo) * (double)samples;
    // vertices far outside the bbox: bounded so that edge functions stay in 64 bits
    w = w < -4194304.0 ? -4194304.0 : (w > 4194304.0 ? 4194304.0 : w);
    return (long long)rint(w * 256.0);
}

__device__ __forceinline__ long long st_floor256(long long a) { return a >= 0 ? a / 256 : -((255 - a) / 256); }

// false when the triangle covers no sample in any slice
__device__ bool st_prepare(const float* __restrict__ tv, const GridDev& g, StTri& s)
{
    const int W = 2 * g.nx, H = 2 * g.ny;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        s.X[i] = st_snap(tv[3 * i], g.bminx, g.bmaxx, W);
        s.Y[i] = st_snap(tv[3 * i + 1], g.bminy, g.bmaxy, H);
    }
    const long long A = (s.X[1] - s.X[0]) * (s.Y[2] - s.Y[0]) - (s.Y[1] - s.Y[0]) * (s.X[2] - s.X[0]);
    if (A == 0) return false;
    if (A < 0) {
        long long t = s.X[1]; s.X[1] = s.X[2]; s.X[2] = t;
        t = s.Y[1]; s.Y[1] = s.Y[2]; s.Y[2] = t;
    }
    s.v0x = (double)tv[0]; s.v0y = (double)tv[1]; s.v0z = (double)tv[2];
    const double e1x = (double)tv[3] - s.v0x, e1y = (double)tv[4] - s.v0y, e1z = (double)tv[5] - s.v0z;
    const double e2x = (double)tv[6] - s.v0x, e2y = (double)tv[7] - s.v0y, e2z = (double)tv[8] - s.v0z;
    s.nx = e1y * e2z - e1z * e2y;
    s.ny = e1z * e2x - e1x * e2z;
    s.nz = e1x * e2y - e1y * e2x;
    if (s.nz == 0.0) return false;
    long long mnx = s.X[0], mxx = s.X[0], mny = s.Y[0], mxy = s.Y[0];
#pragma unroll
    for (int i = 1; i < 3; ++i) {
        mnx = s.X[i] < mnx ? s.X[i] : mnx;
        mxx = s.X[i] > mxx ? s.X[i] : mxx;
        mny = s.Y[i] < mny ? s.Y[i] : mny;
        mxy = s.Y[i] > mxy ? s.Y[i] : mxy;
    }
    const long long xlo = st_floor256(mnx - 128 + 255), xhi = st_floor256(mxx - 128);
    const long long ylo = st_floor256(mny - 128 + 255), yhi = st_floor256(mxy - 128);
    s.pxlo = (int)(xlo < 0 ? 0 : xlo);
    s.pylo = (int)(ylo < 0 ? 0 : ylo);
    s.pxhi = (int)(xhi > W - 1 ? W - 1 : xhi);
    s.pyhi = (int)(yhi > H - 1 ? H - 1 : yhi);
    return s.pxlo <= s.pxhi && s.pylo <= s.pyhi;
}

__device__ __forceinline__ bool st_covers(const StTri& s, int px, int py)
{
    const long long sx = 256ll * px + 128, sy = 256ll * py + 128;
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        const int a = e, b = e == 2 ? 0 : e + 1;
        const long long dx = s.X[b] - s.X[a], dy = s.Y[b] - s.Y[a];
        const long long E = dx * (sy - s.Y[a]) - dy * (sx - s.X[a]);
        if (E < 0 || (E == 0 && !(dy < 0 || (dy == 0 && dx < 0)))) return false;
    }
    return true;
}

// first slice whose clip plane keeps the triangle's point over the sample
// (z_T <= zs); nz when none does
__device__ __forceinline__ int st_slice(const StTri& s, const GridDev& g, int px, int py)
{
    const double l = (double)g.bminx, r = (double)g.bmaxx, b = (double)g.bminy, t = (double)g.bmaxy;
    const double x = l + ((double)px + 0.5) * ((r - l) / (double)(2 * g.nx));
    const double y = b + ((double)py + 0.5) * ((t - b) / (double)(2 * g.ny));
    const double zt = s.v0z - (s.nx * (x - s.v0x) + s.ny * (y - s.v0y)) / s.nz;
    int lo = 0, hi = g.nz;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((double)g.r1z[mid] < zt) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_st_count(const float* __restrict__ tris, int64_t T, GridDev g,
                                                  uint32_t* __restrict__ cnt)
{
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < T; t += (int64_t)gridDim.x * blockDim.x) {
        StTri s;
        cnt[t] = st_prepare(tris + 9 * t, g, s)
                     ? (uint32_t)(s.pxhi - s.pxlo + 1) * (uint32_t)(s.pyhi - s.pylo + 1)
                     : 0u;
    }
}

// the (triangle, sample) items, ST_RUN consecutive ones per thread: one
// search for the run's first triangle, then forward
constexpr int ST_RUN = 8;
__global__ __launch_bounds__(256) void k_st_toggle(const float* __restrict__ tris, int64_t T, GridDev g,
                                                   const uint64_t* __restrict__ off, int wpr,
                                                   unsigned long long* __restrict__ tog)
{
    const uint64_t N = off[T];
    const int H = 2 * g.ny;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x * ST_RUN;
    for (uint64_t i0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * ST_RUN; i0 < N; i0 += stride) {
        const uint64_t i1 = i0 + ST_RUN < N ? i0 + ST_RUN : N;
        int64_t t = find_segment(off, T, i0);
        StTri s;
        st_prepare(tris + 9 * t, g, s);
        for (uint64_t i = i0; i < i1; ++i) {
            if (off[t + 1] <= i) {
                do ++t; while (off[t + 1] <= i);
                st_prepare(tris + 9 * t, g, s);
            }
            const uint32_t li = (uint32_t)(i - off[t]);
            const uint32_t bw = (uint32_t)(s.pxhi - s.pxlo + 1);
            const int py = s.pylo + (int)(li / bw), px = s.pxlo + (int)(li % bw);
            if (!st_covers(s, px, py)) continue;
            const int k = st_slice(s, g, px, py);
            if (k < g.nz)
                atomicXor(&tog[((int64_t)k * H + py) * wpr + (px >> 6)], 1ull << (px & 63));
        }
    }
}

// One wave per Level1 row word (row j, 64 voxels) walks the slices: the
// running XOR of the toggles of sample rows 2j and 2j + 1 is the stencil's
// parity at each slice, and a voxel is inside when more than 2 of its
// samples (2x + {0,1}, 2j + {0,1}) are odd.  The slices' toggle words are
// loaded ST_ZU at a time ahead of their use.
constexpr int ST_ZU = 8;
__global__ __launch_bounds__(256) void k_st_vote(const unsigned long long* __restrict__ tog, GridDev g, int wpr,
                                                 uint64_t* __restrict__ ins)
{
    const int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nxw = g.nxw, ny = g.ny, nx = g.nx, nz = g.nz, H = 2 * g.ny;
    if (w >= (int64_t)ny * nxw) return;
    const int lane = __lane_id();
    const int xw = (int)(w % nxw), j = (int)(w / nxw);
    const int x = xw * 64 + lane;
    const int sx = 2 * x, sh = sx & 63;
    // lanes past the row read a word of the row (their votes are masked)
    const int wi = x < nx ? sx >> 6 : 0;
    const int64_t plane = (int64_t)H * wpr;
    const unsigned long long* r0 = tog + (int64_t)(2 * j) * wpr + wi;
    unsigned long long a = 0, b = 0;
    for (int k0 = 0; k0 < nz; k0 += ST_ZU) {
        unsigned long long ta[ST_ZU], tb[ST_ZU];
#pragma unroll
        for (int u = 0; u < ST_ZU; ++u) {
            const int k = k0 + u < nz ? k0 + u : nz - 1;
            ta[u] = r0[k * plane];
            tb[u] = r0[k * plane + wpr];
        }
#pragma unroll
        for (int u = 0; u < ST_ZU; ++u) {
            if (k0 + u >= nz) break;
            a ^= ta[u];
            b ^= tb[u];
            const int odd = (int)((a >> sh) & 1ull) + (int)((a >> (sh + 1)) & 1ull) + (int)((b >> sh) & 1ull) +
                            (int)((b >> (sh + 1)) & 1ull);
            const uint64_t m = __ballot(x < nx && odd > 2);
            if (lane == 0) ins[((int64_t)(k0 + u) * ny + j) * nxw + xw] = m;
        }
    }
}

size_t l1_stencil_words(const GridDev& g)
{
    return (size_t)g.nz * (size_t)(2 * g.ny) * (size_t)((2 * g.nx + 63) / 64);
}

void launch_l1_stencil(const float* tris, int64_t T, const GridDev& g, uint32_t* cnt, uint64_t* off,
                       uint64_t* scan_tmp, uint64_t* tog, uint64_t* ins, hipStream_t s)
{
    const int wpr = (2 * g.nx + 63) / 64;
    unsigned long long* tg = reinterpret_cast<unsigned long long*>(tog);
    launch_fill_u64(tog, 0, (int64_t)l1_stencil_words(g), s);
    if (T > 0) {
        k_st_count<<<grid_for(T, 256, 16384), 256, 0, s>>>(tris, T, g, cnt);
        launch_scan_u32_u64(cnt, off, T, scan_tmp, s);
        k_st_toggle<<<4096, 256, 0, s>>>(tris, T, g, off, wpr, tg);
    }
    const int64_t rows = (int64_t)g.ny * g.nxw;   // one wave each
    k_st_vote<<<(unsigned)((rows * 64 + 255) / 256), 256, 0, s>>>(tg, g, wpr, ins);
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

// b0 (first boundary rank), l2w0 (first Level2 word) and l2n0 (first Level2
// normal) of every mesh, from the batch-wide word prefix.
__global__ __launch_bounds__(64) void k_mesh_bounds(const uint32_t* __restrict__ wprefix, Meshes M)
{
    // one wave: 64 meshes per round, exclusive scans of the Level2 word and
    // normal counts carried across rounds
    const int lane = __lane_id();
    uint64_t l2 = 0, l2n = 0;
    for (int base = 0; base <= M.nm; base += 64) {
        const int m = base + lane;
        uint32_t b0 = 0;
        uint64_t w = 0, nrm = 0;
        if (m <= M.nm) {
            b0 = wprefix[M.off[m].word0];
            if (m < M.nm) {
                const uint64_t nb = wprefix[M.off[m + 1].word0] - b0;
                const GridDev& g = M.grids[m];
                w = nb * (uint64_t)g.w2;
                nrm = g.normals ? nb * (uint64_t)g.n2 * g.n2 * g.n2 : 0;
            }
        }
        const uint64_t wi = wave_incl_scan<uint64_t>(w), ni = wave_incl_scan<uint64_t>(nrm);
        if (m <= M.nm) {
            M.off[m].b0 = b0;
            M.off[m].l2w0 = l2 + wi - w;
            M.off[m].l2n0 = l2n + ni - nrm;
        }
        l2 += shfl64(wi, 63);
        l2n += shfl64(ni, 63);
    }
}

void launch_mesh_bounds(const uint32_t* wprefix, const Meshes& M, hipStream_t s)
{
    k_mesh_bounds<<<1, 64, 0, s>>>(wprefix, M);
}

// ascending linear indices (within the mesh) of the boundary voxels, and in a
// batch the mesh of each
template <bool kM>
__global__ void k_boundary_index(const uint64_t* __restrict__ bnd, const uint32_t* __restrict__ wprefix,
                                 GridDev g0, Meshes M, int64_t W, int32_t* __restrict__ bidx)
{
    for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < W;
         w += (int64_t)gridDim.x * blockDim.x) {
        uint64_t bits = bnd[w];
        if (!bits) continue;
        uint32_t pos = wprefix[w];
        const int mesh = kM ? mesh_of_word(M.off, M.nm, (uint64_t)w) : 0;
        const GridDev& g = grid_of<kM>(g0, M, mesh);
        const int64_t wl = w - (int64_t)(kM ? M.off[mesh].word0 : 0);
        const int xw = (int)(wl % g.nxw);
        const int64_t row = wl / g.nxw;
        while (bits) {
            const int bit = __ffsll((unsigned long long)bits) - 1;
            bits &= bits - 1;
            if constexpr (kM) M.bmesh[pos] = (uint32_t)mesh;
            bidx[pos++] = (int32_t)(row * g.nx + xw * 64 + bit);
        }
    }
}

void launch_boundary_index(const uint64_t* bnd, const uint32_t* wprefix, const GridDev& g0, const Meshes& M,
                           int64_t W, int32_t* bidx, hipStream_t s)
{
    if (W <= 0) return;
    const int blocks = grid_for(W, 256, 16384);
    if (M.nm > 1)
        k_boundary_index<true><<<blocks, 256, 0, s>>>(bnd, wprefix, g0, M, W, bidx);
    else
        k_boundary_index<false><<<blocks, 256, 0, s>>>(bnd, wprefix, g0, M, W, bidx);
}

__device__ __forceinline__ uint32_t rank_of(const uint64_t* __restrict__ bnd,
                                            const uint32_t* __restrict__ wprefix, uint64_t bitidx)
{
    const uint64_t w = bitidx >> 6;
    const int bit = (int)(bitidx & 63);
    const uint64_t below = bit ? (~0ull >> (64 - bit)) : 0ull;
    return wprefix[w] + (uint32_t)__popcll(bnd[w] & below);
}

// ---------------------------------------------------------------------------
// Counting sort of n items by a key in [0, nkeys) in two levels: tiles of
// 2^shift consecutive keys.  Pass 1 (count) and pass 2 (bucket) walk the same
// contiguous item range per block and histogram the tile ids in LDS, so the
// global counters see one atomic per (block, tile) instead of one per item;
// pass 3 (one block per tile) counts its keys in LDS, writes their CSR
// offsets and places every item at its final position.  The order inside one
// key is unspecified (as with any atomic placement).  Each thread keeps TS_U
// items in flight: their loads are issued together (indices clamped to a
// valid item, the LDS atomics and stores predicated), since the per-item
// chains of dependent loads, not bandwidth, bound these passes.
// ---------------------------------------------------------------------------
constexpr int TS_NT = 256;
constexpr int TS_U = 4;

__device__ __forceinline__ void ts_range(uint64_t n, uint64_t& lo, uint64_t& hi)
{
    const uint64_t per = (n + gridDim.x - 1) / gridDim.x;
    lo = (uint64_t)blockIdx.x * per;
    hi = lo + per < n ? lo + per : n;
}

// LDS histogram of the tile ids of items [lo, hi) of this block
// kCount: the count pass, which takes each item's key from count_key and
// hands it to keep_key (a source may keep it for the later passes); the
// bucket pass reads sort_key
template <bool kCount, typename Src>
__device__ __forceinline__ void ts_histogram(const Src& src, uint64_t lo, uint64_t hi, int shift, uint32_t* s_h)
{
    for (uint64_t i0 = lo + threadIdx.x; i0 < hi; i0 += (uint64_t)TS_U * TS_NT) {
        uint32_t k[TS_U];
#pragma unroll
        for (int u = 0; u < TS_U; ++u) {
            const uint64_t i = i0 + (uint64_t)u * TS_NT;
            k[u] = kCount ? src.count_key(i < hi ? i : i0) : src.sort_key(i < hi ? i : i0);
        }
#pragma unroll
        for (int u = 0; u < TS_U; ++u) {
            const uint64_t i = i0 + (uint64_t)u * TS_NT;
            if (i < hi) {
                atomicAdd(&s_h[k[u] >> shift], 1u);
                if constexpr (kCount) src.keep_key(i, k[u]);
            }
        }
    }
}

template <typename Src>
__global__ __launch_bounds__(TS_NT) void k_ts_count(Src src, uint64_t n, int shift, int ntiles,
                                                    uint32_t* __restrict__ tile_total)
{
    extern __shared__ uint32_t s_h[];   // [ntiles]
    for (int t = threadIdx.x; t < ntiles; t += TS_NT) s_h[t] = 0;
    __syncthreads();
    uint64_t lo, hi;
    ts_range(n, lo, hi);
    ts_histogram<true>(src, lo, hi, shift, s_h);
    __syncthreads();
    for (int t = threadIdx.x; t < ntiles; t += TS_NT)
        if (s_h[t]) atomicAdd(&tile_total[t], s_h[t]);
}

// items (16 bytes, .x = key) to their tile's bucket range; the item is built
// here, where the block reads its inputs in their original order
template <typename Src>
__global__ __launch_bounds__(TS_NT) void k_ts_bucket(Src src, uint64_t n, int shift, int ntiles,
                                                     const uint32_t* __restrict__ tile_off,
                                                     uint32_t* __restrict__ tile_fill, uint4* __restrict__ bucket)
{
    extern __shared__ uint32_t s_h[];   // [ntiles]: counts, then this block's next slot per tile
    for (int t = threadIdx.x; t < ntiles; t += TS_NT) s_h[t] = 0;
    __syncthreads();
    uint64_t lo, hi;
    ts_range(n, lo, hi);
    ts_histogram<false>(src, lo, hi, shift, s_h);
    __syncthreads();
    for (int t = threadIdx.x; t < ntiles; t += TS_NT)
        if (s_h[t]) s_h[t] = tile_off[t] + atomicAdd(&tile_fill[t], s_h[t]);
    __syncthreads();
    for (uint64_t i0 = lo + threadIdx.x; i0 < hi; i0 += (uint64_t)TS_U * TS_NT) {
        uint4 it[TS_U];
#pragma unroll
        for (int u = 0; u < TS_U; ++u) {
            const uint64_t i = i0 + (uint64_t)u * TS_NT;
            it[u] = src.item(i < hi ? i : i0);
        }
#pragma unroll
        for (int u = 0; u < TS_U; ++u)
            if (i0 + (uint64_t)u * TS_NT < hi) bucket[atomicAdd(&s_h[it[u].x >> shift], 1u)] = it[u];
    }
}

template <typename Place>
__global__ __launch_bounds__(TS_NT) void k_ts_place(const uint4* __restrict__ bucket,
                                                    const uint32_t* __restrict__ tile_off, int shift,
                                                    uint32_t nkeys, uint32_t* __restrict__ off,
                                                    unsigned int* __restrict__ max_out, Place place)
{
    extern __shared__ uint32_t s_c[];   // [2^shift] key counts, then next slots
    __shared__ uint32_t s_wave[TS_NT / 64];
    __shared__ uint32_t s_max;
    const uint32_t T = 1u << shift;
    const uint32_t k0 = (uint32_t)blockIdx.x << shift;
    const uint32_t b0 = tile_off[blockIdx.x], b1 = tile_off[blockIdx.x + 1];
    for (uint32_t j = threadIdx.x; j < T; j += TS_NT) s_c[j] = 0;
    if (threadIdx.x == 0) s_max = 0;
    __syncthreads();
    for (uint32_t i0 = b0 + threadIdx.x; i0 < b1; i0 += TS_U * TS_NT) {
        uint32_t k[TS_U];
#pragma unroll
        for (int u = 0; u < TS_U; ++u) {
            const uint32_t i = i0 + (uint32_t)u * TS_NT;
            k[u] = bucket[i < b1 ? i : i0].x;
        }
#pragma unroll
        for (int u = 0; u < TS_U; ++u)
            if (i0 + (uint32_t)u * TS_NT < b1) atomicAdd(&s_c[k[u] - k0], 1u);
    }
    __syncthreads();
    // exclusive scan of the T counts: thread r owns counts [r*C, (r+1)*C)
    const uint32_t C = T / TS_NT;
    uint32_t sum = 0, mx = 0;
    for (uint32_t c = 0; c < C; ++c) {
        const uint32_t v = s_c[threadIdx.x * C + c];
        sum += v;
        mx = v > mx ? v : mx;
    }
    const int lane = __lane_id(), wid = threadIdx.x >> 6;
    const uint32_t inc = wave_incl_scan<uint32_t>(sum);
    if (lane == 63) s_wave[wid] = inc;
    for (unsigned d = 32; d >= 1; d >>= 1) {
        const uint32_t o = __shfl_xor(mx, d);
        mx = o > mx ? o : mx;
    }
    if (lane == 0 && mx) atomicMax(&s_max, mx);
    __syncthreads();
    uint32_t run = b0 + inc - sum;
    for (int w = 0; w < wid; ++w) run += s_wave[w];
    for (uint32_t c = 0; c < C; ++c) {
        const uint32_t j = threadIdx.x * C + c;
        const uint32_t v = s_c[j];
        s_c[j] = run;
        if (k0 + j < nkeys) off[k0 + j] = run;
        run += v;
    }
    if (k0 + T >= nkeys && threadIdx.x == TS_NT - 1) off[nkeys] = b1;
    if (threadIdx.x == 0 && max_out && s_max) atomicMax(max_out, s_max);
    __syncthreads();
    for (uint32_t i0 = b0 + threadIdx.x; i0 < b1; i0 += TS_U * TS_NT) {
        uint4 it[TS_U];
        uint32_t pos[TS_U];
#pragma unroll
        for (int u = 0; u < TS_U; ++u) {
            const uint32_t i = i0 + (uint32_t)u * TS_NT;
            it[u] = bucket[i < b1 ? i : i0];
        }
#pragma unroll
        for (int u = 0; u < TS_U; ++u)
            pos[u] = i0 + (uint32_t)u * TS_NT < b1 ? atomicAdd(&s_c[it[u].x - k0], 1u) : 0u;
#pragma unroll
        for (int u = 0; u < TS_U; ++u) place(pos[u], it[u], i0 + (uint32_t)u * TS_NT < b1);
    }
}

// tile size 2^shift, 2^8 (TS_NT) to 2^13 keys: at most 8192 tiles below 2^26
// keys, 16384 at 2^27.  The count and bucket passes hold a counter per tile
// in every block's LDS (32 KiB at 8192 tiles: five blocks per CU, where
// 16384 tiles allowed two), the place pass one per key of its tile (at most
// 32 KiB).
int ts_shift(uint64_t nkeys)
{
    int lg = 0;
    while ((1ull << lg) < nkeys) ++lg;
    const int sh = lg - 13;
    return sh < 8 ? 8 : sh > 13 ? 13 : sh;
}

template <typename Src, typename Place>
static void ts_sort(Src src, uint64_t n, uint32_t nkeys, const TileSortTmp& tmp, int items_per_block, uint32_t* off,
                    unsigned int* max_out, Place place, hipStream_t s)
{
    const int shift = ts_shift(nkeys);
    const int ntiles = (int)((nkeys + (1u << shift) - 1) >> shift);
    if (ntiles == 0) {   // no keys, hence no items
        launch_fill_u32(off, 0, 1, s);
        return;
    }
    launch_fill_u32(tmp.total, 0, 2 * (int64_t)ntiles + 1, s);   // total[ntiles], fill[ntiles + 1]
    uint32_t* fill = tmp.total + ntiles;
    if (n > 0) {
        const uint64_t ipb = (uint64_t)items_per_block;
        const int blocks = (int)std::min<uint64_t>(2048, std::max<uint64_t>(1, (n + ipb - 1) / ipb));
        const size_t lds = (size_t)ntiles * 4;
        k_ts_count<Src><<<blocks, TS_NT, lds, s>>>(src, n, shift, ntiles, tmp.total);
        launch_scan_u32_u32(tmp.total, tmp.off, ntiles, tmp.scan, nullptr, s);
        k_ts_bucket<Src><<<blocks, TS_NT, lds, s>>>(src, n, shift, ntiles, tmp.off, fill, tmp.bucket);
    } else {
        launch_fill_u32(tmp.off, 0, (int64_t)ntiles + 1, s);
    }
    k_ts_place<Place><<<ntiles, TS_NT, (size_t)4 << shift, s>>>(tmp.bucket, tmp.off, shift, nkeys, off, max_out,
                                                              place);
}

size_t ts_tiles(uint64_t nkeys)
{
    const int shift = ts_shift(nkeys);
    return (size_t)((nkeys + (1u << shift) - 1) >> shift);
}

// A bin pair sorts by the rank of its voxel among the boundary voxels; its
// item carries the triangle and the Level2 box-axis masks: bit p of mask x
// says the SAT's x box-axis predicate holds for the triangle and the Level2
// sub-index p of that voxel (the same float test the full SAT runs), likewise
// y, z; packed x | y << 16 | z << 32.
template <bool kM>
struct PairSrc {
    // per pair, its voxel's bit in the Level1 grid; the count pass writes the
    // voxel's boundary rank over it (ranked: a repeated compaction, whose
    // pairs hold ranks already), so the bucket pass reads the rank directly
    uint64_t* bits;
    bool ranked;
    const uint32_t* tri;
    const uint64_t* bnd;
    const uint32_t* wprefix;
    const float* tris;
    const int32_t* bidx;
    GridDev g0;
    Meshes M;
    bool masks;
    __device__ uint32_t count_key(uint64_t i) const
    {
        return ranked ? (uint32_t)bits[i] : rank_of(bnd, wprefix, bits[i]);
    }
    __device__ void keep_key(uint64_t i, uint32_t k) const
    {
        if (!ranked) bits[i] = k;
    }
    __device__ uint32_t sort_key(uint64_t i) const { return (uint32_t)bits[i]; }
    __device__ uint4 item(uint64_t i) const
    {
        const uint32_t b = (uint32_t)bits[i];
        const uint32_t t = tri[i];
        const int mesh = kM ? (int)M.bmesh[b] : 0;
        const GridDev& g = grid_of<kM>(g0, M, mesh);
        const int n2 = g.n2;
        if (!masks || n2 == 0) return make_uint4(b, t, 0u, 0u);
        const int v = bidx[b];
        const int nxy = g.nx * g.ny;
        const int k1 = v / nxy;
        const int ij = v - k1 * nxy;
        const int j1 = ij / g.nx;
        const int i1 = ij - j1 * g.nx;
        const float* tv = tris + 9 * (uint64_t)t;
        const float a0 = tv[0], a1 = tv[1], a2 = tv[2], a3 = tv[3], a4 = tv[4], a5 = tv[5], a6 = tv[6],
                    a7 = tv[7], a8 = tv[8];
        const float* cx = g.c2x + (int64_t)i1 * n2;
        const float* cy = g.c2y + (int64_t)j1 * n2;
        const float* cz = g.c2z + (int64_t)k1 * n2;
        uint32_t mx = 0, my = 0, mz = 0;
        for (int p = 0; p < n2; ++p) {
            mx |= (uint32_t)sat_axis_ok(a0, a3, a6, cx[p], g.h2x) << p;
            my |= (uint32_t)sat_axis_ok(a1, a4, a7, cy[p], g.h2y) << p;
            mz |= (uint32_t)sat_axis_ok(a2, a5, a8, cz[p], g.h2z) << p;
        }
        return make_uint4(b, t, mx | (my << 16), mz);
    }
};

struct PairPlace {
    uint32_t* list;
    uint64_t* pmask;
    __device__ void operator()(uint32_t pos, const uint4& it, bool ok) const
    {
        if (!ok) return;
        list[pos] = it.y;
        if (pmask) pmask[pos] = (uint64_t)it.z | ((uint64_t)it.w << 32);
    }
};

void launch_bin_cols(const float* tris, const uint64_t* key, const uint32_t* zmax, uint64_t n, uint32_t ncol,
                     const TileSortTmp& tmp, int items_per_block, uint32_t* coff, float4* rec, hipStream_t s)
{
    ts_sort(ColSrc{key, zmax}, n, ncol, tmp, items_per_block, coff, nullptr, ColPlace{tris, rec}, s);
}

void launch_bin_pairs(uint64_t* bits, bool ranked, const uint32_t* tri, uint64_t n, const uint64_t* bnd,
                      const uint32_t* wprefix, uint32_t B, const TileSortTmp& tmp, int items_per_block,
                      uint32_t* boff, unsigned int* max_out, uint32_t* list, const float* tris,
                      const int32_t* bidx, const GridDev& g0, const Meshes& M, uint64_t* pmask, hipStream_t s)
{
    const PairPlace place{list, pmask};
    if (M.nm > 1)
        ts_sort(PairSrc<true>{bits, ranked, tri, bnd, wprefix, tris, bidx, g0, M, pmask != nullptr}, n, B, tmp,
                items_per_block, boff, max_out, place, s);
    else
        ts_sort(PairSrc<false>{bits, ranked, tri, bnd, wprefix, tris, bidx, g0, M, pmask != nullptr}, n, B, tmp,
                items_per_block, boff, max_out, place, s);
}

// Per-mesh bin totals (pairs) and largest bin of a batch.  bmesh is
// non-decreasing, so blocks inside one mesh reduce before one atomic.
__global__ __launch_bounds__(256) void k_mesh_bins(const uint32_t* __restrict__ boff, int64_t B, Meshes M)
{
    __shared__ uint32_t s_max[4];
    const int64_t first = (int64_t)blockIdx.x * blockDim.x;
    const int64_t b = first + threadIdx.x;
    const uint32_t c = b < B ? boff[b + 1] - boff[b] : 0u;
    const int64_t last = (first + (int64_t)blockDim.x < B ? first + (int64_t)blockDim.x : B) - 1;
    const uint32_t ma = M.bmesh[first], mb = M.bmesh[last];
    if (ma == mb) {
        uint32_t mx = c;
        for (unsigned d = 32; d >= 1; d >>= 1) {
            const uint32_t o = __shfl_xor(mx, d);
            mx = o > mx ? o : mx;
        }
        if (__lane_id() == 0) s_max[threadIdx.x >> 6] = mx;
        block_add_u64(c, &M.ctr[ma].pairs);   // synchronises the block
        if (threadIdx.x == 0) {
            uint32_t v = 0;
            for (int i = 0; i < (int)(blockDim.x >> 6); ++i) v = s_max[i] > v ? s_max[i] : v;
            if (v) atomicMax(&M.ctr[ma].max_per_voxel, v);
        }
    } else if (b < B && c) {
        const uint32_t mm = M.bmesh[b];
        atomicAdd(&M.ctr[mm].pairs, (unsigned long long)c);
        atomicMax(&M.ctr[mm].max_per_voxel, c);
    }
}

void launch_mesh_bins(const uint32_t* boff, int64_t B, const Meshes& M, hipStream_t s)
{
    if (B <= 0) return;
    k_mesh_bins<<<grid_for(B, 256, 1 << 30), 256, 0, s>>>(boff, B, M);
}

// ---------------------------------------------------------------------------
// Level2 SAT (src/Object.cpp:2344-2424).  Two forms: one wave per boundary
// voxel (k_l2_sat_voxel), and one wave per group of consecutive boundary
// voxels (k_l2_sat_grp: same mesh, first bin entries in one 64-entry chunk,
// at most L2_GMAX voxels and 64 Level2 words), so voxels with few bin entries
// share the wave's rounds and SAT batches.  Each bin entry carries its box-axis masks
// (the SAT's own box-axis predicate per sub-index, computed with the bins)
// and its triangle a plane slab (gpv_slab.hpp) valid over every Level2
// center of its index AABB.  The lanes take the (entry, sub-column) items its
// x/y masks allow, drop sub-columns whose e x z edge predicates separate, and
// keep the sub-voxels r whose float |N.(c - t0)| the slab admits (the bound
// is grown by that evaluation's error, so the SAT rejects every other one);
// those queue for the full float SAT in 64-lane batches.  Sub-voxels already
// set are skipped (the result is an OR).
// ---------------------------------------------------------------------------
// ring of queued (triangle, sub-voxel) entries per wave: a round pushes at
// most 64 entries onto fewer than 64 already queued, so it holds 128
constexpr int L2_QCAP = 128;
constexpr int L2_GMAX = kL2GroupMax;   // voxels per group
constexpr int L2_CMAX = kL2GroupCenters;   // center floats per group: 3 * voxels * n2

// 1/d for the sub-box widths d = 1..16 (float quotients, as 1.0f / d)
__constant__ float kInvSmall[17] = {0.0f,        1.0f,        0.5f,        1.0f / 3.0f, 0.25f,       0.2f,
                                    1.0f / 6.0f, 1.0f / 7.0f, 0.125f,      1.0f / 9.0f, 0.1f,        1.0f / 11.0f,
                                    1.0f / 12.0f, 1.0f / 13.0f, 1.0f / 14.0f, 1.0f / 15.0f, 0.0625f};

// value of a float register in lane `l` (l uniform across the wave)
__device__ __forceinline__ float lane_f(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// voxels of a group: the capacity for a mesh's resolution
__device__ __forceinline__ uint32_t l2_group_cap(int w2, int n2)
{
    return (uint32_t)l2_group_voxels(w2, n2);
}

// Group starts: flag[b] = 1 where a new group begins (mesh start, a new
// 64-entry chunk of first bin entries, or a full group).
template <bool kM>
__global__ void k_l2_group_flags(const uint32_t* __restrict__ boff, int64_t B, GridDev g0, Meshes M,
                                 uint32_t* __restrict__ flag)
{
    for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < B; b += (int64_t)gridDim.x * blockDim.x) {
        const int mesh = kM ? (int)M.bmesh[b] : 0;
        const GridDev& g = grid_of<kM>(g0, M, mesh);
        const int64_t lb = b - (int64_t)(kM ? M.off[mesh].b0 : 0u);
        const bool st =
            lb == 0 || lb % (int64_t)l2_group_cap(g.w2, g.n2) == 0 || (boff[b] >> 6) != (boff[b - 1] >> 6);
        flag[b] = st ? 1u : 0u;
    }
}

__global__ void k_l2_group_starts(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ gidx, int64_t B,
                                  uint32_t* __restrict__ gstart)
{
    for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < B; b += (int64_t)gridDim.x * blockDim.x) {
        if (flag[b]) gstart[gidx[b]] = (uint32_t)b;
        if (b == B - 1) gstart[gidx[B]] = (uint32_t)B;
    }
}

void launch_l2_groups(const uint32_t* boff, int64_t B, const GridDev& g0, const Meshes& M, uint32_t* flag,
                      uint32_t* gidx, uint32_t* scan_tmp, uint32_t* gstart, hipStream_t s)
{
    if (B <= 0) return;
    const int blocks = grid_for(B, 256, 16384);
    if (M.nm > 1)
        k_l2_group_flags<true><<<blocks, 256, 0, s>>>(boff, B, g0, M, flag);
    else
        k_l2_group_flags<false><<<blocks, 256, 0, s>>>(boff, B, g0, M, flag);
    launch_scan_u32_u32(flag, gidx, B, scan_tmp, nullptr, s);
    k_l2_group_starts<<<blocks, 256, 0, s>>>(flag, gidx, B, gstart);
}

// LDS of one wave of the Level2 SAT
struct L2Lds {
    unsigned long long acc[64];   // Level2 boundary words of the group's voxels
    float tv[64 * 9];             // staged triangles
    uint32_t q[L2_QCAP];          // ring of queued SAT entries
    uint32_t dec[64];             // staged entry: sub-box origin/extent and box bit
    unsigned long long mk[64];    // staged entry: box-axis masks
    float4 sl[64];                // staged entry: Level2 slab
    uint32_t off[64];             // staged entry: first item in the chunk's item space
    uint8_t ev[64];               // staged entry: its voxel in the group
    uint8_t win[64];              // per round: staged entry starting at each item position
    uint32_t vb[L2_GMAX + 1];     // group voxel v: first bin entry
    float c[L2_CMAX];             // centers: x [v*n2+p], y [nv*n2 + ...], z [2*nv*n2 + ...]
};

// Full float SAT for queued entries (one per lane); entry = p | q<<4 | r<<8 |
// j<<12 | box<<18 where j is the staged triangle (its voxel is ev[j]) and box
// says the box-axis predicates are already known to hold.
__device__ __forceinline__ void l2_sat_run(uint32_t ent, bool valid, const L2Lds& L, unsigned long long* acc, int n2,
                                           int nvn2, int w2, float h2x, float h2y, float h2z)
{
    const int p = ent & 15, q = (ent >> 4) & 15, r = (ent >> 8) & 15, j = (ent >> 12) & 63;
    const bool box = (ent >> 18) & 1;
    const int v = valid ? (int)L.ev[j] : 0;
    const int s = (r * n2 + q) * n2 + p;
    const int wi = v * w2 + (s >> 6);
    if (!valid || ((acc[wi] >> (s & 63)) & 1ull)) return;
    const float ccx = L.c[v * n2 + p], ccy = L.c[nvn2 + v * n2 + q], ccz = L.c[2 * nvn2 + v * n2 + r];
    float tv[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) tv[c] = L.tv[j * 9 + c];
    // every queued entry's column passed xy_edges_ok
    const bool hit = box ? tri_box_overlap<false, false>(ccx, ccy, ccz, h2x, h2y, h2z, tv)
                         : tri_box_overlap<true, false>(ccx, ccy, ccz, h2x, h2y, h2z, tv);
    if (hit) atomicOr(&acc[wi], 1ull << (s & 63));
}

template <bool kM, int W>
__global__ __launch_bounds__(64 * W) void k_l2_sat_grp(const float* __restrict__ tris,
                                                          const int32_t* __restrict__ bidx,
                                                          const uint32_t* __restrict__ boff,
                                                          const uint32_t* __restrict__ list,
                                                          const uint64_t* __restrict__ pmask,
                                                          const float4* __restrict__ slab2,
                                                          const uint32_t* __restrict__ gstart,
                                                          const uint32_t* __restrict__ ngroups, GridDev g0,
                                                          Meshes M, uint64_t* __restrict__ sat2)
{
    __shared__ L2Lds s_l[W];
    const int wv = threadIdx.x >> 6;
    const int lane = __lane_id();
    L2Lds& L = s_l[wv];
    unsigned long long* acc = L.acc;
    const uint32_t G = *ngroups;
    for (uint32_t grp = blockIdx.x * W + wv; grp < G; grp += gridDim.x * W) {
        const uint32_t vb0 = gstart[grp];
        const int nv = (int)(gstart[grp + 1] - vb0);
        const int mesh = kM ? uniform_i((int)M.bmesh[vb0]) : 0;
        const GridDev& g = grid_of<kM>(g0, M, mesh);
        const int n2 = g.n2, w2 = g.w2, gnx = g.nx, gny = g.ny;
        const float h2x = g.h2x, h2y = g.h2y, h2z = g.h2z;
        if (kM && n2 == 0) continue;
        const int nvn2 = nv * n2;
        uint64_t* out =
            sat2 + (kM ? M.off[mesh].l2w0 + (uint64_t)(vb0 - M.off[mesh].b0) * w2 : (uint64_t)vb0 * w2);
        for (int i = lane; i < nv * w2; i += 64) acc[i] = 0ull;
        if (lane <= nv) L.vb[lane] = boff[vb0 + lane];
        // the group's Level2 centers: x, y, z tables of each voxel
        for (int i = lane; i < nvn2; i += 64) {
            const int v = i / n2, pp = i - v * n2;
            const int vox = bidx[vb0 + v];
            const int nxy = gnx * gny;
            const int k1 = vox / nxy;
            const int ij = vox - k1 * nxy;
            const int j1 = ij / gnx;
            const int i1 = ij - j1 * gnx;
            L.c[i] = g.c2x[(int64_t)i1 * n2 + pp];
            L.c[nvn2 + i] = g.c2y[(int64_t)j1 * n2 + pp];
            L.c[2 * nvn2 + i] = g.c2z[(int64_t)k1 * n2 + pp];
        }
        __builtin_amdgcn_wave_barrier();
        uint32_t qh = 0, qt = 0;   // ring head/tail (uniform)
        const uint32_t e0 = L.vb[0], e1 = L.vb[nv];
        for (uint32_t base = e0; base < e1; base += 64) {
            // the queue refers to staged triangles by slot: drain it before restaging
            while (qt != qh) {
                const uint32_t n = qt - qh < 64 ? qt - qh : 64;
                const bool valid = (uint32_t)lane < n;
                const uint32_t ent = valid ? L.q[(qh + lane) % L2_QCAP] : 0u;
                l2_sat_run(ent, valid, L, acc, n2, nvn2, w2, h2x, h2y, h2z);
                qh += n;
            }
            __builtin_amdgcn_wave_barrier();
            // staging: lane j owns bin entry base + j -- vertices, masks, slab,
            // sub-box decode and its voxel go to LDS
            const uint32_t i = base + lane;
            uint32_t P = 0;
            if (i < e1) {
                const uint64_t t = list[i];
#pragma unroll
                for (int c = 0; c < 9; ++c) L.tv[lane * 9 + c] = tris[t * 9 + c];
                int lo = 0, hi = nv;   // the voxel whose entries hold i: vb[lo] <= i < vb[lo + 1]
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (L.vb[mid] <= i) lo = mid;
                    else hi = mid;
                }
                L.ev[lane] = (uint8_t)lo;
                const uint64_t mk = pmask[i];
                const uint32_t mx = (uint32_t)(mk & 0xffffu), my = (uint32_t)((mk >> 16) & 0xffffu),
                               mz = (uint32_t)((mk >> 32) & 0xffffu);
                if (mx && my && mz) {
                    const int pa = __ffs(mx) - 1, pb = 31 - __clz(mx);
                    const int qa = __ffs(my) - 1, qb = 31 - __clz(my);
                    const int ra = __ffs(mz) - 1, rb = 31 - __clz(mz);
                    const int dp = pb - pa + 1, dq = qb - qa + 1, dr = rb - ra + 1;
                    // contiguous passing sets (centers are monotone in the index) mean
                    // the box-axis predicates hold on the whole sub-box
                    const uint32_t boxbit =
                        (__popc(mx) == dp && __popc(my) == dq && __popc(mz) == dr) ? 1u : 0u;
                    L.dec[lane] = (uint32_t)pa | ((uint32_t)qa << 4) | ((uint32_t)ra << 8) | ((uint32_t)rb << 12) |
                                  ((uint32_t)dp << 16) | (boxbit << 24);
                    L.mk[lane] = mk;
                    L.sl[lane] = slab2[t];
                    P = (uint32_t)(dp * dq);
                }
            }
            const uint32_t incl = wave_incl_scan<uint32_t>(P);
            const uint32_t so = incl - P;   // this lane's entry: items [so, so + P)
            L.off[lane] = so;
            const uint32_t S = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
            __builtin_amdgcn_wave_barrier();
            // the chunk's (entry, sub-column) items, 64 per round
            for (uint32_t l0 = 0; l0 < S; l0 += 64) {
                // each item's entry, as in k_l2_sat_voxel: entries mark their
                // first position in the window, an item takes the highest mark
                // at or below it
                L.win[lane] = 0xff;
                __builtin_amdgcn_wave_barrier();
                if (P && so < l0 + 64u && so + P > l0) L.win[so > l0 ? so - l0 : 0u] = (uint8_t)lane;
                __builtin_amdgcn_wave_barrier();
                const uint32_t wj = L.win[lane];
                const uint64_t marks = __ballot(wj != 0xffu) & ((2ull << lane) - 1ull);
                const int top = marks ? 63 - __clzll((long long)marks) : 0;
                const int jw = __shfl((int)wj, top);
                __builtin_amdgcn_wave_barrier();
                const uint32_t l = l0 + (uint32_t)lane;
                uint32_t keep = 0;
                int p = 0, q = 0, j = 0, v = 0, ra = 0, rb = -1;
                uint32_t boxbit = 0, mz = 0;
                if (l < S) {
                    j = jw;
                    v = (int)L.ev[j];
                    const uint32_t d = L.dec[j];
                    const int pa = d & 15, qa = (d >> 4) & 15;
                    ra = (d >> 8) & 15;
                    rb = (d >> 12) & 15;
                    const int dp = (d >> 16) & 255;
                    boxbit = (d >> 24) & 1u;
                    const int lj = (int)(l - L.off[j]);
                    // exact small-integer division: (lj + 0.5)/dp is >= 1/32 from an integer
                    const int rem = (int)(((float)lj + 0.5f) * kInvSmall[dp]);
                    p = pa + lj - rem * dp;
                    q = qa + rem;
                    const uint64_t mk = L.mk[j];
                    const uint32_t mx = (uint32_t)(mk & 0xffffu), my = (uint32_t)((mk >> 16) & 0xffffu);
                    mz = (uint32_t)((mk >> 32) & 0xffffu);
                    if ((mx >> p) & (my >> q) & 1u) {
                        const float ccx = L.c[v * n2 + p], ccy = L.c[nvn2 + v * n2 + q];
                        // the e x z edge predicates hold or fail for the whole sub-column
                        if (xy_edges_ok(ccx, ccy, h2x, h2y, L.tv + j * 9)) {
                            const float4 sl = L.sl[j];
                            const float t0x = L.tv[j * 9 + 0], t0y = L.tv[j * 9 + 1], t0z = L.tv[j * 9 + 2];
                            // float_slab_keeps, with the (x, y) part shared by the sub-column
                            const float A = sl.x * (ccx - t0x) + sl.y * (ccy - t0y);
                            const float* cz = L.c + 2 * nvn2 + v * n2;
                            for (int r = ra; r <= rb; ++r) {
                                const float sd = A + sl.z * (cz[r] - t0z);
                                keep |= (uint32_t)(fabsf(sd) <= sl.w && ((mz >> r) & 1u)) << r;
                            }
                        }
                    }
                }
                // queue the kept sub-voxels not yet set, one r per lane per round
                while (__ballot(keep != 0)) {
                    const int r = keep ? __ffs(keep) - 1 : 0;
                    const bool has = keep != 0;
                    keep &= keep - 1;
                    const int sv = (r * n2 + q) * n2 + p;
                    const bool push = has && !((acc[(v * w2 + (sv >> 6)) & 63] >> (sv & 63)) & 1ull);
                    const uint64_t m = __ballot(push);
                    if (push) {
                        const uint32_t at = qt + (uint32_t)__popcll(m & lanes_below_mask());
                        L.q[at % L2_QCAP] = (uint32_t)p | ((uint32_t)q << 4) | ((uint32_t)r << 8) |
                                            ((uint32_t)j << 12) | (boxbit << 18);
                    }
                    qt += (uint32_t)__popcll(m);
                    __builtin_amdgcn_wave_barrier();
                    if (qt - qh >= 64) {   // a full wave of work
                        const uint32_t ent = L.q[(qh + lane) % L2_QCAP];
                        l2_sat_run(ent, true, L, acc, n2, nvn2, w2, h2x, h2y, h2z);
                        qh += 64;
                        __builtin_amdgcn_wave_barrier();
                    }
                }
            }
        }
        while (qt != qh) {
            const uint32_t n = qt - qh < 64 ? qt - qh : 64;
            const bool valid = (uint32_t)lane < n;
            const uint32_t ent = valid ? L.q[(qh + lane) % L2_QCAP] : 0u;
            l2_sat_run(ent, valid, L, acc, n2, nvn2, w2, h2x, h2y, h2z);
            qh += n;
        }
        __builtin_amdgcn_wave_barrier();
        for (int i = lane; i < nv * w2; i += 64) out[i] = acc[i];
        __builtin_amdgcn_wave_barrier();
    }
}

// One wave per boundary voxel: the form for voxels with many bin entries
// (the parent's sub-centers stay in lane registers, entry p in lane p).
// Full float SAT for queued entries (one per lane); entry = p | q<<4 | r<<8 |
// j<<12 | box<<18 where j is the staged triangle and box says the box-axis
// predicates are already known to hold.
__device__ __forceinline__ void l2_sat_run_v(uint32_t ent, bool valid, const float* stv, float cxl, float cyl,
                                           float czl, int n2, float h2x, float h2y, float h2z,
                                           unsigned long long* acc)
{
    const int p = ent & 15, q = (ent >> 4) & 15, r = (ent >> 8) & 15, j = (ent >> 12) & 63;
    const bool box = (ent >> 18) & 1;
    const float ccx = __shfl(cxl, p), ccy = __shfl(cyl, q), ccz = __shfl(czl, r);
    const int s = (r * n2 + q) * n2 + p;
    if (!valid || ((acc[s >> 6] >> (s & 63)) & 1ull)) return;
    float tv[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) tv[c] = stv[j * 9 + c];
    // every queued entry's column passed xy_edges_ok
    const bool hit = box ? tri_box_overlap<false, false>(ccx, ccy, ccz, h2x, h2y, h2z, tv)
                         : tri_box_overlap<true, false>(ccx, ccy, ccz, h2x, h2y, h2z, tv);
    if (hit) atomicOr(&acc[s >> 6], 1ull << (s & 63));
}

// Eight waves per SIMD: at most 64 VGPRs and 4.7 KB of LDS per wave (the
// staged entries' masks and item offsets stay in their lanes' registers and
// are read across lanes).
template <bool kM, int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(8))) void k_l2_sat_voxel(
    const float* __restrict__ tris, const int32_t* __restrict__ bidx, int64_t B, const uint32_t* __restrict__ boff,
    const uint32_t* __restrict__ list, const uint64_t* __restrict__ pmask, const float4* __restrict__ slab2,
    GridDev g0, Meshes M, uint64_t* __restrict__ sat2)
{
    __shared__ unsigned long long s_acc[W][64];
    __shared__ float s_tv[W][64 * 9];
    __shared__ uint32_t s_q[W][L2_QCAP];
    // per staged triangle: sub-box origin and extents, and its slab (its
    // masks and the start of its sub-columns in the chunk's flattened
    // (triangle, sub-column) item space stay in the staging lane)
    __shared__ uint32_t s_dec[W][64];
    __shared__ float4 s_sl[W][64];
    __shared__ float s_cz[W][16];
    __shared__ uint8_t s_win[W][64];   // per round: staged entry starting at each item position
    const int wv = threadIdx.x >> 6;
    const int lane = __lane_id();
    const int64_t b = (int64_t)blockIdx.x * W + wv;
    if (b >= B) return;
    const int mesh = kM ? uniform_i((int)M.bmesh[b]) : 0;
    const GridDev& g = grid_of<kM>(g0, M, mesh);
    const int n2 = g.n2, w2 = g.w2, gnx = g.nx, gny = g.ny;
    const float h2x = g.h2x, h2y = g.h2y, h2z = g.h2z;
    if (kM && n2 == 0) return;
    uint64_t* out = sat2 + (kM ? M.off[mesh].l2w0 + (uint64_t)(b - M.off[mesh].b0) * w2 : (uint64_t)b * w2);
    unsigned long long* acc = s_acc[wv];
    for (int i = lane; i < w2; i += 64) acc[i] = 0ull;
    const int v = bidx[b];
    const int nxy = gnx * gny;
    const int k1 = v / nxy;
    const int ij = v - k1 * nxy;
    const int j1 = ij / gnx;
    const int i1 = ij - j1 * gnx;
    // the parent's n2 sub-centers per axis in registers, entry p in lane p;
    // z also in LDS for per-lane layer lookups
    const float cxl = lane < n2 ? g.c2x[(int64_t)i1 * n2 + lane] : 0.0f;
    const float cyl = lane < n2 ? g.c2y[(int64_t)j1 * n2 + lane] : 0.0f;
    const float czl = lane < n2 ? g.c2z[(int64_t)k1 * n2 + lane] : 0.0f;
    float* cz = s_cz[wv];
    if (lane < n2) cz[lane] = czl;
    float* stv = s_tv[wv];
    uint32_t* sq = s_q[wv];
    uint32_t* dec = s_dec[wv];
    float4* ssl = s_sl[wv];
    uint8_t* win = s_win[wv];
    __builtin_amdgcn_wave_barrier();
    uint32_t qh = 0, qt = 0;   // ring head/tail (uniform)
    const uint32_t e0 = boff[b], e1 = boff[b + 1];
    for (uint32_t base = e0; base < e1; base += 64) {
        // the queue refers to staged triangles by slot: drain it before restaging
        while (qt != qh) {
            const uint32_t n = qt - qh < 64 ? qt - qh : 64;
            const bool valid = (uint32_t)lane < n;
            const uint32_t ent = valid ? sq[(qh + lane) % L2_QCAP] : 0u;
            l2_sat_run_v(ent, valid, stv, cxl, cyl, czl, n2, h2x, h2y, h2z, acc);
            qh += n;
        }
        __builtin_amdgcn_wave_barrier();
        // staging: lane j owns bin entry base + j -- vertices, masks, slab and
        // its sub-box decode go to LDS
        const uint32_t i = base + lane;
        uint32_t P = 0;
        uint64_t mk_reg = 0;
        if (i < e1) {
            const uint64_t t = list[i];
#pragma unroll
            for (int c = 0; c < 9; ++c) stv[lane * 9 + c] = tris[t * 9 + c];
            const uint64_t mk = pmask[i];
            const uint32_t mx = (uint32_t)(mk & 0xffffu), my = (uint32_t)((mk >> 16) & 0xffffu),
                           mz = (uint32_t)((mk >> 32) & 0xffffu);
            if (mx && my && mz) {
                const int pa = __ffs(mx) - 1, pb = 31 - __clz(mx);
                const int qa = __ffs(my) - 1, qb = 31 - __clz(my);
                const int ra = __ffs(mz) - 1, rb = 31 - __clz(mz);
                const int dp = pb - pa + 1, dq = qb - qa + 1, dr = rb - ra + 1;
                // contiguous passing sets (centers are monotone in the index) mean
                // the box-axis predicates hold on the whole sub-box
                const uint32_t boxbit =
                    (__popc(mx) == dp && __popc(my) == dq && __popc(mz) == dr) ? 1u : 0u;
                dec[lane] = (uint32_t)pa | ((uint32_t)qa << 4) | ((uint32_t)ra << 8) | ((uint32_t)rb << 12) |
                            ((uint32_t)dp << 16) | (boxbit << 24);
                mk_reg = mk;
                ssl[lane] = slab2[t];
                P = (uint32_t)(dp * dq);
            }
        }
        const uint32_t incl = wave_incl_scan<uint32_t>(P);
        const uint32_t so = incl - P;   // this lane's entry: items [so, so + P)
        const uint32_t S = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
        __builtin_amdgcn_wave_barrier();
        // the chunk's (triangle, sub-column) items, 64 per round
        for (uint32_t l0 = 0; l0 < S; l0 += 64) {
            // each item's entry: every entry with items in [l0, l0 + 64) marks
            // its first position there (the one running into the window marks
            // 0; starts are distinct), and item l takes the mark at the highest
            // marked position <= l
            win[lane] = 0xff;
            __builtin_amdgcn_wave_barrier();
            if (P && so < l0 + 64u && so + P > l0) win[so > l0 ? so - l0 : 0u] = (uint8_t)lane;
            __builtin_amdgcn_wave_barrier();
            const uint32_t wj = win[lane];
            const uint64_t marks = __ballot(wj != 0xffu) & ((2ull << lane) - 1ull);
            const int top = marks ? 63 - __clzll((long long)marks) : 0;
            const int jw = __shfl((int)wj, top);
            // the item's entry fields, read across lanes by every lane
            const uint32_t so_j = (uint32_t)__shfl((int)so, jw);
            const uint32_t mklo_j = (uint32_t)__shfl((int)(uint32_t)mk_reg, jw);
            const uint32_t mkhi_j = (uint32_t)__shfl((int)(uint32_t)(mk_reg >> 32), jw);
            __builtin_amdgcn_wave_barrier();
            const uint32_t l = l0 + (uint32_t)lane;
            uint32_t keep = 0;
            int p = 0, q = 0, j = 0, ra = 0, rb = -1;
            uint32_t boxbit = 0, mz = 0;
            bool in = false;
            if (l < S) {
                j = jw;
                const uint32_t d = dec[j];
                const int pa = d & 15, qa = (d >> 4) & 15;
                ra = (d >> 8) & 15;
                rb = (d >> 12) & 15;
                const int dp = (d >> 16) & 255;
                boxbit = (d >> 24) & 1u;
                const int lj = (int)(l - so_j);
                // exact small-integer division: (lj + 0.5)/dp is >= 1/32 from an integer
                const int rem = (int)(((float)lj + 0.5f) * kInvSmall[dp]);
                p = pa + lj - rem * dp;
                q = qa + rem;
                const uint32_t mx = mklo_j & 0xffffu, my = mklo_j >> 16;
                mz = mkhi_j & 0xffffu;
                in = ((mx >> p) & (my >> q) & 1u) != 0;
            }
            // every lane takes part in the cross-lane reads
            const float ccx = __shfl(cxl, p), ccy = __shfl(cyl, q);
            // the e x z edge predicates hold or fail for the whole sub-column
            if (in) in = xy_edges_ok(ccx, ccy, h2x, h2y, stv + j * 9);
            if (in) {
                const float4 sl = ssl[j];
                const float t0x = stv[j * 9 + 0], t0y = stv[j * 9 + 1], t0z = stv[j * 9 + 2];
                // float_slab_keeps, with the (x, y) part shared by the sub-column
                const float A = sl.x * (ccx - t0x) + sl.y * (ccy - t0y);
                for (int r = ra; r <= rb; ++r) {
                    const float sd = A + sl.z * (cz[r] - t0z);
                    keep |= (uint32_t)(fabsf(sd) <= sl.w && ((mz >> r) & 1u)) << r;
                }
            }
            // queue the kept sub-voxels not yet set, one r per lane per round
            while (__ballot(keep != 0)) {
                const int r = keep ? __ffs(keep) - 1 : 0;
                const bool has = keep != 0;
                keep &= keep - 1;
                const int sv = (r * n2 + q) * n2 + p;
                const bool push = has && !((acc[(sv & 4095) >> 6] >> (sv & 63)) & 1ull);
                const uint64_t m = __ballot(push);
                if (push) {
                    const uint32_t at = qt + (uint32_t)__popcll(m & lanes_below_mask());
                    sq[at % L2_QCAP] = (uint32_t)p | ((uint32_t)q << 4) | ((uint32_t)r << 8) |
                                       ((uint32_t)j << 12) | (boxbit << 18);
                }
                qt += (uint32_t)__popcll(m);
                __builtin_amdgcn_wave_barrier();
                if (qt - qh >= 64) {   // a full wave of work
                    const uint32_t ent = sq[(qh + lane) % L2_QCAP];
                    l2_sat_run_v(ent, true, stv, cxl, cyl, czl, n2, h2x, h2y, h2z, acc);
                    qh += 64;
                    __builtin_amdgcn_wave_barrier();
                }
            }
        }
    }
    while (qt != qh) {
        const uint32_t n = qt - qh < 64 ? qt - qh : 64;
        const bool valid = (uint32_t)lane < n;
        const uint32_t ent = valid ? sq[(qh + lane) % L2_QCAP] : 0u;
        l2_sat_run_v(ent, valid, stv, cxl, cyl, czl, n2, h2x, h2y, h2z, acc);
        qh += n;
    }
    __builtin_amdgcn_wave_barrier();
    for (int i = lane; i < w2; i += 64) out[i] = acc[i];
}

template <int W>
static void launch_l2_sat_w(const float* tris, const int32_t* bidx, int64_t B, const uint32_t* boff,
                            const uint32_t* list, const uint64_t* pmask, const float4* slab2, const uint32_t* gstart,
                            const uint32_t* ngroups, int64_t gmax, const GridDev& g0, const Meshes& M,
                            uint64_t* sat2, hipStream_t s)
{
    const dim3 bs(64 * W);
    if (gstart) {
        // one wave per group, for an upper bound gmax on the group count
        const unsigned blocks = (unsigned)std::max<int64_t>(1, (gmax + W - 1) / W);
        if (M.nm > 1)
            k_l2_sat_grp<true, W><<<blocks, bs, 0, s>>>(tris, bidx, boff, list, pmask, slab2, gstart, ngroups, g0, M,
                                                        sat2);
        else
            k_l2_sat_grp<false, W><<<blocks, bs, 0, s>>>(tris, bidx, boff, list, pmask, slab2, gstart, ngroups, g0, M,
                                                         sat2);
        return;
    }
    const unsigned vblocks = (unsigned)((B + W - 1) / W);
    if (M.nm > 1)
        k_l2_sat_voxel<true, W><<<vblocks, bs, 0, s>>>(tris, bidx, B, boff, list, pmask, slab2, g0, M, sat2);
    else
        k_l2_sat_voxel<false, W><<<vblocks, bs, 0, s>>>(tris, bidx, B, boff, list, pmask, slab2, g0, M, sat2);
}

void launch_l2_sat(const float* tris, const int32_t* bidx, int64_t B, const uint32_t* boff, const uint32_t* list,
                   const uint64_t* pmask, const float4* slab2, const uint32_t* gstart, const uint32_t* ngroups,
                   int64_t gmax, int waves, const GridDev& g0, const Meshes& M, uint64_t* sat2, hipStream_t s)
{
    if (B <= 0) return;
    if (waves <= 1)
        launch_l2_sat_w<1>(tris, bidx, B, boff, list, pmask, slab2, gstart, ngroups, gmax, g0, M, sat2, s);
    else if (waves == 2)
        launch_l2_sat_w<2>(tris, bidx, B, boff, list, pmask, slab2, gstart, ngroups, gmax, g0, M, sat2, s);
    else
        launch_l2_sat_w<4>(tris, bidx, B, boff, list, pmask, slab2, gstart, ngroups, gmax, g0, M, sat2, s);
}

// ---------------------------------------------------------------------------
// Level2 ray parity (src/Object.cpp:1111-1207): one wave per boundary voxel
// (i1, j1, k1).  Candidates are the column's (triangle, zmax) entries with
// zmax >= k1 -- the deduplicated union of the bins at z' >= k1 -- each with
// a precomputed MT record (edges, 1/det, V1).  Per chunk of 64 candidates:
// one coalesced load into LDS; the z-independent det/u/v test runs once per
// (candidate, sub-column) and leaves a pass bit; t > 1e-6 then runs per
// sub-voxel only for the passing pairs.
// ---------------------------------------------------------------------------
struct ColRec {
    float e1x, e1y, e1z, e2x, e2y, e2z, inv, v1x, v1y, v1z;
    int zmax;
};

__device__ __forceinline__ ColRec load_colrec(const float4* __restrict__ rec, uint32_t i)
{
    const float4 a = rec[3 * (uint64_t)i + 0];
    const float4 b = rec[3 * (uint64_t)i + 1];
    const float4 c = rec[3 * (uint64_t)i + 2];
    ColRec r;
    r.e1x = a.x; r.e1y = a.y; r.e1z = a.z; r.e2x = a.w;
    r.e2y = b.x; r.e2z = b.y; r.inv = b.z; r.v1x = b.w;
    r.v1y = c.x; r.v1z = c.y; r.zmax = __float_as_int(c.z);
    return r;
}

__device__ __forceinline__ bool colrec_xy(const ColRec& r, float ox, float oy, float& tx, float& ty)
{
    MtTri m;
    m.e1x = r.e1x; m.e1y = r.e1y; m.e1z = r.e1z;
    m.e2x = r.e2x; m.e2y = r.e2y; m.e2z = r.e2z;
    m.inv_det = r.inv;
    return mt_column(m, r.v1x, r.v1y, ox, oy, tx, ty);
}

// Lane l holds row l of a 64x64 bit matrix (bit j = entry (l, j)); returns
// column l (bit j = entry (j, l)).  Six block swaps of halving size.
__device__ __forceinline__ uint64_t wave_transpose64(uint64_t v)
{
    const int lane = __lane_id();
    const uint64_t masks[6] = {0x00000000FFFFFFFFull, 0x0000FFFF0000FFFFull, 0x00FF00FF00FF00FFull,
                               0x0F0F0F0F0F0F0F0Full, 0x3333333333333333ull, 0x5555555555555555ull};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int s = 32 >> k;
        const uint64_t m = masks[k];
        const uint64_t x = shfl64(v, lane ^ s);
        // branch-free select: the upper lane of each pair keeps the ~m half
        const uint64_t sel = (uint64_t)0 - (uint64_t)((lane >> (5 - k)) & 1);
        const uint64_t lo = (v & m) | ((x & m) << s);
        const uint64_t hi = (v & ~m) | ((x & ~m) >> s);
        v = (lo & ~sel) | (hi & sel);
    }
    return v;
}

// The z-independent half of the MT test (det/u/v, src/TriRayIntersection.cpp:
// 78-131 with D = +z) depends on the column entry and the sub-column's (x, y)
// only, so it runs once per (column entry, sub-column) here: one lane per
// column entry, looping over its column's n2^2 sub-columns
// (sc = q*n2 + p); entry e's pass bits go to cpass[e * stride + sc / 64], bit
// sc % 64.  Every boundary voxel of the column then reads them instead of
// repeating the test.  The entry's last word, cpass[e * stride + stride - 1],
// is its header: bits 0-31 hold kh + 1 and bits 32-63 zc + 1, where zc is
// the last layer whose voxels take the entry as a candidate (zmax, or the
// top layer for the CUDA kernels' whole-column candidates) and kh < zc a
// layer such that t > 1e-6 holds for every ray origin of the column in
// layers [0, kh] (t_decided over that box; -1 if none found).  For a voxel at
// k1 <= kh the entry then flips every layer of each sub-column it passes
// without a t test.
// The column entries are packed 64 to a wave across columns (one wave per
// column would leave most lanes idle: columns hold a few dozen entries).
// Each lane takes its column from its record and its sub-column origins from
// the tables.  In a batch a wave's entries can span meshes (they are in
// column order, and each mesh's columns are contiguous): the wave takes its
// meshes one after another, each with the lanes whose columns it holds.
__device__ __forceinline__ void col_pass_entry(const GridDev& g, uint32_t col, const ColRec& r, bool own,
                                               uint64_t* __restrict__ dst, int stride)
{
    const int n2 = g.n2;
    if (n2 == 0) return;
    const int i1 = (int)(col % (uint32_t)g.nx), j1 = (int)(col / (uint32_t)g.nx);
    const float* rx = g.r2x + (int64_t)i1 * n2;
    const float* ry = g.r2y + (int64_t)j1 * n2;
    {
        const int zc = g.l2cuda ? g.nz - 1 : r.zmax;
        int kh = -1;
        if (g.r2mono != 0) {
            const float oxlo = rx[0], oxhi = rx[n2 - 1], oylo = ry[0], oyhi = ry[n2 - 1];
            const float oz0 = g.r2z[0];
#pragma unroll
            for (int d = 8; d >= 1; d >>= 1) {
                const int k = r.zmax - d;
                if (k >= 0 && t_decided(r.e1x, r.e1y, r.e1z, r.e2x, r.e2y, r.e2z, r.inv, r.v1x, r.v1y, r.v1z, oxlo, oxhi,
                                        oylo, oyhi, oz0, g.r2z[(int64_t)k * n2 + n2 - 1], eps_above()) > 0)
                    kh = k;
            }
        }
        if (own) dst[stride - 1] = (uint64_t)(uint32_t)(kh + 1) | ((uint64_t)(uint32_t)(zc + 1) << 32);
    }
    // the lane's sub-column x origins in registers
    float oxr[16];
#pragma unroll
    for (int p = 0; p < 16; ++p) oxr[p] = p < n2 ? rx[p] : 0.0f;
    uint64_t w = 0;
    int k = 0, slot = 0;
    for (int q = 0; q < n2; ++q) {
        const float oy = ry[q];
#pragma unroll
        for (int p = 0; p < 16; ++p) {
            if (p < n2) {
                float tx, ty;
                w |= (uint64_t)colrec_xy(r, oxr[p], oy, tx, ty) << k;
                if (++k == 64) {
                    if (own) dst[slot] = w;
                    ++slot;
                    w = 0;
                    k = 0;
                }
            }
        }
    }
    if (k && own) dst[slot] = w;
}

template <bool kM, int CP_WAVES>
__global__ __launch_bounds__(64 * CP_WAVES) void k_col_pass(const float4* __restrict__ rec, GridDev g0,
                                                                   Meshes M, uint32_t nent, int stride,
                                                                   uint64_t* __restrict__ cpass)
{
    const uint64_t wbase = ((uint64_t)blockIdx.x * CP_WAVES + (threadIdx.x >> 6)) * 64;
    if (wbase >= nent) return;
    const int lane = __lane_id();
    const uint32_t e = (uint32_t)wbase + (uint32_t)lane;
    const bool own = e < nent;
    const uint32_t ei = own ? e : nent - 1;
    const ColRec r = load_colrec(rec, ei);
    const uint32_t colb = __float_as_uint(rec[3 * (uint64_t)ei + 2].w);   // batch-wide column
    uint64_t* dst = cpass + (uint64_t)ei * stride;
    if constexpr (!kM) {
        col_pass_entry(g0, colb, r, own, dst, stride);
    } else {
        const int lastl = nent - 1 - wbase < 63 ? (int)(nent - 1 - wbase) : 63;
        const uint32_t cf = (uint32_t)__shfl((int)colb, 0), cl = (uint32_t)__shfl((int)colb, lastl);
        auto col0 = [](const MeshOff& o) { return (uint64_t)o.col0; };
        const int mf = wave_mesh_search(M.off, M.nm, cf, col0), ml = wave_mesh_search(M.off, M.nm, cl, col0);
        for (int m = mf; m <= ml; ++m) {
            const uint32_t c0 = M.off[m].col0, c1 = M.off[m + 1].col0;
            if (colb >= c0 && colb < c1) col_pass_entry(M.grids[m], colb - c0, r, own, dst, stride);
        }
    }
}

template <int CW>
static void launch_col_pass_w(const float4* rec, const GridDev& g0, const Meshes& M, uint64_t nent, int stride,
                              uint64_t* cpass, hipStream_t s)
{
    const uint64_t waves = (nent + 63) / 64;
    const uint64_t blocks = (waves + CW - 1) / CW;
    if (M.nm > 1)
        k_col_pass<true, CW><<<(unsigned)blocks, 64 * CW, 0, s>>>(rec, g0, M, (uint32_t)nent, stride, cpass);
    else
        k_col_pass<false, CW><<<(unsigned)blocks, 64 * CW, 0, s>>>(rec, g0, M, (uint32_t)nent, stride, cpass);
}

void launch_col_pass(const float4* rec, const GridDev& g0, const Meshes& M, uint64_t nent, int stride, int waves,
                     uint64_t* cpass, hipStream_t s)
{
    if (nent == 0) return;
    if (waves <= 1) launch_col_pass_w<1>(rec, g0, M, nent, stride, cpass, s);
    else if (waves == 2) launch_col_pass_w<2>(rec, g0, M, nent, stride, cpass, s);
    else launch_col_pass_w<4>(rec, g0, M, nent, stride, cpass, s);
}

// How the t test of a column record moves with the ray origin's z, for every
// sub-column: 0 it passes on a prefix of rising z (or is constant), 1 on a
// suffix, 2 not known to be monotone.  Rounding is monotone, so each float
// operation of mt_t_hit is monotone in a monotone operand: tz rises with z,
// qx = ty e1z - tz e1y moves with -sgn(e1y), qy = tz e1x - tx e1z with
// sgn(e1x), qz is constant; e2x qx and e2y qy then move with -sgn(e2x)
// sgn(e1y) and sgn(e2y) sgn(e1x), their left-to-right sum (plus the constant
// e2z qz) is monotone when those do not oppose, and the product by inv flips
// it for inv < 0.  That needs every intermediate finite: with the edges,
// |tx|, |ty| and |v1z| within 2^42, the ray origins within 2^41 and |inv|
// within 1e7 no product or sum leaves the float range (2^110 at most); the
// record's part is checked here, the origins' where the search is used.
// The directions do not depend on (tx, ty).
__device__ __forceinline__ int sgnf(float x) { return (x > 0.0f) - (x < 0.0f); }

__device__ __forceinline__ int t_run_kind(const ColRec& r)
{
    const float lim = 0x1p42f;
    if (!(fabsf(r.e1x) <= lim && fabsf(r.e1y) <= lim && fabsf(r.e1z) <= lim && fabsf(r.e2x) <= lim &&
          fabsf(r.e2y) <= lim && fabsf(r.e2z) <= lim && fabsf(r.v1z) <= lim && fabsf(r.inv) <= 1e7f))
        return 2;
    const int da = -sgnf(r.e2x) * sgnf(r.e1y), db = sgnf(r.e2y) * sgnf(r.e1x);
    if (da * db < 0) return 2;
    return (da ? da : db) * sgnf(r.inv) > 0 ? 1 : 0;
}

// One boundary voxel (i1, j1, k1) of the Level2 parity.  N2 > 0: the Level2
// resolution as a compile-time constant (layer loops unrolled, the layer
// origins held uniform); N2 == 0: any resolution n2 <= 16.  The kernels use
// it for the resolutions other than 8 (two voxels per wave,
// l2_parity_pack2_8) and at most 4 (four per wave, l2_parity_pack4).
template <int N2, bool kRuns = false>
__device__ __forceinline__ void l2_parity_voxel(int n2_rt, int w2, const GridDev& g, int i1, int j1, int k1,
                                                uint32_t e0, uint32_t e1, const float4* __restrict__ rec,
                                                const uint64_t* __restrict__ cpass, int stride, ColRec* cr,
                                                unsigned long long* acc, float* soz, uint64_t* __restrict__ out)
{
    const int lane = __lane_id();
    const int n2 = N2 > 0 ? N2 : n2_rt;
    constexpr int NMX = N2 > 0 ? (N2 * N2 + 63) / 64 : 4;   // sub-column slots per lane
    const int nsc = n2 * n2;
    const int nm = N2 > 0 ? NMX : (nsc + 63) >> 6;
    const float* rx = g.r2x + (int64_t)i1 * n2;
    const float* ry = g.r2y + (int64_t)j1 * n2;
    const float* rz = g.r2z + (int64_t)k1 * n2;
    // this lane's sub-columns sc = lane + 64*mm: ray origins
    float ox[NMX], oy[NMX];
#pragma unroll
    for (int mm = 0; mm < NMX; ++mm) {
        const int sc = lane + 64 * mm;
        ox[mm] = (mm < nm && sc < nsc) ? rx[sc % n2] : 0.0f;
        oy[mm] = (mm < nm && sc < nsc) ? ry[sc / n2] : 0.0f;
    }
    // layer origins of the sub-voxels: entry r in lane r (and, for a fixed
    // resolution, in uniform registers)
    const float ozl = lane < n2 ? rz[lane] : 0.0f;
    float oz[N2 > 0 ? N2 : 1];
#pragma unroll
    for (int rr = 0; rr < (N2 > 0 ? N2 : 1); ++rr) oz[rr] = lane_f(ozl, rr);
    // resolutions without a fixed layer loop: the layer origins also in LDS,
    // for per-lane searches along the layers (tables non-decreasing)
    const bool zruns = kRuns && N2 == 0 && g.r2mono != 0 && n2 >= 8 && fabsf(lane_f(ozl, 0)) <= 0x1p41f &&
                       fabsf(lane_f(ozl, n2 - 1)) <= 0x1p41f;
    if (kRuns && N2 == 0 && lane < n2) soz[lane] = ozl;
    __builtin_amdgcn_wave_barrier();
    const uint64_t layers = (1ull << n2) - 1;
    uint64_t bits = 0;   // bit (mm*n2 + r) for sub-column lane + 64*mm
    for (uint32_t base = e0; base < e1; base += 64) {
        const uint32_t i = base + lane;
        const int k1c = k1;
        uint64_t row[NMX];   // candidate's det/u/v pass bits over the sub-columns
#pragma unroll
        for (int mm = 0; mm < NMX; ++mm) row[mm] = 0;
        bool hit = false;   // every layer's t passes: no t test, no record
        if (i < e1) {
            // header (k_col_pass): candidate while k1 <= zc
            const uint64_t hw = cpass[(uint64_t)i * stride + stride - 1];
            const int kh = (int)(uint32_t)hw - 1, zc = (int)(uint32_t)(hw >> 32) - 1;
            if (k1c <= zc) {
#pragma unroll
                for (int mm = 0; mm < NMX; ++mm)
                    if (mm < nm) row[mm] = cpass[(uint64_t)i * stride + mm];
                hit = k1c <= kh;
                if (!hit) {
                    ColRec r = load_colrec(rec, i);
                    if (kRuns) r.zmax = t_run_kind(r);   // the slot is free here: the header holds zc
                    cr[lane] = r;
                }
            }
        }
        const uint64_t all_hit = __ballot(hit);
        __builtin_amdgcn_wave_barrier();
        // pass bits per sub-column slot: bit j = candidate j passes at this
        // lane's sub-column
        uint64_t pass[NMX];
#pragma unroll
        for (int mm = 0; mm < NMX; ++mm) pass[mm] = mm < nm ? wave_transpose64(row[mm]) : 0;
        // t per sub-voxel for the passing pairs not decided for the voxel
#pragma unroll
        for (int mm = 0; mm < NMX; ++mm) {
            if (__popcll(pass[mm] & all_hit) & 1) bits ^= layers << (mm * n2);
            for (uint64_t m = pass[mm] & ~all_hit; m; m &= m - 1) {
                const int j = __ffsll((unsigned long long)m) - 1;
                const ColRec r = cr[j];
                const float tx = ox[mm] - r.v1x, ty = oy[mm] - r.v1y;
                uint32_t hm = 0;   // bit rr: t > 1e-6 at layer rr
                if constexpr (N2 > 0) {
#pragma unroll
                    for (int rr = 0; rr < N2; ++rr)
                        hm |= (uint32_t)mt_t_hit(r.e1x, r.e1y, r.e1z, r.e2x, r.e2y, r.e2z, r.inv, tx, ty, r.v1z,
                                                 oz[rr])
                              << rr;
                } else {
                    if (kRuns && zruns && r.zmax < 2 && fabsf(tx) <= 0x1p42f && fabsf(ty) <= 0x1p42f) {
                        // the passing layers are one run at an end: find its edge
                        const bool pre = r.zmax == 0;
                        int lo = 0, hi = n2;
                        while (lo < hi) {
                            const int mid = (lo + hi) >> 1;
                            const bool t = mt_t_hit(r.e1x, r.e1y, r.e1z, r.e2x, r.e2y, r.e2z, r.inv, tx, ty, r.v1z,
                                                    soz[mid]);
                            if (t == pre) lo = mid + 1;
                            else hi = mid;
                        }
                        hm = pre ? (uint32_t)((1ull << lo) - 1) : (uint32_t)(layers & ~((1ull << lo) - 1));
                    } else {
                        for (int rr = 0; rr < n2; ++rr)
                            hm |= (uint32_t)mt_t_hit(r.e1x, r.e1y, r.e1z, r.e2x, r.e2y, r.e2z, r.inv, tx, ty, r.v1z,
                                                     lane_f(ozl, rr))
                                  << rr;
                    }
                }
                bits ^= (uint64_t)hm << (mm * n2);
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (nsc % 64 == 0) {
        // word r * (nsc / 64) + mm holds layer r of sub-columns 64mm + lane:
        // one ballot per word, lane w keeps word w, one coalesced store
        uint64_t mine = 0;
#pragma unroll
        for (int mm = 0; mm < NMX; ++mm) {
            if (mm >= nm) break;
            for (int r = 0; r < n2; ++r) {
                const uint64_t wd = __ballot((bits >> (mm * n2 + r)) & 1ull);
                if (lane == r * nm + mm) mine = wd;
            }
        }
        if (lane < w2) out[lane] = mine;
        return;
    }
    for (int i = lane; i < w2; i += 64) acc[i] = 0ull;
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
    for (int i = lane; i < w2; i += 64) out[i] = acc[i];
}

// kM: batch (resolution per mesh, N2 = 0); single mesh: N2 = its resolution
// when it is 4 or 8 (a kernel of its own, so each keeps its own register
// budget), else 0
// Four boundary voxels of one mesh with n2 <= 4 (n2^2 <= 16 sub-columns) in
// one wave: lane 16v + sc is sub-column sc of voxel v.  Each quarter stages
// 16 of its own column's candidates per round; a candidate's pass bits sit at
// bits [16v, 16v + 16) of its lane's row, so one 64x64 transpose gives every
// lane the pass bits of its own voxel's candidates (the rows of the other
// quarters have no bits there).  N2 = 4: layers unrolled; 0: any n2 <= 4.
template <int N2>
__device__ __forceinline__ void l2_parity_pack4(int n2rt, const GridDev& g, uint32_t col0, int64_t bfirst, int nv,
                                                const int32_t* __restrict__ bidx, const uint32_t* __restrict__ coff,
                                                const float4* __restrict__ rec, const uint64_t* __restrict__ cpass,
                                                int stride, ColRec* cr, uint64_t* __restrict__ out)
{
    const int lane = __lane_id();
    const int n2 = N2 > 0 ? N2 : n2rt;
    const int nsc = n2 * n2;
    const int v = lane >> 4, sc = lane & 15;
    const bool lv = v < nv;
    int k1 = 0, ij = 0, i1 = 0, j1 = 0;
    if (lv) {
        const int vox = bidx[bfirst + v];
        const int nxy = g.nx * g.ny;
        k1 = vox / nxy;
        ij = vox - k1 * nxy;
        j1 = ij / g.nx;
        i1 = ij - j1 * g.nx;
    }
    const uint32_t e0 = lv ? coff[col0 + ij] : 0u, e1 = lv ? coff[col0 + ij + 1] : 0u;
    const bool ls = lv && sc < nsc;
    const float ox = ls ? g.r2x[(int64_t)i1 * n2 + sc % n2] : 0.0f;
    const float oy = ls ? g.r2y[(int64_t)j1 * n2 + sc / n2] : 0.0f;
    float oz[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) oz[r] = (lv && r < n2) ? g.r2z[(int64_t)k1 * n2 + r] : 0.0f;
    const uint32_t layers = (1u << n2) - 1;
    // rounds of 16 candidates: the most any quarter needs
    uint32_t nr = (e1 - e0 + 15) >> 4;
#pragma unroll
    for (int d = 16; d <= 32; d <<= 1) {
        const uint32_t o = (uint32_t)__shfl_xor((int)nr, d);
        nr = o > nr ? o : nr;
    }
    const uint32_t rounds = (uint32_t)__builtin_amdgcn_readfirstlane((int)nr);
    uint32_t par = 0;   // bit r: parity at layer r of this lane's sub-column
    for (uint32_t rd = 0; rd < rounds; ++rd) {
        const uint32_t i = e0 + rd * 16 + (uint32_t)sc;
        uint64_t row = 0;
        bool hit = false;   // every layer's t passes for this slot's voxel
        if (lv && i < e1) {
            // header (k_col_pass): candidate while k1 <= zc
            const uint64_t hw = cpass[(uint64_t)i * stride + stride - 1];
            const int kh = (int)(uint32_t)hw - 1, zc = (int)(uint32_t)(hw >> 32) - 1;
            if (k1 <= zc) {
                row = (cpass[(uint64_t)i * stride] & 0xffffull) << (16 * v);
                hit = k1 <= kh;
                if (!hit) cr[lane] = load_colrec(rec, i);
            }
        }
        // slot j's flag is for slot j's own voxel, the only one whose lanes
        // see bit j after the transpose
        const uint64_t all_hit = __ballot(hit);
        __builtin_amdgcn_wave_barrier();
        const uint64_t pm = wave_transpose64(row);
        if (__popcll(pm & all_hit) & 1) par ^= layers;
        for (uint64_t m = pm & ~all_hit; m; m &= m - 1) {
            const int j = __ffsll((unsigned long long)m) - 1;
            const ColRec r = cr[j];
            const float tx = ox - r.v1x, ty = oy - r.v1y;
            uint32_t hm = 0;
            if constexpr (N2 > 0) {
#pragma unroll
                for (int rr = 0; rr < N2; ++rr)
                    hm |= (uint32_t)mt_t_hit(r.e1x, r.e1y, r.e1z, r.e2x, r.e2y, r.e2z, r.inv, tx, ty, r.v1z, oz[rr])
                          << rr;
            } else {
                for (int rr = 0; rr < n2; ++rr)
                    hm |= (uint32_t)mt_t_hit(r.e1x, r.e1y, r.e1z, r.e2x, r.e2y, r.e2z, r.inv, tx, ty, r.v1z,
                                             rr == 0 ? oz[0] : rr == 1 ? oz[1] : rr == 2 ? oz[2] : oz[3])
                          << rr;
            }
            par ^= hm;
        }
        __builtin_amdgcn_wave_barrier();
    }
    // voxel v's word: bit r*nsc + sc for layer r (n2^3 <= 64, one word)
    const uint64_t smask = nsc >= 64 ? ~0ull : ((1ull << nsc) - 1);
    uint64_t word = 0;
    for (int r = 0; r < n2; ++r) {
        const uint64_t m = __ballot((par >> r) & 1u);
        word |= ((m >> (16 * v)) & smask) << (r * nsc);
    }
    if (lv && sc == 0) out[v] = word;
}

// Waves: single mesh, one per boundary voxel, or (n2 <= 4) one per four;
// batch, per mesh from MeshOff::pw0 likewise.
// Two boundary voxels of a Level2-8 mesh per wave.  Lanes 32h .. 32h + 31
// load voxel h's column entries, 32 per round; one 64x64 transpose of their
// pass-bit rows leaves lane l with sub-column l of both voxels (bit j < 32:
// voxel 0's slot j passes there; bit j >= 32: voxel 1's slot j - 32), so the
// two voxels' dependent loads and tests overlap in one wave and share its
// transposes.  Voxel h's words go to out[8h .. 8h + 7] (word r holds layer r
// of its 64 sub-columns, bit = sub-column).
__device__ __forceinline__ void l2_parity_pack2_8(const GridDev& g, uint32_t col0, int64_t bfirst, int nv,
                                                  const int32_t* __restrict__ bidx, const uint32_t* __restrict__ coff,
                                                  const float4* __restrict__ rec, const uint64_t* __restrict__ cpass,
                                                  int stride, ColRec* cr, uint64_t* __restrict__ out)
{
    const int lane = __lane_id();
    const int h = lane >> 5, slot = lane & 31;
    const int nxy = g.nx * g.ny;
    int k1v[2];
    uint32_t e0v[2], e1v[2];
    float ox[2], oy[2], oz[2][8];
#pragma unroll
    for (int hh = 0; hh < 2; ++hh) {
        const int vox = uniform_i(bidx[bfirst + (hh < nv ? hh : 0)]);
        const int k1 = vox / nxy;
        const int ij = vox - k1 * nxy;
        const int j1 = ij / g.nx;
        const int i1 = ij - j1 * g.nx;
        k1v[hh] = k1;
        e0v[hh] = (uint32_t)uniform_i((int)coff[col0 + ij]);
        e1v[hh] = hh < nv ? (uint32_t)uniform_i((int)coff[col0 + ij + 1]) : e0v[hh];
        ox[hh] = g.r2x[(int64_t)i1 * 8 + (lane & 7)];
        oy[hh] = g.r2y[(int64_t)j1 * 8 + (lane >> 3)];
#pragma unroll
        for (int rr = 0; rr < 8; ++rr)
            oz[hh][rr] = __int_as_float(uniform_i(__float_as_int(g.r2z[(int64_t)k1 * 8 + rr])));
    }
    const int k1 = h ? k1v[1] : k1v[0];
    const uint32_t e0 = h ? e0v[1] : e0v[0], e1 = h ? e1v[1] : e1v[0];
    const uint32_t n0 = (e1v[0] - e0v[0] + 31) >> 5, n1 = (e1v[1] - e0v[1] + 31) >> 5;
    const uint32_t rounds = n0 > n1 ? n0 : n1;
    uint32_t par[2] = {0u, 0u};   // bit r: parity at layer r of sub-column lane of voxel h
    for (uint32_t rd = 0; rd < rounds; ++rd) {
        const uint32_t i = e0 + rd * 32 + (uint32_t)slot;
        uint64_t row = 0;
        bool hit = false;   // every layer's t passes: no t test, no record
        if (i < e1) {
            // header (k_col_pass): candidate while k1 <= zc
            const uint64_t hw = cpass[(uint64_t)i * stride + stride - 1];
            const int kh = (int)(uint32_t)hw - 1, zc = (int)(uint32_t)(hw >> 32) - 1;
            if (k1 <= zc) {
                row = cpass[(uint64_t)i * stride];
                hit = k1 <= kh;
                if (!hit) cr[lane] = load_colrec(rec, i);
            }
        }
        const uint64_t all_hit = __ballot(hit);
        __builtin_amdgcn_wave_barrier();
        const uint64_t pm = wave_transpose64(row);
#pragma unroll
        for (int hh = 0; hh < 2; ++hh) {
            const uint64_t mh = pm & (hh ? 0xffffffff00000000ull : 0x00000000ffffffffull);
            if (__popcll(mh & all_hit) & 1) par[hh] ^= 0xffu;
            for (uint64_t m = mh & ~all_hit; m; m &= m - 1) {
                const int j = __ffsll((unsigned long long)m) - 1;
                const ColRec r = cr[j];
                const float tx = ox[hh] - r.v1x, ty = oy[hh] - r.v1y;
                uint32_t hm = 0;
#pragma unroll
                for (int rr = 0; rr < 8; ++rr)
                    hm |= (uint32_t)mt_t_hit(r.e1x, r.e1y, r.e1z, r.e2x, r.e2y, r.e2z, r.inv, tx, ty, r.v1z, oz[hh][rr])
                          << rr;
                par[hh] ^= hm;
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    uint64_t mine = 0;
#pragma unroll
    for (int hh = 0; hh < 2; ++hh) {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const uint64_t wd = __ballot((par[hh] >> r) & 1u);
            if (lane == hh * 8 + r) mine = wd;
        }
    }
    if (lane < 8 * nv) out[lane] = mine;
}

// kRuns: some mesh has a resolution above 8 without a fixed layer loop,
// whose passing layers are found by a search along the run (the other
// instantiations leave that path out and keep its registers)
// Eight waves per SIMD (at most 64 VGPRs) for the latency of each voxel's
// chain of dependent loads.
template <bool kM, int N2, int L2P_WAVES, bool kRuns>
__global__ __launch_bounds__(64 * L2P_WAVES) __attribute__((amdgpu_waves_per_eu(8))) void k_l2_parity(const int32_t* __restrict__ bidx, int64_t B,
                                                              int64_t nwaves, const uint32_t* __restrict__ coff,
                                                              const float4* __restrict__ rec,
                                                              const uint64_t* __restrict__ cpass, int stride,
                                                              GridDev g0, Meshes M, uint64_t* __restrict__ par2)
{
    __shared__ unsigned long long s_acc[L2P_WAVES][64];
    __shared__ ColRec s_rec[L2P_WAVES][64];
    __shared__ float s_oz[L2P_WAVES][16];
    const int wv = threadIdx.x >> 6;
    const int64_t w = (int64_t)blockIdx.x * L2P_WAVES + wv;
    if (w >= nwaves) return;
    int mesh = 0;
    if constexpr (kM)   // the mesh whose waves hold w
        mesh = wave_mesh_search(M.off, M.nm, (uint64_t)w, [](const MeshOff& o) { return (uint64_t)o.pw0; });
    const GridDev& g = grid_of<kM>(g0, M, mesh);
    const int n2 = g.n2, w2 = g.w2, gnx = g.nx, gny = g.ny;
    if (kM && n2 == 0) return;
    const int64_t mb0 = kM ? (int64_t)M.off[mesh].b0 : 0;
    const int64_t mb1 = kM ? (int64_t)M.off[mesh + 1].b0 : B;
    const int64_t lw = w - (kM ? (int64_t)M.off[mesh].pw0 : 0);
    const uint64_t l2w0 = kM ? M.off[mesh].l2w0 : 0;
    const uint32_t col0 = kM ? M.off[mesh].col0 : 0;
    if ((N2 == 0 || N2 == 8) && n2 == 8) {
        const int64_t bfirst = mb0 + 2 * lw;
        const int nv = (int)(mb1 - bfirst < 2 ? mb1 - bfirst : 2);
        l2_parity_pack2_8(g, col0, bfirst, nv, bidx, coff, rec, cpass, stride, s_rec[wv],
                          par2 + l2w0 + (uint64_t)(bfirst - mb0) * 8);
        return;
    }
    if ((N2 == 0 || N2 == 4) && n2 <= 4) {
        const int64_t bfirst = mb0 + 4 * lw;
        const int nv = (int)(mb1 - bfirst < 4 ? mb1 - bfirst : 4);
        uint64_t* out = par2 + l2w0 + (uint64_t)(bfirst - mb0);   // w2 == 1
        if (n2 == 4)
            l2_parity_pack4<4>(n2, g, col0, bfirst, nv, bidx, coff, rec, cpass, stride, s_rec[wv], out);
        else
            l2_parity_pack4<0>(n2, g, col0, bfirst, nv, bidx, coff, rec, cpass, stride, s_rec[wv], out);
        return;
    }
    const int64_t b = mb0 + lw;
    uint64_t* out = par2 + l2w0 + (uint64_t)(b - mb0) * w2;
    const int v = bidx[b];
    const int nxy = gnx * gny;
    const int k1 = v / nxy;
    const int ij = v - k1 * nxy;
    const int j1 = ij / gnx;
    const int i1 = ij - j1 * gnx;
    const uint32_t e0 = coff[col0 + ij], e1 = coff[col0 + ij + 1];
    // the other resolutions (Level2 8 and at most 4 returned above)
    if constexpr (N2 == 0)
        l2_parity_voxel<0, kRuns>(n2, w2, g, i1, j1, k1, e0, e1, rec, cpass, stride, s_rec[wv], s_acc[wv], s_oz[wv],
                                  out);
}

// waves of the Level2 parity for nb boundary voxels at resolution n2
int64_t l2_parity_waves(int n2, int64_t nb)
{
    if (n2 <= 0) return 0;
    return n2 <= 4 ? (nb + 3) / 4 : n2 == 8 ? (nb + 1) / 2 : nb;
}

template <int PW, bool kRuns>
static void launch_l2_parity_r(const int32_t* bidx, int64_t B, int64_t nwaves, const uint32_t* coff,
                               const float4* rec, const uint64_t* cpass, int stride, const GridDev& g0,
                               const Meshes& M, uint64_t* par2, hipStream_t s)
{
    const unsigned blocks = (unsigned)((nwaves + PW - 1) / PW);
    const dim3 bs(64 * PW);
    if (M.nm > 1)
        k_l2_parity<true, 0, PW, kRuns><<<blocks, bs, 0, s>>>(bidx, B, nwaves, coff, rec, cpass, stride, g0, M, par2);
    else if (g0.n2 == 8)
        k_l2_parity<false, 8, PW, false><<<blocks, bs, 0, s>>>(bidx, B, nwaves, coff, rec, cpass, stride, g0, M,
                                                                par2);
    else if (g0.n2 == 4)
        k_l2_parity<false, 4, PW, false><<<blocks, bs, 0, s>>>(bidx, B, nwaves, coff, rec, cpass, stride, g0, M,
                                                                par2);
    else
        k_l2_parity<false, 0, PW, kRuns><<<blocks, bs, 0, s>>>(bidx, B, nwaves, coff, rec, cpass, stride, g0, M,
                                                                par2);
}

template <int PW>
static void launch_l2_parity_w(const int32_t* bidx, int64_t B, int64_t nwaves, const uint32_t* coff,
                               const float4* rec, const uint64_t* cpass, int stride, int n2max, const GridDev& g0,
                               const Meshes& M, uint64_t* par2, hipStream_t s)
{
    if (n2max > 8)
        launch_l2_parity_r<PW, true>(bidx, B, nwaves, coff, rec, cpass, stride, g0, M, par2, s);
    else
        launch_l2_parity_r<PW, false>(bidx, B, nwaves, coff, rec, cpass, stride, g0, M, par2, s);
}

void launch_l2_parity(const int32_t* bidx, int64_t B, int64_t nwaves, const uint32_t* coff, const float4* rec,
                      const uint64_t* cpass, int stride, int n2max, int waves, const GridDev& g0, const Meshes& M,
                      uint64_t* par2, hipStream_t s)
{
    if (B <= 0 || nwaves <= 0) return;
    if (waves <= 1) launch_l2_parity_w<1>(bidx, B, nwaves, coff, rec, cpass, stride, n2max, g0, M, par2, s);
    else if (waves == 2) launch_l2_parity_w<2>(bidx, B, nwaves, coff, rec, cpass, stride, n2max, g0, M, par2, s);
    else launch_l2_parity_w<4>(bidx, B, nwaves, coff, rec, cpass, stride, n2max, g0, M, par2, s);
}

// ---------------------------------------------------------------------------
// normals.  Face normal: side1 = v1 - v0, side2 = v2 - v0, their cross
// product, VectorNormalize (src/Object.cpp:3233-3240; the load-time
// faceNormal of :521-526 is the same float expression on the same points).
// VectorNormalize (includes/FloatVector.h:332-341) divides by the float
// magnitude of the left-to-right sum of squares when it is nonzero; sqrtf and
// the division are correctly rounded here.  The reference sums in bin order,
// which its CPU sweep makes ascending in triangle id, so bins are sorted
// before any sum.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void normalize3(float& x, float& y, float& z)
{
    const float m = sqrtf(x * x + y * y + z * z);
    if (m != 0.0f) {
        x = x / m;
        y = y / m;
        z = z / m;
    }
}

__global__ void k_face_normals(const float* __restrict__ tris, int64_t T, float* __restrict__ fn)
{
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < T; t += (int64_t)gridDim.x * blockDim.x) {
        const float* tv = tris + 9 * t;
        const float a0 = tv[3] - tv[0], a1 = tv[4] - tv[1], a2 = tv[5] - tv[2];
        const float b0 = tv[6] - tv[0], b1 = tv[7] - tv[1], b2 = tv[8] - tv[2];
        float x = a1 * b2 - b1 * a2;
        float y = a2 * b0 - b2 * a0;
        float z = a0 * b1 - b0 * a1;
        normalize3(x, y, z);
        fn[3 * t + 0] = x;
        fn[3 * t + 1] = y;
        fn[3 * t + 2] = z;
    }
}

void launch_face_normals(const float* tris, int64_t T, float* fn, hipStream_t s)
{
    if (T <= 0) return;
    k_face_normals<<<grid_for(T, 256, 16384), 256, 0, s>>>(tris, T, fn);
}

// true when boundary voxel b's mesh asks for normals (wave-uniform)
template <bool kM>
__device__ __forceinline__ bool wants_normals(const Meshes& M, int64_t b)
{
    if constexpr (kM) return M.grids[uniform_i((int)M.bmesh[b])].normals != 0;
    else return true;
}

// One wave per boundary voxel: the bin's triangle ids (distinct) ranked by
// counting smaller ids, written in ascending order to slist.
template <bool kM>
__global__ __launch_bounds__(256) void k_sort_bins(const uint32_t* __restrict__ list,
                                                   const uint32_t* __restrict__ boff, int64_t B, Meshes M,
                                                   uint32_t* __restrict__ slist)
{
    const int64_t b = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (b >= B || !wants_normals<kM>(M, b)) return;
    const int lane = __lane_id();
    const uint32_t e0 = boff[b], n = boff[b + 1] - e0;
    for (uint32_t bi = 0; bi < n; bi += 64) {
        const uint32_t ti = bi + lane < n ? list[e0 + bi + lane] : 0xffffffffu;
        uint32_t rank = 0;
        for (uint32_t bj = 0; bj < n; bj += 64) {
            const uint32_t tj = bj + lane < n ? list[e0 + bj + lane] : 0xffffffffu;
            const uint32_t m = n - bj < 64 ? n - bj : 64;
            for (uint32_t k = 0; k < m; ++k)
                rank += (uint32_t)((uint32_t)__builtin_amdgcn_readlane((int)tj, (int)k) < ti);
        }
        if (bi + lane < n) slist[e0 + rank] = ti;
    }
}

void launch_sort_bins(const uint32_t* list, const uint32_t* boff, int64_t B, const GridDev& g0, const Meshes& M,
                      uint32_t* slist, hipStream_t s)
{
    (void)g0;
    if (B <= 0) return;
    const unsigned blocks = (unsigned)((B * 64 + 255) / 256);
    if (M.nm > 1) k_sort_bins<true><<<blocks, 256, 0, s>>>(list, boff, B, M, slist);
    else k_sort_bins<false><<<blocks, 256, 0, s>>>(list, boff, B, M, slist);
}

// Level1 normal of a boundary voxel (src/Object.cpp:3219-3253): the face
// normals of its sorted bin summed in order, divided by the count, normalized.
template <bool kM>
__global__ __launch_bounds__(256) void k_l1_normals(const uint32_t* __restrict__ slist,
                                                    const uint32_t* __restrict__ boff, int64_t B,
                                                    const float* __restrict__ fn, Meshes M,
                                                    float* __restrict__ l1n)
{
    const int64_t b = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (b >= B || !wants_normals<kM>(M, b)) return;
    const int lane = __lane_id();
    const uint32_t e0 = boff[b], n = boff[b + 1] - e0;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    for (uint32_t bi = 0; bi < n; bi += 64) {
        float fx = 0.0f, fy = 0.0f, fz = 0.0f;
        if (bi + lane < n) {
            const uint64_t t = slist[e0 + bi + lane];
            fx = fn[3 * t];
            fy = fn[3 * t + 1];
            fz = fn[3 * t + 2];
        }
        const uint32_t m = n - bi < 64 ? n - bi : 64;
        for (uint32_t k = 0; k < m; ++k) {
            sx = sx + lane_f(fx, (int)k);
            sy = sy + lane_f(fy, (int)k);
            sz = sz + lane_f(fz, (int)k);
        }
    }
    const float cnt = (float)n;
    float x = sx / cnt, y = sy / cnt, z = sz / cnt;
    normalize3(x, y, z);
    if (lane == 0) {
        l1n[3 * b + 0] = x;
        l1n[3 * b + 1] = y;
        l1n[3 * b + 2] = z;
    }
}

void launch_l1_normals(const uint32_t* slist, const uint32_t* boff, int64_t B, const float* fn, const GridDev& g0,
                       const Meshes& M, float* l1n, hipStream_t s)
{
    (void)g0;
    if (B <= 0) return;
    const unsigned blocks = (unsigned)((B * 64 + 255) / 256);
    if (M.nm > 1) k_l1_normals<true><<<blocks, 256, 0, s>>>(slist, boff, B, fn, M, l1n);
    else k_l1_normals<false><<<blocks, 256, 0, s>>>(slist, boff, B, fn, M, l1n);
}

// Level2 normals (AddPointToVoxel src/Object.cpp:2705-2720 under the loop nest
// of :2372-2400, averaged at :2404-2421): per Level2 voxel, the face normals
// of the bin triangles whose SAT hits it, in bin order, over their count,
// normalized; (0, 0, 0) where none hits.  One wave per boundary voxel, each
// lane owning L2N_G Level2 voxels per pass; the box-axis predicates are
// evaluated once per (triangle, sub-index) by ballot, as in k_l2_sat.
constexpr int L2N_G = 8;

template <bool kM>
__global__ __launch_bounds__(256) void k_l2_normals(const float* __restrict__ tris,
                                                    const int32_t* __restrict__ bidx, int64_t B,
                                                    const uint32_t* __restrict__ boff,
                                                    const uint32_t* __restrict__ slist,
                                                    const float* __restrict__ fn, GridDev g0, Meshes M,
                                                    float* __restrict__ l2n)
{
    const int64_t b = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (b >= B) return;
    const int mesh = kM ? uniform_i((int)M.bmesh[b]) : 0;
    const GridDev& g = grid_of<kM>(g0, M, mesh);
    const int n2 = g.n2, gnx = g.nx, gny = g.ny;
    if (n2 == 0 || (kM && !g.normals)) return;
    const float h2x = g.h2x, h2y = g.h2y, h2z = g.h2z;
    const int S = n2 * n2 * n2;
    float* out = l2n + 3 * (kM ? M.off[mesh].l2n0 + (uint64_t)(b - M.off[mesh].b0) * S : (uint64_t)b * S);
    const int v = bidx[b];
    const int nxy = gnx * gny;
    const int k1 = v / nxy;
    const int ij = v - k1 * nxy;
    const int j1 = ij / gnx;
    const int i1 = ij - j1 * gnx;
    const float* c2x = g.c2x + (int64_t)i1 * n2;
    const float* c2y = g.c2y + (int64_t)j1 * n2;
    const float* c2z = g.c2z + (int64_t)k1 * n2;
    const int lane = __lane_id();
    const float cxl = lane < n2 ? c2x[lane] : 0.0f;
    const float cyl = lane < n2 ? c2y[lane] : 0.0f;
    const float czl = lane < n2 ? c2z[lane] : 0.0f;
    const uint32_t e0 = boff[b], n = boff[b + 1] - e0;
    for (int s0 = 0; s0 < S; s0 += 64 * L2N_G) {
        float ax[L2N_G], ay[L2N_G], az[L2N_G], ac[L2N_G];
        float ccx[L2N_G], ccy[L2N_G], ccz[L2N_G];
        int sp[L2N_G], sq[L2N_G], sr[L2N_G];
        // centers of this lane's Level2 voxels, gathered while every lane is active
#pragma unroll
        for (int k = 0; k < L2N_G; ++k) {
            const int sv = s0 + k * 64 + lane;
            ax[k] = ay[k] = az[k] = ac[k] = 0.0f;
            sp[k] = sv % n2;
            sq[k] = (sv / n2) % n2;
            sr[k] = (sv / (n2 * n2)) % n2;
            ccx[k] = __shfl(cxl, sp[k]);
            ccy[k] = __shfl(cyl, sq[k]);
            ccz[k] = __shfl(czl, sr[k]);
        }
        for (uint32_t e = 0; e < n; ++e) {
            const uint64_t t = slist[e0 + e];
            float tv[9];
#pragma unroll
            for (int c = 0; c < 9; ++c) tv[c] = tris[t * 9 + c];
            const uint64_t mx = __ballot(lane < n2 && sat_axis_ok(tv[0], tv[3], tv[6], cxl, h2x));
            const uint64_t my = __ballot(lane < n2 && sat_axis_ok(tv[1], tv[4], tv[7], cyl, h2y));
            const uint64_t mz = __ballot(lane < n2 && sat_axis_ok(tv[2], tv[5], tv[8], czl, h2z));
            if (!mx || !my || !mz) continue;
            const float fx = fn[3 * t], fy = fn[3 * t + 1], fz = fn[3 * t + 2];
#pragma unroll
            for (int k = 0; k < L2N_G; ++k) {
                const int sv = s0 + k * 64 + lane;
                if (sv >= S || !((mx >> sp[k]) & (my >> sq[k]) & (mz >> sr[k]) & 1ull)) continue;
                if (tri_box_overlap<false>(ccx[k], ccy[k], ccz[k], h2x, h2y, h2z, tv)) {
                    ax[k] = ax[k] + fx;
                    ay[k] = ay[k] + fy;
                    az[k] = az[k] + fz;
                    ac[k] = ac[k] + 1.0f;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < L2N_G; ++k) {
            const int sv = s0 + k * 64 + lane;
            if (sv >= S) continue;
            float x = 0.0f, y = 0.0f, z = 0.0f;
            if (ac[k] > 0.0f) {
                x = ax[k] / ac[k];
                y = ay[k] / ac[k];
                z = az[k] / ac[k];
                normalize3(x, y, z);
            }
            out[3 * (int64_t)sv + 0] = x;
            out[3 * (int64_t)sv + 1] = y;
            out[3 * (int64_t)sv + 2] = z;
        }
    }
}

void launch_l2_normals(const float* tris, const int32_t* bidx, int64_t B, const uint32_t* boff,
                       const uint32_t* slist, const float* fn, const GridDev& g0, const Meshes& M, float* l2n,
                       hipStream_t s)
{
    if (B <= 0) return;
    const unsigned blocks = (unsigned)((B * 64 + 255) / 256);
    if (M.nm > 1) k_l2_normals<true><<<blocks, 256, 0, s>>>(tris, bidx, B, boff, slist, fn, g0, M, l2n);
    else k_l2_normals<false><<<blocks, 256, 0, s>>>(tris, bidx, B, boff, slist, fn, g0, M, l2n);
}

// .raw normal byte: outputDType(n * normScale + normBias), normScale =
// 256.0/3.0 and normBias = 127.0 as floats (src/Object.cpp:2940-2942, 3032-3034)
__device__ __forceinline__ uint8_t normal_byte(float v)
{
    const float y = v * 85.33333587646484375f;
    return (uint8_t)(int)(y + 127.0f);
}

// Level1Normal.raw: 3 bytes per Level1 voxel, [z][y][x]; only boundary voxels
// carry a normal, the rest are 0 -> 127
__global__ void k_export_l1_normals(const uint64_t* __restrict__ bnd, const uint32_t* __restrict__ wprefix,
                                    uint32_t base, const float* __restrict__ l1n, GridDev g, int64_t n1,
                                    uint8_t* __restrict__ out)
{
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n1; v += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(v % g.nx);
        const int64_t row = v / g.nx;
        const int64_t w = row * g.nxw + (x >> 6);
        float nx = 0.0f, ny = 0.0f, nz = 0.0f;
        if ((bnd[w] >> (x & 63)) & 1ull) {
            const uint64_t r = rank_of(bnd, wprefix, (uint64_t)w * 64 + (uint64_t)(x & 63)) - base;
            nx = l1n[3 * r];
            ny = l1n[3 * r + 1];
            nz = l1n[3 * r + 2];
        }
        out[3 * v + 0] = normal_byte(nx);
        out[3 * v + 1] = normal_byte(ny);
        out[3 * v + 2] = normal_byte(nz);
    }
}

void launch_export_l1_normals(const uint64_t* bnd, const uint32_t* wprefix, uint32_t base, const float* l1n,
                              const GridDev& g, uint8_t* out, hipStream_t s)
{
    const int64_t n1 = (int64_t)g.nx * g.ny * g.nz;
    if (n1 <= 0) return;
    k_export_l1_normals<<<grid_for(n1, 256, 65536), 256, 0, s>>>(bnd, wprefix, base, l1n, g, n1, out);
}

__global__ void k_export_l2_normals(const float* __restrict__ l2n, int64_t n, uint8_t* __restrict__ out)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = normal_byte(l2n[i]);
}

void launch_export_l2_normals(const float* l2n, int64_t n, uint8_t* out, hipStream_t s)
{
    if (n <= 0) return;
    k_export_l2_normals<<<grid_for(n, 256, 65536), 256, 0, s>>>(l2n, n, out);
}

// ---------------------------------------------------------------------------
// counts (src/Object.cpp:3353-3378)
// ---------------------------------------------------------------------------
// Single-mesh form: a grid-stride pass over both planes.  Multi-mesh form:
// blockIdx.y is the mesh, blockIdx.x strides over that mesh's words.
template <bool kM>
__global__ void k_count_states(const uint64_t* __restrict__ bnd, const uint64_t* __restrict__ ins, int64_t W,
                               const uint64_t* __restrict__ sat2, const uint64_t* __restrict__ par2,
                               int64_t W2, Meshes M)
{
    const int mesh = kM ? (int)blockIdx.y : 0;
    const uint64_t w_lo = kM ? M.off[mesh].word0 : 0, w_hi = kM ? M.off[mesh + 1].word0 : (uint64_t)W;
    const uint64_t l_lo = kM ? M.off[mesh].l2w0 : 0, l_hi = kM ? M.off[mesh + 1].l2w0 : (uint64_t)W2;
    uint64_t a = 0, b = 0, c = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = w_lo + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < w_hi; i += stride)
        a += (uint64_t)__popcll(ins[i] & ~bnd[i]);
    // the Level2 planes four words per lane at a time, their loads issued
    // together (the blocks are few: they add into one counter each)
    constexpr int U = 4;
    uint64_t i = l_lo + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + (U - 1) * stride < l_hi; i += U * stride) {
        uint64_t p[U], q[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            p[u] = par2[i + u * stride];
            q[u] = sat2[i + u * stride];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            b += (uint64_t)__popcll(p[u] & ~q[u]);
            c += (uint64_t)__popcll(q[u]);
        }
    }
    for (; i < l_hi; i += stride) {
        b += (uint64_t)__popcll(par2[i] & ~sat2[i]);
        c += (uint64_t)__popcll(sat2[i]);
    }
    block_add_u64(a, &M.ctr[mesh].l1_inside);
    block_add_u64(b, &M.ctr[mesh].l2_inside);
    block_add_u64(c, &M.ctr[mesh].l2_boundary);
}

void launch_count_states(const uint64_t* bnd, const uint64_t* ins, int64_t W, const uint64_t* sat2,
                         const uint64_t* par2, int64_t W2, const Meshes& M, hipStream_t s)
{
    const int64_t n = W > W2 ? W : W2;
    if (n <= 0) return;
    if (M.nm > 1) {
        const int64_t per = (n + M.nm - 1) / M.nm;
        const dim3 grid((unsigned)grid_for(per, 256, 64), (unsigned)M.nm);
        k_count_states<true><<<grid, 256, 0, s>>>(bnd, ins, W, sat2, par2, W2, M);
    } else {
        k_count_states<false><<<grid_for(n, 256, 512), 256, 0, s>>>(bnd, ins, W, sat2, par2, W2, M);
    }
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

__global__ void k_export_prefix(const uint64_t* __restrict__ bnd, const uint32_t* __restrict__ wprefix,
                                uint32_t base, GridDev g, int64_t n1, int32_t* __restrict__ out)
{
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n1; v += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(v % g.nx);
        const int64_t row = v / g.nx;
        const uint64_t bitidx = (uint64_t)(row * g.nxw + (x >> 6)) * 64 + (uint64_t)(x & 63);
        out[v] = (int32_t)(rank_of(bnd, wprefix, bitidx) - base);
    }
}

void launch_export_prefix(const uint64_t* bnd, const uint32_t* wprefix, uint32_t base, const GridDev& g,
                          int32_t* out, hipStream_t s)
{
    const int64_t n1 = (int64_t)g.nx * g.ny * g.nz;
    if (n1 <= 0) return;
    k_export_prefix<<<grid_for(n1, 256, 65536), 256, 0, s>>>(bnd, wprefix, base, g, n1, out);
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
