// gpv_runtime.cpp -- host side of libgpv: contexts, device pool, grid
// derivation, the staged pipeline and the C ABI (include/gpv.h).
//
// Host float semantics matter here too: bbox, grid sizes and center tables
// are computed with the reference's float/double promotions (see the
// citations per function) and this file is compiled with -ffp-contract=off.
#include "gpv.h"
#include "gpv_internal.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

using namespace gpv;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(e_ == hipErrorOutOfMemory ? GPV_ERR_OOM : GPV_ERR_HIP, "%s: %s", \
                        #expr, hipGetErrorString(e_));                                   \
    } while (0)

// Device memory pool: size classes (powers of two up to 256 MiB, 16 MiB
// granules above), blocks cached until the context is destroyed, so the
// steady state of repeated voxelizations allocates nothing.
class Pool {
public:
    ~Pool() { trim(); }
    static size_t size_class(size_t n)
    {
        if (n < 256) n = 256;
        if (n <= (size_t(256) << 20)) {
            size_t c = 256;
            while (c < n) c <<= 1;
            return c;
        }
        const size_t gran = size_t(16) << 20;
        return (n + gran - 1) / gran * gran;
    }
    hipError_t get(size_t n, void** out)
    {
        const size_t c = size_class(n);
        auto it = free_.find(c);
        if (it != free_.end()) {
            *out = it->second;
            free_.erase(it);
            live_[*out] = c;
            return hipSuccess;
        }
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, c);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            trim();
            e = hipMalloc(&p, c);
            if (e != hipSuccess) return e;
        }
        live_[p] = c;
        bytes_ += c;
        *out = p;
        return hipSuccess;
    }
    void put(void* p)
    {
        if (!p) return;
        auto it = live_.find(p);
        if (it == live_.end()) return;
        free_.emplace(it->second, p);
        live_.erase(it);
    }
    void trim()
    {
        for (auto& kv : free_) {
            (void)hipFree(kv.second);
            bytes_ -= kv.first;
        }
        free_.clear();
    }
    size_t bytes() const { return bytes_; }

private:
    std::multimap<size_t, void*> free_;
    std::unordered_map<void*, size_t> live_;
    size_t bytes_ = 0;
};

}  // namespace

struct gpv_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    Pool pool;
    Counters* h_ctr = nullptr;   // pinned read-back slots
    uint32_t* h_nb = nullptr;
    int num_cu = 256;
};

struct gpv_grid {
    gpv_ctx* ctx = nullptr;
    gpv_grid_info info{};
    GridDev g{};
    gpv_params params{};
    int64_t T = 0;
    const float* tris = nullptr;
    bool own_tris = false;
    float* tables = nullptr;
    // per triangle
    int4 *box0 = nullptr, *box1 = nullptr, *rng = nullptr;
    uint32_t *sat_items = nullptr, *ray_items = nullptr;
    uint64_t *sat_off = nullptr, *ray_off = nullptr, *scan64 = nullptr;
    uint32_t* scan32 = nullptr;
    size_t scan_cap = 0;
    Counters* ctr = nullptr;
    // Level1 bitplanes
    int64_t W = 0;
    uint64_t *bnd = nullptr, *ins = nullptr;
    // append buffers
    uint64_t* pair_key = nullptr;
    uint64_t pair_cap = 0;
    uint64_t* col_key = nullptr;
    uint32_t* col_z = nullptr;
    uint64_t col_cap = 0;
    uint64_t* hit_key = nullptr;
    uint64_t hit_cap = 0;
    uint64_t n_pairs = 0, n_cols = 0, n_hits = 0;
    // compaction and bins
    uint32_t *wcnt = nullptr, *wprefix = nullptr;
    int32_t* bidx = nullptr;
    int64_t B = 0;
    uint32_t *pb = nullptr, *cntb = nullptr, *boff = nullptr, *list = nullptr;
    uint32_t *ccnt = nullptr, *coff = nullptr;
    float4* crec = nullptr;   // 3 float4 per column entry
    uint32_t *hcnt = nullptr, *hoff = nullptr;
    float4* hrec = nullptr;
    uint64_t *sat2 = nullptr, *par2 = nullptr;
    // progress
    bool setup_done = false, sat_done = false, inout_done = false, compact_done = false,
         l2p_done = false, l2s_done = false, finished = false;
    // stream events: 0 setup, 1 L1 SAT, 2 ray expansion, 3 parity fill,
    // 4 compaction+bins, 5 L2 parity, 6 L2 SAT start, 7 L2 SAT end, 8 finish
    hipEvent_t ev[9] = {};
    std::vector<void*> bufs;
};

namespace {

template <typename T>
int alloc(gpv_grid* G, T** p, size_t count)
{
    void* q = nullptr;
    hipError_t e = G->ctx->pool.get(std::max<size_t>(count, 1) * sizeof(T), &q);
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? GPV_ERR_OOM : GPV_ERR_HIP, "device allocation of %zu bytes: %s",
                    count * sizeof(T), hipGetErrorString(e));
    *p = static_cast<T*>(q);
    G->bufs.push_back(q);
    return GPV_OK;
}

template <typename T>
void release(gpv_grid* G, T*& p)
{
    if (!p) return;
    void* q = (void*)p;
    G->ctx->pool.put(q);
    auto it = std::find(G->bufs.begin(), G->bufs.end(), q);
    if (it != G->bufs.end()) G->bufs.erase(it);
    p = nullptr;
}

#define TRY(expr)                 \
    do {                          \
        int r_ = (expr);          \
        if (r_ != GPV_OK) return r_; \
    } while (0)

int next_div4(int a) { return (a % 4 == 0) ? a : a + (4 - a % 4); }

// Object::PerformVoxelization grid derivation, src/Object.cpp:3094-3136.
int derive_grid(const float bmin[3], const float bmax[3], float mms, const gpv_params* p,
                gpv_grid_info* info)
{
    if (!p || p->voxel_count < 1) return fail(GPV_ERR_INVALID, "voxel_count must be >= 1");
    if (p->level2 && p->voxel_count2 < 1) return fail(GPV_ERR_INVALID, "voxel_count2 must be >= 1");
    if (p->level2 && p->voxel_count2 > 16)
        return fail(GPV_ERR_INVALID, "voxel_count2 %d exceeds the Level2 limit of 16", p->voxel_count2);
    if (!(mms > 0.0f) || !std::isfinite(mms)) return fail(GPV_ERR_INVALID, "max_model_size must be > 0");
    memset(info, 0, sizeof(*info));
    for (int q = 0; q < 3; q++) {
        info->bmin[q] = bmin[q];
        info->bmax[q] = bmax[q];
    }
    info->max_model_size = mms;
    const float nominal = (float)((double)mms / (1.0 * p->voxel_count));
    for (int q = 0; q < 3; q++) {
        const float extent = bmax[q] - bmin[q];
        const float ratio = extent / nominal;
        if (!(ratio >= 0.0f) || ratio > 16384.0f)
            return fail(GPV_ERR_INVALID, "axis %d: %g divisions is outside [0, 16384]", q, (double)ratio);
        int n = (int)ratio;
        if (n == 0) n++;
        n = next_div4(n);
        info->n[q] = n;
        info->g[q] = (float)((double)(bmax[q] - bmin[q]) / (n * 1.0));
    }
    info->n1 = (int64_t)info->n[0] * info->n[1] * info->n[2];
    if (p->level2) {
        info->n2 = p->voxel_count2;
        for (int q = 0; q < 3; q++) info->g2[q] = (float)((double)info->g[q] / (p->voxel_count2 * 1.0));
    }
    return GPV_OK;
}

// Center tables (float, computed in double exactly as the reference does):
//   L1 SAT  (p + 0.5)*ext*2 + min          src/Object.cpp:2329-2331
//   L1 ray  min + (i + 0.5)*g              src/Object.cpp:743-745
//   L2 SAT  (p + 0.5)*g2 + mid - ext1      src/Object.cpp:2369-2371, 2389-2391
//   L2 ray  minPoint + (p + 0.5)*g2        src/Object.cpp:1134-1136, 1158-1160
void build_tables(const gpv_grid_info& I, std::vector<float>& tab, size_t off[12])
{
    const int n2 = I.n2 > 0 ? I.n2 : 0;
    size_t pos = 0;
    for (int a = 0; a < 3; a++) { off[a] = pos; pos += I.n[a]; }
    for (int a = 0; a < 3; a++) { off[3 + a] = pos; pos += I.n[a]; }
    for (int a = 0; a < 3; a++) { off[6 + a] = pos; pos += (size_t)I.n[a] * n2; }
    for (int a = 0; a < 3; a++) { off[9 + a] = pos; pos += (size_t)I.n[a] * n2; }
    tab.assign(pos + 1, 0.0f);
    for (int a = 0; a < 3; a++) {
        const float ext1 = (float)((double)I.g[a] / 2.0);
        for (int p = 0; p < I.n[a]; p++) {
            tab[off[a] + p] = (float)((p + 0.5) * (double)ext1 * 2 + (double)I.bmin[a]);
            tab[off[3 + a] + p] = (float)((double)I.bmin[a] + (p + 0.5) * (double)I.g[a]);
        }
        for (int i = 0; i < I.n[a] && n2 > 0; i++) {
            const float mid = (float)((i + 0.5) * (double)I.g[a] + (double)I.bmin[a]);
            const float minp = (float)i * I.g[a] + I.bmin[a];
            for (int p = 0; p < n2; p++) {
                tab[off[6 + a] + (size_t)i * n2 + p] =
                    (float)((p + 0.5) * (double)I.g2[a] + (double)mid - (double)ext1);
                tab[off[9 + a] + (size_t)i * n2 + p] = (float)((double)minp + (p + 0.5) * (double)I.g2[a]);
            }
        }
    }
}

int expand_blocks(const gpv_ctx* c) { return std::max(1, c->num_cu) * 8; }

int ensure_scan_tmp(gpv_grid* G, int64_t n)
{
    const size_t need = scan_tmp_elems(n);
    if (need <= G->scan_cap) return GPV_OK;
    release(G, G->scan64);
    G->scan_cap = need;
    TRY(alloc(G, &G->scan64, need));
    G->scan32 = reinterpret_cast<uint32_t*>(G->scan64);   // u32 scans use the same storage
    return GPV_OK;
}

// per-triangle setup shared by the SAT and the ray paths
int stage_setup(gpv_grid* G)
{
    if (G->setup_done) return GPV_OK;
    gpv_ctx* c = G->ctx;
    hipStream_t s = c->stream;
    HIP_TRY(hipEventRecord(G->ev[0], s));
    launch_fill_u64(reinterpret_cast<uint64_t*>(G->ctr), 0, sizeof(Counters) / 8, s);
    launch_tri_setup(G->tris, G->T, G->g, G->box0, G->box1, G->sat_items, G->rng, G->ray_items, G->ctr, s);
    TRY(ensure_scan_tmp(G, G->T));
    launch_scan_u32_u64(G->sat_items, G->sat_off, G->T, G->scan64, s);
    launch_scan_u32_u64(G->ray_items, G->ray_off, G->T, G->scan64, s);
    HIP_TRY(hipGetLastError());
    G->setup_done = true;
    return GPV_OK;
}

int make_appends(gpv_grid* G)
{
    TRY(alloc(G, &G->pair_key, G->pair_cap));
    TRY(alloc(G, &G->col_key, G->col_cap));
    TRY(alloc(G, &G->col_z, G->col_cap));
    return GPV_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* gpv_last_error(void) { return g_err.c_str(); }

const char* gpv_status_string(int s)
{
    switch (s) {
    case GPV_OK: return "ok";
    case GPV_ERR_INVALID: return "invalid argument";
    case GPV_ERR_OOM: return "out of memory";
    case GPV_ERR_CAPACITY: return "bin capacity did not converge";
    case GPV_ERR_HIP: return "HIP runtime error";
    case GPV_ERR_NO_DEVICE: return "no usable device";
    case GPV_ERR_IO: return "I/O error";
    case GPV_ERR_STATE: return "stage called out of order";
    default: return "unknown status";
    }
}

int gpv_ctx_create(int device, unsigned flags, gpv_ctx** out)
{
    (void)flags;
    if (!out) return fail(GPV_ERR_INVALID, "out is NULL");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return fail(GPV_ERR_NO_DEVICE, "no HIP device visible");
    }
    if (device < 0) HIP_TRY(hipGetDevice(&device));
    if (device >= n) return fail(GPV_ERR_INVALID, "device %d out of range (%d devices)", device, n);
    HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(GPV_ERR_NO_DEVICE, "device %d is %s; this build targets gfx950", device, prop.gcnArchName);
    gpv_ctx* c = new gpv_ctx();
    c->device = device;
    c->num_cu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipHostMalloc((void**)&c->h_ctr, sizeof(Counters) + 64, hipHostMallocDefault) != hipSuccess) {
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return fail(GPV_ERR_HIP, "stream / pinned buffer creation failed");
    }
    c->h_nb = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(c->h_ctr) + sizeof(Counters) + 32);
    *out = c;
    return GPV_OK;
}

int gpv_ctx_destroy(gpv_ctx* c)
{
    if (!c) return GPV_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    c->pool.trim();
    if (c->h_ctr) (void)hipHostFree(c->h_ctr);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return GPV_OK;
}

int gpv_device_upload(gpv_ctx* c, const void* host, int64_t bytes, void** dev)
{
    if (!c || !dev || bytes < 0 || (bytes > 0 && !host)) return fail(GPV_ERR_INVALID, "bad upload arguments");
    HIP_TRY(hipSetDevice(c->device));
    void* p = nullptr;
    hipError_t e = c->pool.get((size_t)std::max<int64_t>(bytes, 1), &p);
    if (e != hipSuccess) return fail(GPV_ERR_OOM, "device allocation of %lld bytes failed", (long long)bytes);
    if (bytes > 0) {
        e = hipMemcpyAsync(p, host, (size_t)bytes, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) {
            c->pool.put(p);
            return fail(GPV_ERR_HIP, "upload: %s", hipGetErrorString(e));
        }
    }
    *dev = p;
    return GPV_OK;
}

int gpv_device_free(gpv_ctx* c, void* dev)
{
    if (!c) return fail(GPV_ERR_INVALID, "NULL context");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->pool.put(dev);
    return GPV_OK;
}

// Object::ReadObject bbox and pad, src/Object.cpp:441-450 and 572-583.
int gpv_mesh_bbox(const float* verts, int64_t nv, float bmin[3], float bmax[3], float* mms)
{
    if (!verts || nv <= 0 || !bmin || !bmax || !mms) return fail(GPV_ERR_INVALID, "empty vertex list");
    float mn[3], mx[3];
    for (int q = 0; q < 3; q++) mn[q] = mx[q] = verts[q];
    for (int64_t i = 1; i < nv; i++) {
        const float* p = verts + 3 * i;
        for (int q = 0; q < 3; q++) {
            mn[q] = mn[q] < p[q] ? mn[q] : p[q];   // MinFloat3, includes/FloatVector.h:344-351
            mx[q] = mx[q] > p[q] ? mx[q] : p[q];   // MaxFloat3, includes/FloatVector.h:353-360
        }
    }
    const float d0 = mx[0] - mn[0], d1 = mx[1] - mn[1], d2 = mx[2] - mn[2];
    const float sq = d0 * d0 + d1 * d1 + d2 * d2;   // VectorMagnitude, includes/FloatVector.h:323-326
    const float mag = (float)std::sqrt((double)sq);
    const double model_size = mag;
    const float offset = (float)(0.001 * model_size);
    for (int q = 0; q < 3; q++) {
        bmax[q] = offset + mx[q];
        bmin[q] = mn[q] - offset;
    }
    const float ex = bmax[0] - bmin[0], ey = bmax[1] - bmin[1], ez = bmax[2] - bmin[2];
    const float m12 = ey > ez ? ey : ez;
    *mms = ex > m12 ? ex : m12;
    return GPV_OK;
}

int gpv_grid_derive(const float bmin[3], const float bmax[3], float mms, const gpv_params* params,
                    gpv_grid_info* info)
{
    if (!bmin || !bmax || !info) return fail(GPV_ERR_INVALID, "NULL argument");
    return derive_grid(bmin, bmax, mms, params, info);
}

int gpv_grid_create(gpv_ctx* c, const float* tris, int64_t T, int on_device, const float bmin[3],
                    const float bmax[3], float mms, const gpv_params* params, gpv_grid** out)
{
    if (!c || !out || !bmin || !bmax || !params) return fail(GPV_ERR_INVALID, "NULL argument");
    if (T < 0 || (T > 0 && !tris)) return fail(GPV_ERR_INVALID, "bad triangle soup");
    if (T >= (int64_t(1) << 31)) return fail(GPV_ERR_INVALID, "at most 2^31-1 triangles");
    *out = nullptr;
    gpv_grid_info info;
    TRY(derive_grid(bmin, bmax, mms, params, &info));
    HIP_TRY(hipSetDevice(c->device));
    gpv_grid* G = new gpv_grid();
    G->ctx = c;
    G->info = info;
    G->params = *params;
    G->T = T;
    auto bail = [&](int r) {
        for (void* p : G->bufs) c->pool.put(p);
        for (auto& e : G->ev) if (e) (void)hipEventDestroy(e);
        delete G;
        return r;
    };
    for (auto& e : G->ev)
        if (hipEventCreate(&e) != hipSuccess) return bail(fail(GPV_ERR_HIP, "event creation failed"));

    const int nx = info.n[0], ny = info.n[1], nz = info.n[2];
    const int nxw = (nx + 63) / 64;
    const int n2 = info.n2;
    GridDev& g = G->g;
    g.nx = nx; g.ny = ny; g.nz = nz; g.nxw = nxw;
    g.n2 = n2;
    g.w2 = n2 > 0 ? (n2 * n2 * n2 + 63) / 64 : 0;
    g.bminx = info.bmin[0]; g.bminy = info.bmin[1]; g.bminz = info.bmin[2];
    g.bmaxx = info.bmax[0]; g.bmaxy = info.bmax[1]; g.bmaxz = info.bmax[2];
    g.h1x = (float)((double)info.g[0] / 2.0); g.h1y = (float)((double)info.g[1] / 2.0);
    g.h1z = (float)((double)info.g[2] / 2.0);
    g.h2x = (float)((double)info.g2[0] / 2.0); g.h2y = (float)((double)info.g2[1] / 2.0);
    g.h2z = (float)((double)info.g2[2] / 2.0);
    if ((int64_t)nx * ny >= (int64_t(1) << 31)) return bail(fail(GPV_ERR_INVALID, "too many columns"));

    std::vector<float> tab;
    size_t off[12];
    build_tables(info, tab, off);
    int r;
    if ((r = alloc(G, &G->tables, tab.size())) != GPV_OK) return bail(r);
    if (hipMemcpyAsync(G->tables, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice, c->stream) !=
        hipSuccess)
        return bail(fail(GPV_ERR_HIP, "table upload failed"));
    const float* tb = G->tables;
    g.c1x = tb + off[0]; g.c1y = tb + off[1]; g.c1z = tb + off[2];
    g.r1x = tb + off[3]; g.r1y = tb + off[4]; g.r1z = tb + off[5];
    g.c2x = tb + off[6]; g.c2y = tb + off[7]; g.c2z = tb + off[8];
    g.r2x = tb + off[9]; g.r2y = tb + off[10]; g.r2z = tb + off[11];

    if (on_device) {
        G->tris = tris;
    } else if (T > 0) {
        float* d = nullptr;
        if ((r = alloc(G, &d, (size_t)T * 9)) != GPV_OK) return bail(r);
        if (hipMemcpyAsync(d, tris, (size_t)T * 9 * sizeof(float), hipMemcpyHostToDevice, c->stream) != hipSuccess)
            return bail(fail(GPV_ERR_HIP, "soup upload failed"));
        G->tris = d;
        G->own_tris = true;
    }
    const size_t Tn = (size_t)std::max<int64_t>(T, 1);
    G->W = (int64_t)nz * ny * nxw;
    if ((r = alloc(G, &G->box0, Tn)) || (r = alloc(G, &G->box1, Tn)) || (r = alloc(G, &G->rng, Tn)) ||
        (r = alloc(G, &G->sat_items, Tn)) || (r = alloc(G, &G->ray_items, Tn)) ||
        (r = alloc(G, &G->sat_off, Tn + 1)) || (r = alloc(G, &G->ray_off, Tn + 1)) ||
        (r = alloc(G, &G->ctr, 1)) || (r = alloc(G, &G->bnd, (size_t)G->W)) ||
        (r = alloc(G, &G->ins, (size_t)G->W)) || (r = alloc(G, &G->wcnt, (size_t)G->W)) ||
        (r = alloc(G, &G->wprefix, (size_t)G->W + 1)))
        return bail(r);
    const int64_t cap0 = params->initial_capacity > 0 ? params->initial_capacity
                                                      : std::max<int64_t>(int64_t(1) << 16, 16 * T);
    G->pair_cap = (uint64_t)cap0;
    G->col_cap = (uint64_t)cap0;
    G->hit_cap = (uint64_t)std::max<int64_t>(int64_t(1) << 16, 8 * T);
    if ((r = make_appends(G)) != GPV_OK) return bail(r);
    if ((r = alloc(G, &G->hit_key, G->hit_cap)) != GPV_OK) return bail(r);
    G->info.n1 = (int64_t)nx * ny * nz;
    *out = G;
    return GPV_OK;
}

// ---------------------------------------------------------------------------
// stages, split into launch and completion halves so that the fused path
// (gpv_voxelize) reads every data-dependent size in one host round trip
// ---------------------------------------------------------------------------
namespace {

// Level1 SAT sweep + the boundary popcount scan (ranks need the final bits)
int launch_sat_half(gpv_grid* G)
{
    gpv_ctx* c = G->ctx;
    hipStream_t s = c->stream;
    HIP_TRY(hipEventRecord(G->ev[1], s));
    launch_fill_u64(G->bnd, 0, G->W, s);
    launch_fill_u64(reinterpret_cast<uint64_t*>(&G->ctr->pairs), 0, 2, s);  // pairs, colents
    Append64 pa{&G->ctr->pairs, G->pair_key, nullptr, G->pair_cap};
    Append64 co{&G->ctr->colents, G->col_key, G->col_z, G->col_cap};
    launch_sat_expand(G->tris, G->T, G->box0, G->box1, G->sat_off, G->g, G->bnd, pa, co, expand_blocks(c), s);
    launch_popc_words(G->bnd, G->W, G->wcnt, s);
    TRY(ensure_scan_tmp(G, G->W));
    launch_scan_u32_u32(G->wcnt, G->wprefix, G->W, G->scan32, nullptr, s);
    HIP_TRY(hipGetLastError());
    return GPV_OK;
}

int launch_ray_half(gpv_grid* G)
{
    gpv_ctx* c = G->ctx;
    hipStream_t s = c->stream;
    HIP_TRY(hipEventRecord(G->ev[2], s));
    launch_fill_u64(reinterpret_cast<uint64_t*>(&G->ctr->hits), 0, 1, s);
    Append64 hi{&G->ctr->hits, G->hit_key, nullptr, G->hit_cap};
    launch_ray_expand(G->tris, G->T, G->rng, G->ray_off, G->g, hi, expand_blocks(c), s);
    HIP_TRY(hipGetLastError());
    return GPV_OK;
}

// one round trip: the counters and, when asked, the boundary count B
int read_counts(gpv_grid* G, bool with_b)
{
    gpv_ctx* c = G->ctx;
    HIP_TRY(hipMemcpyAsync(c->h_ctr, G->ctr, sizeof(Counters), hipMemcpyDeviceToHost, c->stream));
    if (with_b) HIP_TRY(hipMemcpyAsync(c->h_nb, G->wprefix + G->W, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return GPV_OK;
}

// after read_counts: accept the sweep or regrow the bin buffers to the exact
// count (the reference's retry, src/Object.cpp:3198-3214)
int complete_sat(gpv_grid* G, int64_t* needed)
{
    gpv_ctx* c = G->ctx;
    const uint64_t np = c->h_ctr->pairs, nc = c->h_ctr->colents;
    G->info.tests_l1 = (int64_t)c->h_ctr->tests_l1;
    *needed = 0;
    if (np > G->pair_cap || nc > G->col_cap) {
        const uint64_t need = std::max(np, nc);
        release(G, G->pair_key);
        release(G, G->col_key);
        release(G, G->col_z);
        G->pair_cap = need + need / 8 + 1024;
        G->col_cap = G->pair_cap;
        TRY(make_appends(G));
        G->info.regrows++;
        *needed = (int64_t)need;
        return GPV_OK;
    }
    G->n_pairs = np;
    G->n_cols = nc;
    G->info.pairs = (int64_t)np;
    G->B = *c->h_nb;
    G->info.nb = G->B;
    G->info.l1_boundary = G->B;
    G->sat_done = true;
    return GPV_OK;
}

// after read_counts: false (and a regrown buffer) when the hits overflowed
int complete_ray(gpv_grid* G, bool* ok)
{
    const uint64_t nh = G->ctx->h_ctr->hits;
    *ok = nh <= G->hit_cap;
    if (*ok) {
        G->n_hits = nh;
        return GPV_OK;
    }
    release(G, G->hit_key);
    G->hit_cap = nh + nh / 8 + 1024;
    return alloc(G, &G->hit_key, G->hit_cap);
}

// hits grouped by column, then the parity of every Level1 voxel
int launch_parity(gpv_grid* G)
{
    hipStream_t s = G->ctx->stream;
    HIP_TRY(hipEventRecord(G->ev[3], s));
    const int64_t ncol = (int64_t)G->g.nx * G->g.ny;
    release(G, G->hcnt);
    release(G, G->hoff);
    release(G, G->hrec);
    TRY(alloc(G, &G->hcnt, (size_t)ncol));
    TRY(alloc(G, &G->hoff, (size_t)ncol + 1));
    TRY(alloc(G, &G->hrec, (size_t)std::max<uint64_t>(G->n_hits, 1) * 3));
    launch_fill_u32(G->hcnt, 0, ncol, s);
    launch_count_keys_hi(G->hit_key, G->n_hits, G->hcnt, s);
    TRY(ensure_scan_tmp(G, ncol));
    launch_scan_u32_u32(G->hcnt, G->hoff, ncol, G->scan32, nullptr, s);
    launch_scatter_hits(G->tris, G->hit_key, G->n_hits, G->g, G->hoff, G->hcnt, G->hrec, s);
    launch_l1_parity(G->hrec, G->hoff, G->g, G->ins, s);
    HIP_TRY(hipGetLastError());
    G->inout_done = true;
    return GPV_OK;
}

// boundary index list, bins per boundary voxel, column entries with records
int launch_compact(gpv_grid* G)
{
    hipStream_t s = G->ctx->stream;
    HIP_TRY(hipEventRecord(G->ev[4], s));
    release(G, G->bidx);
    TRY(alloc(G, &G->bidx, (size_t)std::max<int64_t>(G->B, 1)));
    launch_boundary_index(G->bnd, G->wprefix, G->g, G->bidx, s);
    release(G, G->pb);
    release(G, G->cntb);
    release(G, G->boff);
    release(G, G->list);
    TRY(alloc(G, &G->pb, (size_t)std::max<uint64_t>(G->n_pairs, 1)));
    TRY(alloc(G, &G->cntb, (size_t)std::max<int64_t>(G->B, 1)));
    TRY(alloc(G, &G->boff, (size_t)G->B + 1));
    TRY(alloc(G, &G->list, (size_t)std::max<uint64_t>(G->n_pairs, 1)));
    launch_fill_u32(G->cntb, 0, G->B, s);
    launch_fill_u32(&G->ctr->max_per_voxel, 0, 1, s);
    launch_pair_rank(G->pair_key, G->n_pairs, G->bnd, G->wprefix, G->pb, G->cntb, s);
    TRY(ensure_scan_tmp(G, G->B));
    launch_scan_u32_u32(G->cntb, G->boff, G->B, G->scan32, &G->ctr->max_per_voxel, s);
    launch_pair_scatter(G->pair_key, G->pb, G->n_pairs, G->boff, G->cntb, G->list, s);
    const int64_t ncol = (int64_t)G->g.nx * G->g.ny;
    release(G, G->ccnt);
    release(G, G->coff);
    release(G, G->crec);
    TRY(alloc(G, &G->ccnt, (size_t)ncol));
    TRY(alloc(G, &G->coff, (size_t)ncol + 1));
    TRY(alloc(G, &G->crec, (size_t)std::max<uint64_t>(G->n_cols, 1) * 3));
    launch_fill_u32(G->ccnt, 0, ncol, s);
    launch_count_keys_hi(G->col_key, G->n_cols, G->ccnt, s);
    TRY(ensure_scan_tmp(G, ncol));
    launch_scan_u32_u32(G->ccnt, G->coff, ncol, G->scan32, nullptr, s);
    launch_scatter_cols(G->tris, G->col_key, G->col_z, G->n_cols, G->coff, G->ccnt, G->crec, s);
    HIP_TRY(hipGetLastError());
    if (G->g.n2 > 0) {
        release(G, G->sat2);
        release(G, G->par2);
        const size_t w2n = (size_t)std::max<int64_t>(G->B * G->g.w2, 1);
        TRY(alloc(G, &G->sat2, w2n));
        TRY(alloc(G, &G->par2, w2n));
    }
    G->compact_done = true;
    return GPV_OK;
}

}  // namespace

int gpv_classify_tessellation(gpv_ctx* c, gpv_grid* G, int64_t* needed)
{
    if (!c || !G || G->ctx != c) return fail(GPV_ERR_INVALID, "bad context/grid");
    int64_t nd = 0;
    HIP_TRY(hipSetDevice(c->device));
    TRY(stage_setup(G));
    TRY(launch_sat_half(G));
    TRY(read_counts(G, true));
    TRY(complete_sat(G, &nd));
    if (needed) *needed = nd;
    return GPV_OK;
}

int gpv_classify_inout(gpv_ctx* c, gpv_grid* G)
{
    if (!c || !G || G->ctx != c) return fail(GPV_ERR_INVALID, "bad context/grid");
    HIP_TRY(hipSetDevice(c->device));
    TRY(stage_setup(G));
    bool ok = false;
    for (int attempt = 0; attempt < 2 && !ok; ++attempt) {
        TRY(launch_ray_half(G));
        TRY(read_counts(G, false));
        TRY(complete_ray(G, &ok));
    }
    if (!ok) return fail(GPV_ERR_CAPACITY, "ray hit capacity did not converge");
    return launch_parity(G);
}

int gpv_compact_boundary(gpv_ctx* c, gpv_grid* G)
{
    if (!c || !G || G->ctx != c) return fail(GPV_ERR_INVALID, "bad context/grid");
    if (!G->sat_done) return fail(GPV_ERR_STATE, "gpv_classify_tessellation has not completed");
    HIP_TRY(hipSetDevice(c->device));
    return launch_compact(G);
}

int gpv_classify_inout_level2(gpv_ctx* c, gpv_grid* G)
{
    if (!c || !G || G->ctx != c) return fail(GPV_ERR_INVALID, "bad context/grid");
    if (!G->compact_done) return fail(GPV_ERR_STATE, "gpv_compact_boundary has not completed");
    if (G->g.n2 <= 0) return GPV_OK;
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipEventRecord(G->ev[5], c->stream));
    launch_l2_parity(G->bidx, G->B, G->coff, G->crec, G->g, G->par2, c->stream);
    HIP_TRY(hipGetLastError());
    G->l2p_done = true;
    return GPV_OK;
}

int gpv_classify_tessellation_level2(gpv_ctx* c, gpv_grid* G)
{
    if (!c || !G || G->ctx != c) return fail(GPV_ERR_INVALID, "bad context/grid");
    if (!G->compact_done) return fail(GPV_ERR_STATE, "gpv_compact_boundary has not completed");
    if (G->g.n2 <= 0) return GPV_OK;
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipEventRecord(G->ev[6], c->stream));
    launch_l2_sat(G->tris, G->bidx, G->B, G->boff, G->list, G->g, G->sat2, c->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(G->ev[7], c->stream));
    G->l2s_done = true;
    return GPV_OK;
}

int gpv_grid_finish(gpv_ctx* c, gpv_grid* G)
{
    if (!c || !G || G->ctx != c) return fail(GPV_ERR_INVALID, "bad context/grid");
    if (!G->compact_done || !G->inout_done) return fail(GPV_ERR_STATE, "Level1 stages have not completed");
    if (G->g.n2 > 0 && (!G->l2p_done || !G->l2s_done)) return fail(GPV_ERR_STATE, "Level2 stages have not completed");
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    launch_fill_u64(reinterpret_cast<uint64_t*>(&G->ctr->l1_inside), 0, 3, s);
    const int64_t W2 = G->g.n2 > 0 ? G->B * G->g.w2 : 0;
    launch_count_states(G->bnd, G->ins, G->W, G->sat2, G->par2, W2, G->ctr, s);
    HIP_TRY(hipEventRecord(G->ev[8], s));
    TRY(read_counts(G, false));
    const Counters& h = *c->h_ctr;
    G->info.l1_inside = (int64_t)h.l1_inside;
    G->info.l1_boundary = G->B;
    G->info.l2_inside = (int64_t)h.l2_inside;
    G->info.l2_boundary = (int64_t)h.l2_boundary;
    G->info.max_tris_per_voxel = (int32_t)h.max_per_voxel;
    G->info.tests_l2 = G->g.n2 > 0 ? (int64_t)G->n_pairs * G->g.n2 * G->g.n2 * G->g.n2 : 0;
    auto el = [&](int a, int b) -> float {
        float v = 0.0f;
        if (hipEventElapsedTime(&v, G->ev[a], G->ev[b]) != hipSuccess) {
            (void)hipGetLastError();
            v = 0.0f;
        }
        return v;
    };
    // setup, L1 SAT (+ boundary scan), L1 parity (ray expansion, the size
    // read-back and the parity fill), compaction + bins, L2 parity, L2 SAT
    G->info.ms_stage[0] = el(0, 1);
    G->info.ms_stage[1] = el(1, 2);
    G->info.ms_stage[2] = el(2, 4);
    if (G->g.n2 > 0) {
        G->info.ms_stage[3] = el(4, 5);
        G->info.ms_stage[4] = el(5, 6);
        G->info.ms_stage[5] = el(6, 7);
    } else {
        G->info.ms_stage[3] = el(4, 8);
    }
    G->info.ms_total = el(0, 8);
    G->finished = true;
    return GPV_OK;
}

int gpv_voxelize(gpv_ctx* c, const float* tris, int64_t T, int on_device, const float bmin[3],
                 const float bmax[3], float mms, const gpv_params* params, gpv_grid** out)
{
    gpv_grid* G = nullptr;
    TRY(gpv_grid_create(c, tris, T, on_device, bmin, bmax, mms, params, &G));
    auto run = [&]() -> int {
        TRY(stage_setup(G));
        // SAT sweep, ray expansion and boundary scan back to back, one read-back
        TRY(launch_sat_half(G));
        TRY(launch_ray_half(G));
        TRY(read_counts(G, true));
        int64_t needed = 0;
        bool ray_ok = false;
        TRY(complete_sat(G, &needed));
        TRY(complete_ray(G, &ray_ok));
        for (int attempt = 0; attempt < 2 && (needed > 0 || !ray_ok); ++attempt) {
            if (needed > 0) TRY(launch_sat_half(G));
            if (!ray_ok) TRY(launch_ray_half(G));
            TRY(read_counts(G, true));
            if (needed > 0) TRY(complete_sat(G, &needed));
            if (!ray_ok) TRY(complete_ray(G, &ray_ok));
        }
        if (needed > 0 || !ray_ok) return fail(GPV_ERR_CAPACITY, "buffer capacity did not converge");
        TRY(launch_parity(G));
        TRY(launch_compact(G));
        if (G->g.n2 > 0) {
            TRY(gpv_classify_inout_level2(c, G));
            TRY(gpv_classify_tessellation_level2(c, G));
        }
        return gpv_grid_finish(c, G);
    };
    const int r = run();
    if (r != GPV_OK) {
        std::string keep = g_err;
        gpv_grid_release(c, G);
        g_err = keep;
        return r;
    }
    *out = G;
    return GPV_OK;
}

int gpv_grid_get_info(const gpv_grid* G, gpv_grid_info* info)
{
    if (!G || !info) return fail(GPV_ERR_INVALID, "NULL argument");
    *info = G->info;
    return GPV_OK;
}

int gpv_grid_device_views(const gpv_grid* G, const uint64_t** l1b, const uint64_t** l1i, const uint64_t** l2b,
                          const uint64_t** l2i, const int32_t** bidx)
{
    if (!G) return fail(GPV_ERR_INVALID, "NULL grid");
    if (l1b) *l1b = G->bnd;
    if (l1i) *l1i = G->ins;
    if (l2b) *l2b = G->sat2;
    if (l2i) *l2i = G->par2;
    if (bidx) *bidx = G->bidx;
    return GPV_OK;
}

static int download_impl(gpv_ctx* c, const gpv_grid* Gc, uint8_t unit, uint8_t* l1, int32_t* prefix,
                         int32_t* bidx, uint8_t* l2)
{
    if (!c || !Gc || Gc->ctx != c) return fail(GPV_ERR_INVALID, "bad context/grid");
    if (!Gc->finished) return fail(GPV_ERR_STATE, "grid is not finished");
    gpv_grid* G = const_cast<gpv_grid*>(Gc);
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int64_t n1 = G->info.n1;
    const int64_t n23 = (int64_t)G->g.n2 * G->g.n2 * G->g.n2;
    if (l1) {
        uint8_t* d = nullptr;
        TRY(alloc(G, &d, (size_t)n1));
        launch_export_l1(G->bnd, G->ins, G->g, unit, d, s);
        HIP_TRY(hipMemcpyAsync(l1, d, (size_t)n1, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        release(G, d);
    }
    if (prefix) {
        int32_t* d = nullptr;
        TRY(alloc(G, &d, (size_t)n1));
        launch_export_prefix(G->bnd, G->wprefix, G->g, d, s);
        HIP_TRY(hipMemcpyAsync(prefix, d, (size_t)n1 * 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        release(G, d);
    }
    if (bidx && G->B > 0) {
        HIP_TRY(hipMemcpyAsync(bidx, G->bidx, (size_t)G->B * 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    if (l2 && G->g.n2 > 0 && G->B > 0) {
        uint8_t* d = nullptr;
        TRY(alloc(G, &d, (size_t)(G->B * n23)));
        launch_export_l2(G->sat2, G->par2, G->B, G->g, unit, d, s);
        HIP_TRY(hipMemcpyAsync(l2, d, (size_t)(G->B * n23), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        release(G, d);
    }
    return GPV_OK;
}

// states 0 outside / 1 inside / 2 boundary
int gpv_grid_download(gpv_ctx* c, const gpv_grid* G, uint8_t* l1, int32_t* prefix, int32_t* bidx, uint8_t* l2)
{
    return download_impl(c, G, 1, l1, prefix, bidx, l2);
}

// Object::SaveVoxelization, src/Object.cpp:2934-3075 (in/out, prefix and
// config files; the config uses the ostream default of 6 significant digits).
int gpv_save_raw(gpv_ctx* c, const gpv_grid* G, int obj_id, const char* dir)
{
    if (!c || !G) return fail(GPV_ERR_INVALID, "NULL argument");
    const gpv_grid_info& I = G->info;
    const bool l2 = G->g.n2 > 0;
    const int64_t n23 = (int64_t)G->g.n2 * G->g.n2 * G->g.n2;
    std::vector<uint8_t> l1v((size_t)I.n1);
    std::vector<int32_t> pre(l2 ? (size_t)I.n1 : 0);
    std::vector<uint8_t> l2v(l2 ? (size_t)(G->B * n23) : 0);
    // uint8(state * 127.0f), src/Object.cpp:3031-3034, 3046-3050
    TRY(download_impl(c, G, 127, l1v.data(), l2 ? pre.data() : nullptr, nullptr, l2 ? l2v.data() : nullptr));
    const std::string base = std::string(dir && *dir ? dir : ".") + "/Obj" + std::to_string(obj_id);
    FILE* f = fopen((base + "VoxelConfig.txt").c_str(), "w");
    if (!f) return fail(GPV_ERR_IO, "cannot open %sVoxelConfig.txt", base.c_str());
    fprintf(f, "Obj%d\n", obj_id);
    fprintf(f, "%g\t%g\t%g\n", I.bmin[0], I.bmin[1], I.bmin[2]);
    fprintf(f, "%g\t%g\t%g\n", I.bmax[0], I.bmax[1], I.bmax[2]);
    fprintf(f, "%d\t%d\t%d\n", I.n[0], I.n[1], I.n[2]);
    fprintf(f, "%g\t%g\t%g\n", I.g[0], I.g[1], I.g[2]);
    fprintf(f, "%lld\n%lld\n", (long long)I.l1_inside, (long long)I.l1_boundary);
    if (l2) {
        fprintf(f, "%d\t%d\t%d\n", I.n2, I.n2, I.n2);
        fprintf(f, "%g\t%g\t%g\n", I.g2[0], I.g2[1], I.g2[2]);
        fprintf(f, "%lld\n%lld\n", (long long)I.l2_inside, (long long)I.l2_boundary);
    }
    fclose(f);
    auto put = [&](const std::string& name, const void* data, size_t bytes) -> int {
        FILE* o = fopen((base + name).c_str(), "wb");
        if (!o) return fail(GPV_ERR_IO, "cannot open %s%s", base.c_str(), name.c_str());
        const size_t w = bytes ? fwrite(data, 1, bytes, o) : 0;
        fclose(o);
        return w == bytes ? GPV_OK : fail(GPV_ERR_IO, "short write on %s%s", base.c_str(), name.c_str());
    };
    TRY(put("Level1InOut.raw", l1v.data(), l1v.size()));
    if (l2) {
        TRY(put("Level1BoundaryPrefixSum.raw", pre.data(), pre.size() * 4));
        TRY(put("Level2InOut.raw", l2v.data(), l2v.size()));
    }
    return GPV_OK;
}

int gpv_grid_release(gpv_ctx* c, gpv_grid* G)
{
    if (!G) return GPV_OK;
    if (!c || G->ctx != c) return fail(GPV_ERR_INVALID, "grid belongs to another context");
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);   // pooled blocks may be reused right away
    for (void* p : G->bufs) c->pool.put(p);
    for (auto& e : G->ev)
        if (e) (void)hipEventDestroy(e);
    delete G;
    return GPV_OK;
}

}  // extern "C"
