// gpv_runtime.cpp -- host side of libgpv: contexts, device pool, grid
// derivation, the staged pipeline and the C ABI (include/gpv.h).
//
// Host float semantics matter here too: bbox, grid sizes and center tables
// are computed with the reference's float/double promotions (see the
// citations per function) and this file is compiled with -ffp-contract=off.
#include "gpv.h"
#include "gpv_internal.hpp"
#include "gpv_serve.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
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

}  // namespace

namespace gpv {
int set_error(int code, const char* msg)
{
    g_err = msg ? msg : "";
    return code;
}
}  // namespace gpv

namespace {

#define HIP_TRY(expr)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(e_ == hipErrorOutOfMemory ? GPV_ERR_OOM : GPV_ERR_HIP, "%s: %s", \
                        #expr, hipGetErrorString(e_));                                   \
    } while (0)

// Device memory pool: size classes (powers of two up to 256 MiB, 16 MiB
// granules above), blocks cached until the context is destroyed, so the
// steady state of repeated voxelizations allocates nothing.  A request takes
// the smallest cached block of at least its class and at most twice it, so
// batches of varying size reuse each other's blocks.  Cached (free) bytes are
// held under a cap: a miss that would exceed it first frees the largest
// cached blocks, and a failed allocation frees every cached block and tries
// once more.
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
    void set_cap(size_t cap) { cap_ = cap; }
    // device bytes the pool may hold in use at once (0: no limit); a request
    // beyond it fails as if the device were out of memory
    void set_limit(size_t limit) { limit_ = limit; }
    void reset_peak()
    {
        peak_live_ = live_bytes_;
        peak_req_ = req_bytes_;
    }
    hipError_t get(size_t n, void** out)
    {
        const size_t c = size_class(n);
        auto it = free_.lower_bound(c);
        if (it != free_.end() && it->first <= 2 * c && (!limit_ || live_bytes_ + it->first <= limit_)) {
            *out = it->second;
            live_[*out] = {it->first, n};
            cached_ -= it->first;
            free_.erase(it);
            ++reuses_;
            note_live(it->first, n);
            return hipSuccess;
        }
        if (limit_ && live_bytes_ + c > limit_) return hipErrorOutOfMemory;
        while (!free_.empty() && cached_ + c > cap_) evict(std::prev(free_.end()));
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, c);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            trim();
            e = hipMalloc(&p, c);
            if (e != hipSuccess) return e;
        }
        live_[p] = {c, n};
        bytes_ += c;
        ++allocs_;
        note_live(c, n);
        *out = p;
        return hipSuccess;
    }
    void put(void* p)
    {
        if (!p) return;
        auto it = live_.find(p);
        if (it == live_.end()) return;
        free_.emplace(it->second.first, p);
        cached_ += it->second.first;
        live_bytes_ -= it->second.first;
        req_bytes_ -= it->second.second;
        live_.erase(it);
    }
    void trim()
    {
        while (!free_.empty()) evict(free_.begin());
    }
    void stats(gpv_pool_stats* s) const
    {
        s->held_bytes = (int64_t)bytes_;
        s->cached_bytes = (int64_t)cached_;
        s->cap_bytes = (int64_t)cap_;
        s->allocations = allocs_;
        s->reuses = reuses_;
        s->evictions = evictions_;
        s->live_bytes = (int64_t)live_bytes_;
        s->peak_live_bytes = (int64_t)peak_live_;
        s->limit_bytes = (int64_t)limit_;
        s->requested_bytes = (int64_t)req_bytes_;
        s->peak_requested_bytes = (int64_t)peak_req_;
    }
    size_t live_bytes() const { return live_bytes_; }
    size_t peak_live() const { return peak_live_; }

private:
    void note_live(size_t c, size_t n)
    {
        live_bytes_ += c;
        peak_live_ = std::max(peak_live_, live_bytes_);
        req_bytes_ += n;
        peak_req_ = std::max(peak_req_, req_bytes_);
    }
    void evict(std::multimap<size_t, void*>::iterator it)
    {
        (void)hipFree(it->second);
        bytes_ -= it->first;
        cached_ -= it->first;
        ++evictions_;
        free_.erase(it);
    }
    std::multimap<size_t, void*> free_;
    std::unordered_map<void*, std::pair<size_t, size_t>> live_;   // block -> (class bytes, requested bytes)
    size_t bytes_ = 0, cached_ = 0, cap_ = ~size_t(0), limit_ = 0, live_bytes_ = 0, peak_live_ = 0;
    size_t req_bytes_ = 0, peak_req_ = 0;
    int64_t allocs_ = 0, reuses_ = 0, evictions_ = 0;
};

// Working memory of gpv_voxelize_noalloc: every transient buffer of that
// entry point is carved from this array, so it calls no hipMalloc.  Each
// device that loads the library holds its own copy until the process ends.
// In host code `scratch` only names the symbol; its device address comes
// from hipGetSymbolAddress, once per device (scratch_setup), and only that
// address is ever handed to kernels or copies.
constexpr size_t kScratchBytes = size_t(GPV_NOALLOC_SCRATCH_BYTES);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};
// gpv_voxelize_noalloc calls on one device are serialised: they share the array
std::mutex g_scratch_call_mu[kMaxDevices];

// Looks up this device's copy of the scratch array (the device must be
// current).  Called by gpv_ctx_create; a failure is kept and reported by
// gpv_voxelize_noalloc, the only user.
void scratch_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        g_scratch_err[device] = e;
        return;
    }
    g_scratch_ptr[device] = p;
    g_scratch_err[device] = hipSuccess;
}

}  // namespace

struct gpv_ctx {
    // every entry point that touches the context holds this, so a context
    // may be shared by threads (calls on it are serialised)
    std::recursive_mutex mu;
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;   // second stream: Level2 parity beside the Level2 SAT
    hipStream_t copy = nullptr;   // uploads of a batch's host soups ahead of its runs (created on first use)
    Pool pool;
    char* h_buf = nullptr;   // pinned read-back staging
    size_t h_cap = 0;
    float* tab_buf = nullptr;   // pinned staging of a job's center tables
    size_t tab_cap = 0;
    int num_cu = 256;
    gpv_tune tune{16, 8, 2048, 1, 1, 2, 128};
    std::vector<hipEvent_t> events;   // recycled stage events of finished jobs
    hipEvent_t na_ev[16] = {};        // stage events of gpv_voxelize_noalloc (created with the context)
};

namespace gpv {

void pool_reset_peak(gpv_ctx* c)
{
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    c->pool.reset_peak();
}

// One pipeline run over nm meshes; nm = 1 is a plain voxelization.  Every
// stage is one launch (or one short launch sequence) over all meshes.  The
// per-mesh grids handed to callers (struct gpv_grid) are views into a job;
// the job's device memory returns to the pool with the last view.
struct Job {
    gpv_ctx* ctx = nullptr;
    int refs = 0;
    int nm = 1;
    std::vector<gpv_grid_info> info;   // [nm]
    std::vector<GridDev> g;            // [nm], device table pointers
    std::vector<float> h_tab;          // center tables of a no-allocation call (ordinary host memory)
    std::vector<MeshOff> off;          // [nm + 1]
    int64_t T = 0;
    const float* tris = nullptr;
    float* own_soup = nullptr;
    SoupPiece* soup_pieces = nullptr;   // the gather table of a batch of device soups
    std::vector<SoupPiece> h_pieces;    // its host copy (the source of an asynchronous upload)
    float* tables = nullptr;
    GridDev* d_grids = nullptr;
    MeshOff* d_off = nullptr;
    uint32_t* tri_mesh = nullptr;
    uint32_t* bmesh = nullptr;
    MeshCtr* mctr = nullptr;
    // per triangle
    int4 *box0 = nullptr, *box1 = nullptr, *rng = nullptr;
    double4* slab1 = nullptr;   // plane slab of triangles spanning several layers
    float4* slab2 = nullptr;    // plane slab over the Level2 centers of the index AABB (float form)
    uint32_t *sat_items = nullptr, *ray_items = nullptr;
    uint64_t *sat_off = nullptr, *ray_off = nullptr, *scan64 = nullptr;
    uint64_t* scan_ray = nullptr;   // scan temporary of the ray half of the set-up (side stream)
    uint32_t* scan32 = nullptr;
    size_t scan_cap = 0;
    Counters* ctr = nullptr;
    // Level1 bitplanes (all meshes) and xy columns
    int64_t W = 0, ncol = 0;
    uint64_t *bnd = nullptr, *ins = nullptr;
    // append buffers
    uint64_t* pair_key = nullptr;     // per pair, its voxel's bit; its boundary rank once binned
    uint32_t* pair_tri = nullptr;
    uint64_t pair_cap = 0;
    bool pairs_ranked = false;        // pair_key holds boundary ranks (a compaction has binned the pairs)
    uint64_t* col_key = nullptr;
    uint32_t* col_z = nullptr;
    uint64_t col_cap = 0;
    uint64_t* hit_key = nullptr;
    uint64_t hit_cap = 0;
    uint64_t n_pairs = 0, n_cols = 0, n_hits = 0;
    int regrows = 0;
    // compaction and bins
    uint32_t *wcnt = nullptr, *wprefix = nullptr;
    int32_t* bidx = nullptr;
    int64_t B = 0, W2 = 0;
    uint32_t *boff = nullptr, *list = nullptr;
    uint64_t* pmask = nullptr;   // Level2 box-axis masks per bin entry
    uint32_t* coff = nullptr;
    uint32_t* scan_side = nullptr;   // scan temporary of the parity fill (it may run beside compaction)
    float4* crec = nullptr;   // 3 float4 per column entry
    // column-entry sort on the side stream (beside the bins): its own scratch
    bool cols_side = false;
    uint32_t* cs_cnt = nullptr;
    uint4* cs_bucket = nullptr;
    uint64_t* cpass = nullptr;   // det/u/v pass bits per column entry and sub-column
    uint32_t *gflag = nullptr, *gidx = nullptr, *gstart = nullptr;   // Level2 SAT voxel groups
    int64_t l2p_waves = 0;   // Level2 parity waves (l2_parity_plan)
    uint32_t *hcnt = nullptr, *hoff = nullptr;
    float4* hrec = nullptr;
    // GL-stencil-equivalent Level1 fill (meshes with params.l1_stencil):
    // per-triangle sample counts and offsets, and the slice toggle planes
    bool any_stencil = false, all_stencil = false, stencil_done = false;
    uint32_t* st_cnt = nullptr;
    uint64_t *st_off = nullptr, *st_scan = nullptr, *st_tog = nullptr;
    uint64_t *sat2 = nullptr, *par2 = nullptr;
    int n2max = 0;
    // normals (meshes with params.normals): face normals, sorted bins, and the
    // resident results per boundary voxel / per Level2 voxel
    bool any_normals = false;
    float* fnorm = nullptr;
    uint32_t* slist = nullptr;
    float *l1n = nullptr, *l2n = nullptr;
    int64_t NL2 = 0;
    // progress
    bool setup_done = false, sat_done = false, inout_done = false, compact_done = false,
         l2p_done = false, l2s_done = false, finished = false;
    // stream events: 0 setup, 1 L1 SAT, 2 ray expansion, 3 parity fill,
    // 4 compaction+bins, 5 L2 parity, 6 L2 SAT start, 7 L2 SAT end, 8 finish,
    // 9 normals start, 10 normals end, 11 L2 parity end, 12 L1 SAT end,
    // 13 ray expansion end, 14 parity fill end, 15 fork of the ray expansion
    hipEvent_t ev[16] = {};
    std::vector<void*> bufs;
    // gpv_voxelize_noalloc: buffers are carved from the device scratch array
    // (arena != nullptr) instead of the pool, the result planes are the
    // caller's, and the events are the context's own set
    char* arena = nullptr;
    size_t arena_cap = 0, arena_used = 0;
    bool borrowed_events = false;
    int64_t nb_cap = 0;                 // caller's capacity in boundary voxels
    int32_t* out_bidx = nullptr;
    uint64_t *out_sat2 = nullptr, *out_par2 = nullptr;

    Meshes view() const { return Meshes{d_grids, d_off, tri_mesh, bmesh, mctr, nm}; }
    int64_t nb(int m) const { return (int64_t)off[m + 1].b0 - (int64_t)off[m].b0; }
};

}  // namespace gpv

struct gpv_grid {
    gpv_ctx* ctx = nullptr;
    gpv::Job* job = nullptr;
    int m = 0;
};

namespace {

template <typename T>
int alloc(Job* J, T** p, size_t count)
{
    void* q = nullptr;
    if (J->arena) {
        // bump allocation in 256-byte steps; nothing is returned before the
        // call ends (release is a no-op), so the arena never fragments
        const size_t bytes = (std::max<size_t>(count, 1) * sizeof(T) + 255) / 256 * 256;
        if (bytes > J->arena_cap - J->arena_used)
            return fail(GPV_ERR_TOO_LARGE,
                        "gpv_voxelize_noalloc: this mesh needs more than the %zu bytes of scratch "
                        "(%zu in use, %zu more asked); use gpv_voxelize",
                        J->arena_cap, J->arena_used, bytes);
        *p = reinterpret_cast<T*>(J->arena + J->arena_used);
        J->arena_used += bytes;
        return GPV_OK;
    }
    hipError_t e = J->ctx->pool.get(std::max<size_t>(count, 1) * sizeof(T), &q);
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? GPV_ERR_OOM : GPV_ERR_HIP, "device allocation of %zu bytes: %s",
                    count * sizeof(T), hipGetErrorString(e));
    *p = static_cast<T*>(q);
    J->bufs.push_back(q);
    return GPV_OK;
}

template <typename T>
void release(Job* J, T*& p)
{
    if (!p) return;
    if (J->arena) {
        p = nullptr;
        return;
    }
    void* q = (void*)p;
    J->ctx->pool.put(q);
    auto it = std::find(J->bufs.begin(), J->bufs.end(), q);
    if (it != J->bufs.end()) J->bufs.erase(it);
    p = nullptr;
}

#define TRY(expr)                    \
    do {                             \
        int r_ = (expr);             \
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
// cuda_l2: both Level2 tables in the CUDA kernels' float expression
//   (2p+1)*ext2 + mid - ext1               cuda/CUDAClassifyTessellation.cu:423-425, 472-474
// Returns whether each Level2 ray-origin table is non-decreasing over its
// whole length (the parity kernels then bound a range of origins by its
// first and last entries).
// the twelve tables' offsets; returns the floats they take (plus one)
size_t table_layout(const gpv_grid_info& I, size_t off[12])
{
    const int n2 = I.n2 > 0 ? I.n2 : 0;
    size_t pos = 0;
    for (int a = 0; a < 3; a++) { off[a] = pos; pos += I.n[a]; }
    for (int a = 0; a < 3; a++) { off[3 + a] = pos; pos += I.n[a]; }
    for (int a = 0; a < 3; a++) { off[6 + a] = pos; pos += (size_t)I.n[a] * n2; }
    for (int a = 0; a < 3; a++) { off[9 + a] = pos; pos += (size_t)I.n[a] * n2; }
    return pos + 1;
}

bool fill_tables(const gpv_grid_info& I, bool cuda_l2, float* tab, const size_t off[12])
{
    const int n2 = I.n2 > 0 ? I.n2 : 0;
    for (int a = 0; a < 3; a++) {
        const float ext1 = (float)((double)I.g[a] / 2.0);
        for (int p = 0; p < I.n[a]; p++) {
            tab[off[a] + p] = (float)((p + 0.5) * (double)ext1 * 2 + (double)I.bmin[a]);
            tab[off[3 + a] + p] = (float)((double)I.bmin[a] + (p + 0.5) * (double)I.g[a]);
        }
        for (int i = 0; i < I.n[a] && n2 > 0; i++) {
            const float mid = (float)((i + 0.5) * (double)I.g[a] + (double)I.bmin[a]);
            const float minp = (float)i * I.g[a] + I.bmin[a];
            const float ext2 = (float)((double)I.g2[a] / 2.0);
            for (int p = 0; p < n2; p++) {
                if (cuda_l2) {
                    const float two_p1 = (float)(2 * p + 1);
                    const float prod = two_p1 * ext2;
                    const float sum = prod + mid;
                    const float c = sum - ext1;
                    tab[off[6 + a] + (size_t)i * n2 + p] = c;
                    tab[off[9 + a] + (size_t)i * n2 + p] = c;
                    continue;
                }
                tab[off[6 + a] + (size_t)i * n2 + p] =
                    (float)((p + 0.5) * (double)I.g2[a] + (double)mid - (double)ext1);
                tab[off[9 + a] + (size_t)i * n2 + p] = (float)((double)minp + (p + 0.5) * (double)I.g2[a]);
            }
        }
    }
    bool mono = true;
    for (int a = 0; a < 3; a++)
        for (size_t k = 1; k < (size_t)I.n[a] * n2; k++) mono = mono && tab[off[9 + a] + k - 1] <= tab[off[9 + a] + k];
    return mono;
}

bool build_tables(const gpv_grid_info& I, bool cuda_l2, std::vector<float>& tab, size_t off[12])
{
    tab.assign(table_layout(I, off), 0.0f);
    return fill_tables(I, cuda_l2, tab.data(), off);
}

int expand_blocks(const gpv_ctx* c) { return std::max(1, c->num_cu) * std::max(1, c->tune.expand_blocks_per_cu); }

int ensure_hbuf(gpv_ctx* c, size_t bytes)
{
    if (bytes <= c->h_cap) return GPV_OK;
    if (c->h_buf) (void)hipHostFree(c->h_buf);
    c->h_buf = nullptr;
    c->h_cap = 0;
    bytes = std::max<size_t>(bytes, 4096);
    HIP_TRY(hipHostMalloc((void**)&c->h_buf, bytes, hipHostMallocDefault));
    c->h_cap = bytes;
    return GPV_OK;
}

// pinned staging layout: Counters | MeshOff[nm + 1] | MeshCtr[nm]
constexpr size_t kHOff = 64;
size_t hctr_at(int nm) { return (kHOff + sizeof(MeshOff) * (size_t)(nm + 1) + 63) / 64 * 64; }

int ensure_scan_tmp(Job* J, int64_t n)
{
    const size_t need = scan_tmp_elems(n);
    if (need <= J->scan_cap) return GPV_OK;
    release(J, J->scan64);
    J->scan_cap = need;
    TRY(alloc(J, &J->scan64, need));
    J->scan32 = reinterpret_cast<uint32_t*>(J->scan64);   // u32 scans use the same storage
    return GPV_OK;
}

void job_free(Job* J)
{
    gpv_ctx* c = J->ctx;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);   // pooled blocks may be reused right away
    (void)hipStreamSynchronize(c->side);
    for (void* p : J->bufs) c->pool.put(p);
    if (!J->borrowed_events)
        for (auto& e : J->ev)
            if (e) c->events.push_back(e);   // every use is complete (the streams were synchronised)
    delete J;
}

// GridDev of one mesh from its derived grid (device table pointers set later)
GridDev grid_dev(const gpv_grid_info& info)
{
    GridDev g{};
    g.nx = info.n[0]; g.ny = info.n[1]; g.nz = info.n[2];
    g.nxw = (g.nx + 63) / 64;
    g.n2 = info.n2;
    g.w2 = info.n2 > 0 ? (info.n2 * info.n2 * info.n2 + 63) / 64 : 0;
    g.bminx = info.bmin[0]; g.bminy = info.bmin[1]; g.bminz = info.bmin[2];
    g.bmaxx = info.bmax[0]; g.bmaxy = info.bmax[1]; g.bmaxz = info.bmax[2];
    g.h1x = (float)((double)info.g[0] / 2.0); g.h1y = (float)((double)info.g[1] / 2.0);
    g.h1z = (float)((double)info.g[2] / 2.0);
    g.h2x = (float)((double)info.g2[0] / 2.0); g.h2y = (float)((double)info.g2[1] / 2.0);
    g.h2z = (float)((double)info.g2[2] / 2.0);
    return g;
}

// What gpv_voxelize_noalloc hands a job: the scratch arena and the caller's
// result buffers (Level1 planes of l1_words words each; bidx and the Level2
// planes for nb_cap boundary voxels).
struct NoAlloc {
    char* arena;
    size_t arena_cap;
    uint64_t *l1b, *l1i;
    int64_t l1_words;
    int32_t* bidx;
    uint64_t *l2b, *l2i;
    int64_t nb_cap;
};

// Grids, offsets, tables, the concatenated soup and every size-independent
// buffer of a job.  nm == 1 with a device soup uses the caller's pointer.
// With na (one mesh, device soup) nothing comes from the pool: buffers are
// carved from na->arena and the Level1 planes are the caller's.
int job_create(gpv_ctx* c, const gpv_mesh_desc* d, int nm, int on_device, Job** out, const NoAlloc* na = nullptr)
{
    *out = nullptr;
    if (nm < 1) return fail(GPV_ERR_INVALID, "a batch needs at least one mesh");
    std::vector<gpv_grid_info> info((size_t)nm);
    int64_t T = 0, W = 0, ncol = 0, vox = 0;
    uint64_t pair_cap = 0, hit_cap = 0;
    for (int m = 0; m < nm; ++m) {
        if (d[m].ntris < 0 || (d[m].ntris > 0 && !d[m].tris))
            return fail(GPV_ERR_INVALID, "mesh %d: bad triangle soup", m);
        TRY(derive_grid(d[m].bmin, d[m].bmax, d[m].max_model_size, &d[m].params, &info[m]));
        const int64_t nx = info[m].n[0], ny = info[m].n[1], nz = info[m].n[2];
        if (nx * ny >= (int64_t(1) << 31)) return fail(GPV_ERR_INVALID, "mesh %d: too many columns", m);
        if (d[m].params.l1_stencil && (nx > (int64_t(1) << 19) || ny > (int64_t(1) << 19) || nx * ny >= (int64_t(1) << 29)))
            return fail(GPV_ERR_INVALID, "mesh %d: the stencil fill takes at most 2^19 divisions per axis and 2^29 columns", m);
        info[m].n1 = nx * ny * nz;
        T += d[m].ntris;
        W += nz * ny * ((nx + 63) / 64);
        ncol += nx * ny;
        vox += nz * ny * ((nx + 63) / 64) * 64;
        const int64_t nmax = std::max(nx, std::max(ny, nz));
        // bin-pair capacity: the caller's, or an estimate from the triangle
        // count and the surface voxels a resolution can hold
        const int64_t pc = d[m].params.initial_capacity > 0
                               ? d[m].params.initial_capacity
                               : std::max<int64_t>(int64_t(1) << 16, 16 * d[m].ntris + 8 * nmax * nmax);
        pair_cap += (uint64_t)pc;
        hit_cap += (uint64_t)(8 * d[m].ntris + 4 * nx * ny);
    }
    if (T >= (int64_t(1) << 31)) return fail(GPV_ERR_INVALID, "at most 2^31-1 triangles per batch");
    if (ncol > (int64_t(1) << 27)) return fail(GPV_ERR_INVALID, "at most 2^27 xy columns per batch");
    if (vox >= (int64_t(1) << 32)) return fail(GPV_ERR_INVALID, "at most 2^32-1 Level1 voxels per batch");
    if (T > 0 && on_device && nm == 1 && !d[0].tris) return fail(GPV_ERR_INVALID, "NULL soup");
    HIP_TRY(hipSetDevice(c->device));
    const size_t hneed = hctr_at(nm) + sizeof(MeshCtr) * (size_t)nm + 64;
    if (na) {
        if (nm != 1 || !on_device) return fail(GPV_ERR_INVALID, "gpv_voxelize_noalloc takes one device soup");
        if (na->l1_words < W)
            return fail(GPV_ERR_INVALID, "Level1 planes hold %lld words; this grid needs %lld",
                        (long long)na->l1_words, (long long)W);
        if (hneed > c->h_cap) return fail(GPV_ERR_STATE, "the context's read-back buffer is too small");
    } else {
        TRY(ensure_hbuf(c, hneed));
    }

    Job* J = new Job();
    if (na) {
        J->arena = na->arena;
        J->arena_cap = na->arena_cap;
        J->borrowed_events = true;
        for (int k = 0; k < 16; ++k) J->ev[k] = c->na_ev[k];
        J->nb_cap = na->nb_cap;
        J->out_bidx = na->bidx;
        J->out_sat2 = na->l2b;
        J->out_par2 = na->l2i;
        // bin and hit buffers within a quarter and a sixteenth of the arena;
        // a larger count regrows (from what is left) like any other run
        pair_cap = std::min<uint64_t>(pair_cap, na->arena_cap / 4 / 24);
        hit_cap = std::min<uint64_t>(hit_cap, na->arena_cap / 16 / 8);
    }
    J->ctx = c;
    J->nm = nm;
    J->info = std::move(info);
    J->T = T;
    J->W = W;
    J->ncol = ncol;
    auto bail = [&](int r) {
        job_free(J);
        return r;
    };
    for (auto& e : J->ev) {
        if (J->borrowed_events) break;
        if (!c->events.empty()) {
            e = c->events.back();
            c->events.pop_back();
        } else if (hipEventCreate(&e) != hipSuccess) {
            e = nullptr;
            return bail(fail(GPV_ERR_HIP, "event creation failed"));
        }
    }

    // per-mesh grids, offsets and the concatenated center tables
    J->g.resize((size_t)nm);
    J->off.assign((size_t)nm + 1, MeshOff{});
    std::vector<size_t> tab_at((size_t)nm * 12);
    size_t ntab = 0;
    uint64_t word0 = 0;
    uint32_t col0 = 0, tri0 = 0, ptask0 = 0;
    for (int m = 0; m < nm; ++m) {
        GridDev& g = J->g[m];
        g = grid_dev(J->info[m]);
        g.normals = d[m].params.normals ? 1 : 0;
        g.l2cuda = d[m].params.cuda_level2 ? 1 : 0;
        g.l1st = d[m].params.l1_stencil ? 1 : 0;
        J->any_stencil = J->any_stencil || g.l1st;
        J->all_stencil = (m == 0 || J->all_stencil) && g.l1st;
        J->info[m].has_normals = g.normals;
        J->any_normals = J->any_normals || g.normals;
        J->n2max = std::max(J->n2max, g.n2);
        J->off[m].word0 = word0;
        J->off[m].col0 = col0;
        J->off[m].tri0 = tri0;
        J->off[m].ptask0 = ptask0;
        ptask0 += (uint32_t)l1_parity_tasks(g);
        word0 += (uint64_t)g.nz * g.ny * g.nxw;
        col0 += (uint32_t)(g.nx * g.ny);
        tri0 += (uint32_t)d[m].ntris;
        size_t o[12];
        const size_t n = table_layout(J->info[m], o);
        for (int k = 0; k < 12; ++k) tab_at[(size_t)m * 12 + k] = ntab + o[k];
        ntab += n;
    }
    // the tables are written straight into the context's page-locked
    // staging, so their upload is asynchronous (the previous job on this
    // context has finished); the no-allocation entry point, which may not
    // grow that staging, uses ordinary host memory instead
    float* tab_src = nullptr;
    if (J->arena) {
        J->h_tab.assign(ntab, 0.0f);
        tab_src = J->h_tab.data();
    } else {
        if (ntab * sizeof(float) > c->tab_cap) {
            const size_t want = std::max<size_t>(ntab * sizeof(float), 2 * c->tab_cap);
            if (c->tab_buf) (void)hipHostFree(c->tab_buf);
            c->tab_buf = nullptr;
            c->tab_cap = 0;
            if (hipHostMalloc(reinterpret_cast<void**>(&c->tab_buf), want, hipHostMallocDefault) != hipSuccess) {
                (void)hipGetLastError();
                c->tab_buf = nullptr;
                return bail(fail(GPV_ERR_OOM, "page-locked table staging of %zu bytes failed", want));
            }
            c->tab_cap = want;
        }
        tab_src = c->tab_buf;
    }
    for (int m = 0; m < nm; ++m) {
        GridDev& g = J->g[m];
        size_t o[12];
        table_layout(J->info[m], o);
        const size_t base = tab_at[(size_t)m * 12];   // the first table starts at offset 0 of the mesh's block
        g.r2mono = fill_tables(J->info[m], g.l2cuda != 0, tab_src + base, o) ? 1 : 0;
    }
    J->off[nm].word0 = word0;
    J->off[nm].col0 = col0;
    J->off[nm].tri0 = tri0;
    J->off[nm].ptask0 = ptask0;
    int r;
    if ((r = alloc(J, &J->tables, ntab)) != GPV_OK) return bail(r);
    if (hipMemcpyAsync(J->tables, tab_src, ntab * sizeof(float), hipMemcpyHostToDevice, c->stream) != hipSuccess)
        return bail(fail(GPV_ERR_HIP, "table upload failed"));
    for (int m = 0; m < nm; ++m) {
        const float* tb = J->tables;
        const size_t* o = &tab_at[(size_t)m * 12];
        GridDev& g = J->g[m];
        g.c1x = tb + o[0]; g.c1y = tb + o[1]; g.c1z = tb + o[2];
        g.r1x = tb + o[3]; g.r1y = tb + o[4]; g.r1z = tb + o[5];
        g.c2x = tb + o[6]; g.c2y = tb + o[7]; g.c2z = tb + o[8];
        g.r2x = tb + o[9]; g.r2y = tb + o[10]; g.r2z = tb + o[11];
    }

    // the soup; device soups that already lie back to back in mesh order (a
    // staged batch upload) are used in place (empty meshes may point anywhere;
    // the first non-empty one starts the soup, its first triangle being 0)
    const float* base = nullptr;
    bool contiguous = on_device != 0 && nm > 1;
    for (int m = 0; m < nm && contiguous; ++m) {
        if (d[m].ntris == 0) continue;
        if (!base) base = d[m].tris;
        else contiguous = d[m].tris == base + (size_t)J->off[m].tri0 * 9;
    }
    if (nm == 1 && on_device) {
        J->tris = d[0].tris;
    } else if (contiguous && base) {
        J->tris = base;
    } else if (T > 0 && on_device) {
        // device soups in one gather launch instead of a copy per mesh
        float* s = nullptr;
        SoupPiece* pc = nullptr;
        if ((r = alloc(J, &s, (size_t)T * 9)) != GPV_OK || (r = alloc(J, &pc, (size_t)nm)) != GPV_OK) return bail(r);
        std::vector<SoupPiece>& hp = J->h_pieces;
        hp.resize((size_t)nm);
        for (int m = 0; m < nm; ++m)
            hp[(size_t)m] = SoupPiece{d[m].tris, (uint64_t)J->off[m].tri0 * 9, (uint64_t)d[m].ntris * 9};
        if (hipMemcpyAsync(pc, hp.data(), sizeof(SoupPiece) * (size_t)nm, hipMemcpyHostToDevice, c->stream) !=
            hipSuccess)
            return bail(fail(GPV_ERR_HIP, "soup table upload failed"));
        launch_gather_soups(pc, nm, s, 64, false, c->stream);
        if (hipGetLastError() != hipSuccess) return bail(fail(GPV_ERR_HIP, "soup gather failed"));
        J->tris = s;
        J->own_soup = s;
        J->soup_pieces = pc;
    } else if (T > 0) {
        float* s = nullptr;
        if ((r = alloc(J, &s, (size_t)T * 9)) != GPV_OK) return bail(r);
        for (int m = 0; m < nm; ++m)
            if (d[m].ntris > 0 &&
                hipMemcpyAsync(s + (size_t)J->off[m].tri0 * 9, d[m].tris, (size_t)d[m].ntris * 9 * sizeof(float),
                               hipMemcpyHostToDevice, c->stream) != hipSuccess)
                return bail(fail(GPV_ERR_HIP, "soup copy of mesh %d failed", m));
        J->tris = s;
        J->own_soup = s;
    }

    // batch descriptors
    if ((r = alloc(J, &J->mctr, (size_t)nm)) != GPV_OK) return bail(r);
    if (nm > 1) {
        if ((r = alloc(J, &J->d_grids, (size_t)nm)) || (r = alloc(J, &J->d_off, (size_t)nm + 1)) ||
            (r = alloc(J, &J->tri_mesh, (size_t)std::max<int64_t>(T, 1))))
            return bail(r);
        if (hipMemcpyAsync(J->d_grids, J->g.data(), sizeof(GridDev) * (size_t)nm, hipMemcpyHostToDevice,
                           c->stream) != hipSuccess ||
            hipMemcpyAsync(J->d_off, J->off.data(), sizeof(MeshOff) * ((size_t)nm + 1), hipMemcpyHostToDevice,
                           c->stream) != hipSuccess)
            return bail(fail(GPV_ERR_HIP, "batch descriptor upload failed"));
    }

    const size_t Tn = (size_t)std::max<int64_t>(T, 1);
    if ((r = alloc(J, &J->box0, Tn)) || (r = alloc(J, &J->box1, Tn)) || (r = alloc(J, &J->rng, Tn)) ||
        (r = alloc(J, &J->slab1, Tn)) || (r = alloc(J, &J->slab2, Tn)) ||
        (r = alloc(J, &J->sat_items, Tn)) || (r = alloc(J, &J->ray_items, Tn)) ||
        (r = alloc(J, &J->sat_off, Tn + 1)) || (r = alloc(J, &J->ray_off, Tn + 1)) ||
        (r = alloc(J, &J->ctr, 1)) || (r = alloc(J, &J->wcnt, (size_t)W)) ||
        (r = alloc(J, &J->wprefix, (size_t)W + 1)))
        return bail(r);
    if (na) {
        J->bnd = na->l1b;
        J->ins = na->l1i;
    } else if ((r = alloc(J, &J->bnd, (size_t)W)) || (r = alloc(J, &J->ins, (size_t)W))) {
        return bail(r);
    }
    J->pair_cap = pair_cap;
    J->col_cap = J->pair_cap;
    J->hit_cap = std::max<uint64_t>(hit_cap, 1u << 16);
    if ((r = alloc(J, &J->pair_key, J->pair_cap)) || (r = alloc(J, &J->pair_tri, J->pair_cap)) ||
        (r = alloc(J, &J->col_key, J->col_cap)) || (r = alloc(J, &J->col_z, J->col_cap)) ||
        (r = alloc(J, &J->hit_key, J->hit_cap)))
        return bail(r);
    if (J->any_stencil) {
        // the stencil fill's scratch: one set for the stencil meshes in turn
        // (stream-ordered), sized for the largest; it needs no count
        int64_t tmax = 1;
        size_t words = 1;
        for (int m = 0; m < nm; ++m) {
            if (!J->g[m].l1st) continue;
            tmax = std::max<int64_t>(tmax, d[m].ntris);
            words = std::max(words, l1_stencil_words(J->g[m]));
        }
        if ((r = alloc(J, &J->st_cnt, (size_t)tmax)) || (r = alloc(J, &J->st_off, (size_t)tmax + 1)) ||
            (r = alloc(J, &J->st_scan, scan_tmp_elems(tmax))) || (r = alloc(J, &J->st_tog, words)))
            return bail(r);
    }
    *out = J;
    return GPV_OK;
}

// per-triangle setup shared by the SAT and the ray paths
int stage_setup(Job* J)
{
    if (J->setup_done) return GPV_OK;
    gpv_ctx* c = J->ctx;
    hipStream_t s = c->stream;
    HIP_TRY(hipEventRecord(J->ev[0], s));
    launch_fill_u64(reinterpret_cast<uint64_t*>(J->mctr), 0, (int64_t)(sizeof(MeshCtr) / 8) * J->nm, s);
    launch_fill_u64(reinterpret_cast<uint64_t*>(&J->ctr->bad_input), 0, 1, s);
    if (J->nm > 1) launch_mesh_ids(J->d_off, J->nm, J->T, J->tri_mesh, s);
    launch_tri_setup(J->tris, J->T, J->g[0], J->view(), J->box0, J->box1, J->sat_items, J->rng, J->ray_items,
                     J->slab1, J->slab2, 3, &J->ctr->bad_input, s);
    TRY(ensure_scan_tmp(J, J->T));
    launch_scan_u32_u64(J->sat_items, J->sat_off, J->T, J->scan64, s);
    launch_scan_u32_u64(J->ray_items, J->ray_off, J->T, J->scan64, s);
    HIP_TRY(hipGetLastError());
    J->setup_done = true;
    return GPV_OK;
}

// The set-up in two halves: the SAT half on the main stream, the ray half
// (and, after it, the ray expansion) on the side stream, forked at ev[15].
int stage_setup_split(Job* J)
{
    gpv_ctx* c = J->ctx;
    hipStream_t s = c->stream;
    HIP_TRY(hipEventRecord(J->ev[0], s));
    launch_fill_u64(reinterpret_cast<uint64_t*>(J->mctr), 0, (int64_t)(sizeof(MeshCtr) / 8) * J->nm, s);
    launch_fill_u64(reinterpret_cast<uint64_t*>(&J->ctr->bad_input), 0, 1, s);
    if (J->nm > 1) launch_mesh_ids(J->d_off, J->nm, J->T, J->tri_mesh, s);
    TRY(ensure_scan_tmp(J, J->T));
    release(J, J->scan_ray);
    TRY(alloc(J, &J->scan_ray, scan_tmp_elems(J->T)));
    HIP_TRY(hipEventRecord(J->ev[15], s));
    launch_tri_setup(J->tris, J->T, J->g[0], J->view(), J->box0, J->box1, J->sat_items, J->rng, J->ray_items,
                     J->slab1, J->slab2, 1, &J->ctr->bad_input, s);
    launch_scan_u32_u64(J->sat_items, J->sat_off, J->T, J->scan64, s);
    HIP_TRY(hipStreamWaitEvent(c->side, J->ev[15], 0));
    if (!J->all_stencil) {   // the ray half feeds only the ray parity
        launch_tri_setup(J->tris, J->T, J->g[0], J->view(), J->box0, J->box1, J->sat_items, J->rng, J->ray_items,
                         J->slab1, J->slab2, 2, &J->ctr->bad_input, c->side);
        launch_scan_u32_u64(J->ray_items, J->ray_off, J->T, J->scan_ray, c->side);
    }
    HIP_TRY(hipGetLastError());
    J->setup_done = true;
    return GPV_OK;
}

// Level1 SAT sweep + the boundary popcount scan (ranks need the final bits)
int launch_sat_half(Job* J)
{
    gpv_ctx* c = J->ctx;
    hipStream_t s = c->stream;
    HIP_TRY(hipEventRecord(J->ev[1], s));
    // boundary flags as bytes (plain stores from the SAT), packed to bits after
    uint8_t* bbytes = nullptr;
    TRY(alloc(J, &bbytes, (size_t)J->W * 64));
    launch_fill_u64(reinterpret_cast<uint64_t*>(bbytes), 0, J->W * 8, s);
    launch_fill_u64(reinterpret_cast<uint64_t*>(&J->ctr->pairs), 0, 2, s);   // pairs, colents
    Append64 pa{&J->ctr->pairs, J->pair_key, J->pair_tri, J->pair_cap};
    Append64 co{&J->ctr->colents, J->col_key, J->col_z, J->col_cap};
    J->pairs_ranked = false;   // the sweep appends voxel bits
    launch_sat_expand(J->tris, J->T, J->box0, J->box1, J->slab1, J->sat_off, J->g[0], J->view(), bbytes, pa, co,
                      expand_blocks(c), c->tune.expand_stage_tris, s);
    launch_pack_bytes(bbytes, J->W, J->bnd, s);
    release(J, bbytes);
    launch_popc_words(J->bnd, J->W, J->wcnt, s);
    TRY(ensure_scan_tmp(J, J->W));
    launch_scan_u32_u32(J->wcnt, J->wprefix, J->W, J->scan32, nullptr, s);
    if (J->nm > 1) launch_mesh_bounds(J->wprefix, J->view(), s);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(J->ev[12], s));
    return GPV_OK;
}

int launch_stencil(Job* J, hipStream_t s);

// Level1 ray expansion on stream s (allocates nothing)
int launch_ray_half(Job* J, hipStream_t s)
{
    gpv_ctx* c = J->ctx;
    HIP_TRY(hipEventRecord(J->ev[2], s));
    if (J->all_stencil) {
        // no ray parity, hence no hits (the counter is cleared on the stream
        // of the count read-back): the stencil fill takes the ray half's place
        launch_fill_u64(reinterpret_cast<uint64_t*>(&J->ctr->hits), 0, 1, J->ctx->stream);
        TRY(launch_stencil(J, s));
        HIP_TRY(hipEventRecord(J->ev[13], s));
        return GPV_OK;
    }
    launch_fill_u64(reinterpret_cast<uint64_t*>(&J->ctr->hits), 0, 1, s);
    Append64 hi{&J->ctr->hits, J->hit_key, nullptr, J->hit_cap};
    launch_ray_expand(J->tris, J->T, J->rng, J->ray_off, J->g[0], J->view(), hi, expand_blocks(c),
                      c->tune.expand_stage_tris, s);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(J->ev[13], s));
    return GPV_OK;
}

int launch_ray_half(Job* J) { return launch_ray_half(J, J->ctx->stream); }

// one round trip: the counters and, when asked, the boundary offsets
int read_counts(Job* J, bool with_b)
{
    gpv_ctx* c = J->ctx;
    hipStream_t s = c->stream;
    HIP_TRY(hipMemcpyAsync(c->h_buf, J->ctr, sizeof(Counters), hipMemcpyDeviceToHost, s));
    if (with_b) {
        if (J->nm > 1)
            HIP_TRY(hipMemcpyAsync(c->h_buf + kHOff, J->d_off, sizeof(MeshOff) * ((size_t)J->nm + 1),
                                   hipMemcpyDeviceToHost, s));
        else
            HIP_TRY(hipMemcpyAsync(c->h_buf + kHOff, J->wprefix + J->W, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    }
    HIP_TRY(hipStreamSynchronize(s));
    // the set-up kernel flags coordinates outside the input domain (NaN and
    // infinities would reach the index AABBs and the pruning bounds, where
    // no float expression of the reference decides the outcome)
    if (reinterpret_cast<const Counters*>(c->h_buf)->bad_input)
        return fail(GPV_ERR_INVALID, "a vertex coordinate is not finite or exceeds 2^40 in magnitude");
    return GPV_OK;
}

const Counters& hctr(const Job* J) { return *reinterpret_cast<const Counters*>(J->ctx->h_buf); }

// after read_counts: accept the sweep or regrow the bin buffers to the exact
// count (the reference's retry, src/Object.cpp:3198-3214)
int complete_sat(Job* J, int64_t* needed)
{
    gpv_ctx* c = J->ctx;
    const uint64_t np = hctr(J).pairs, nc = hctr(J).colents;
    *needed = 0;
    if (np > J->pair_cap || nc > J->col_cap) {
        const uint64_t need = std::max(np, nc);
        release(J, J->pair_key);
        release(J, J->pair_tri);
        release(J, J->col_key);
        release(J, J->col_z);
        J->pair_cap = need + need / 8 + 1024;
        J->col_cap = J->pair_cap;
        TRY(alloc(J, &J->pair_key, J->pair_cap));
        TRY(alloc(J, &J->pair_tri, J->pair_cap));
        TRY(alloc(J, &J->col_key, J->col_cap));
        TRY(alloc(J, &J->col_z, J->col_cap));
        J->regrows++;
        *needed = (int64_t)need;
        return GPV_OK;
    }
    if (np >= (uint64_t(1) << 32) || nc >= (uint64_t(1) << 32))
        return fail(GPV_ERR_CAPACITY, "%llu bin entries exceed the 2^32-1 a run can index; split the batch",
                    (unsigned long long)std::max(np, nc));
    J->n_pairs = np;
    J->n_cols = nc;
    if (J->nm > 1) {
        memcpy(J->off.data(), c->h_buf + kHOff, sizeof(MeshOff) * ((size_t)J->nm + 1));
    } else {
        const uint32_t B = *reinterpret_cast<const uint32_t*>(c->h_buf + kHOff);
        J->off[0].b0 = 0;
        J->off[0].l2w0 = 0;
        J->off[1].b0 = B;
        J->off[1].l2w0 = (uint64_t)B * (uint64_t)J->g[0].w2;
        J->off[0].l2n0 = 0;
        J->off[1].l2n0 = J->g[0].normals ? (uint64_t)B * J->g[0].n2 * J->g[0].n2 * J->g[0].n2 : 0;
    }
    J->B = J->off[J->nm].b0;
    if (J->B > (int64_t(1) << 27))
        return fail(GPV_ERR_INVALID, "%lld boundary voxels: at most 2^27 per batch (split the batch)",
                    (long long)J->B);
    J->W2 = (int64_t)J->off[J->nm].l2w0;
    J->NL2 = (int64_t)J->off[J->nm].l2n0;
    J->sat_done = true;
    return GPV_OK;
}

// after read_counts: false (and a regrown buffer) when the hits overflowed
int complete_ray(Job* J, bool* ok)
{
    const uint64_t nh = hctr(J).hits;
    *ok = nh <= J->hit_cap && nh < (uint64_t(1) << 32);
    if (*ok) {
        J->n_hits = nh;
        return GPV_OK;
    }
    if (nh >= (uint64_t(1) << 32)) return fail(GPV_ERR_CAPACITY, "more than 2^32-1 ray hits; split the batch");
    release(J, J->hit_key);
    J->hit_cap = nh + nh / 8 + 1024;
    return alloc(J, &J->hit_key, J->hit_cap);
}

// hits grouped by column, then the parity of every Level1 voxel
// buffers of the parity fill (allocated before it may fork onto the side stream)
int alloc_parity(Job* J)
{
    release(J, J->hcnt);
    release(J, J->hoff);
    release(J, J->hrec);
    release(J, J->scan_side);
    TRY(alloc(J, &J->hcnt, (size_t)J->ncol));
    TRY(alloc(J, &J->hoff, (size_t)J->ncol + 1));
    TRY(alloc(J, &J->hrec, (size_t)std::max<uint64_t>(J->n_hits, 1) * 3));
    return alloc(J, &J->scan_side, scan_tmp_elems(J->ncol));
}

// the GL-stencil-equivalent Level1 fill of every mesh that asks for it, on
// stream s: it reads only the soup and the grid, so the fused pipeline runs
// it beside the Level1 SAT; it writes the inside bits of those meshes,
// which the ray-parity fill leaves alone
int launch_stencil(Job* J, hipStream_t s)
{
    if (!J->any_stencil || J->stencil_done) return GPV_OK;
    for (int m = 0; m < J->nm; ++m)
        if (J->g[m].l1st)
            launch_l1_stencil(J->tris + (size_t)J->off[m].tri0 * 9, (int64_t)J->off[m + 1].tri0 - J->off[m].tri0,
                              J->g[m], J->st_cnt, J->st_off, J->st_scan, J->st_tog, J->ins + J->off[m].word0, s);
    HIP_TRY(hipGetLastError());
    J->stencil_done = true;
    return GPV_OK;
}

// buffers of the column-entry sort when it runs on the side stream
// (allocated before the fork, released after the Level2 join)
int alloc_cols(Job* J)
{
    release(J, J->coff);
    release(J, J->crec);
    release(J, J->cs_cnt);
    release(J, J->cs_bucket);
    TRY(alloc(J, &J->coff, (size_t)J->ncol + 1));
    TRY(alloc(J, &J->crec, (size_t)std::max<uint64_t>(J->n_cols, 1) * 3));
    TRY(alloc(J, &J->cs_cnt, 3 * ts_tiles((uint64_t)J->ncol) + 2));
    return alloc(J, &J->cs_bucket, (size_t)std::max<uint64_t>(J->n_cols, 1));
}

// column entries by column with their MT records, on stream s; scan is a
// scan temporary of at least scan_tmp_elems(ncol) that no other stream uses
int col_sort_kernels(Job* J, hipStream_t s, uint32_t* scan)
{
    const size_t tiles = ts_tiles((uint64_t)J->ncol);
    const TileSortTmp ts{J->cs_cnt, J->cs_cnt + 2 * tiles + 1, scan, J->cs_bucket};
    const int ipb = std::max(256, J->ctx->tune.sort_items_per_block);
    launch_bin_cols(J->tris, J->col_key, J->col_z, J->n_cols, (uint32_t)J->ncol, ts, ipb, J->coff, J->crec, s);
    HIP_TRY(hipGetLastError());
    return GPV_OK;
}

// the parity fill on stream s: hits by column, their MT records, the fill
int parity_kernels(Job* J, hipStream_t s)
{
    if (!J->all_stencil) {   // the ray parity skips the stencil meshes
        launch_fill_u32(J->hcnt, 0, J->ncol, s);
        launch_count_keys_hi(J->hit_key, J->n_hits, J->hcnt, s);
        launch_scan_u32_u32(J->hcnt, J->hoff, J->ncol, J->scan_side, nullptr, s);
        launch_scatter_hits(J->tris, J->hit_key, J->n_hits, J->g[0], J->view(), J->hoff, J->hcnt, J->hrec, s);
        launch_l1_parity(J->hrec, J->hoff, J->g[0], J->view(), (int64_t)J->off[J->nm].ptask0, J->ins, s);
    }
    TRY(launch_stencil(J, s));   // already launched beside the Level1 SAT in the fused pipeline
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(J->ev[14], s));
    return GPV_OK;
}

int launch_parity(Job* J)
{
    hipStream_t s = J->ctx->stream;
    HIP_TRY(hipEventRecord(J->ev[3], s));
    TRY(alloc_parity(J));
    TRY(parity_kernels(J, s));
    J->inout_done = true;
    return GPV_OK;
}

// boundary index list, bins per boundary voxel, column entries with records
int launch_compact(Job* J)
{
    hipStream_t s = J->ctx->stream;
    HIP_TRY(hipEventRecord(J->ev[4], s));
    const size_t Bn = (size_t)std::max<int64_t>(J->B, 1);
    release(J, J->bidx);
    release(J, J->bmesh);
    if (J->out_bidx) {
        if (J->B > J->nb_cap)
            return fail(GPV_ERR_TOO_LARGE, "%lld boundary voxels; the result buffers hold %lld",
                        (long long)J->B, (long long)J->nb_cap);
        J->bidx = J->out_bidx;
    } else {
        TRY(alloc(J, &J->bidx, Bn));
    }
    if (J->nm > 1) TRY(alloc(J, &J->bmesh, Bn));
    launch_boundary_index(J->bnd, J->wprefix, J->g[0], J->view(), J->W, J->bidx, s);
    release(J, J->boff);
    release(J, J->list);
    release(J, J->pmask);
    TRY(alloc(J, &J->boff, (size_t)J->B + 1));
    TRY(alloc(J, &J->list, (size_t)std::max<uint64_t>(J->n_pairs, 1)));
    if (J->n2max > 0) TRY(alloc(J, &J->pmask, (size_t)std::max<uint64_t>(J->n_pairs, 1)));
    // one scratch for both counting sorts (bins by boundary rank, column
    // entries by column)
    const size_t tiles = J->cols_side ? ts_tiles((uint64_t)J->B)
                                      : std::max(ts_tiles((uint64_t)J->B), ts_tiles((uint64_t)J->ncol));
    TileSortTmp ts{};
    uint32_t* ts_cnt = nullptr;
    uint4* ts_bucket = nullptr;
    TRY(alloc(J, &ts_cnt, 3 * tiles + 2));
    TRY(alloc(J, &ts_bucket,
              (size_t)std::max<uint64_t>(J->cols_side ? J->n_pairs : std::max(J->n_pairs, J->n_cols), 1)));
    TRY(ensure_scan_tmp(J, (int64_t)tiles));
    ts.total = ts_cnt;
    ts.off = ts_cnt + 2 * tiles + 1;
    ts.scan = J->scan32;
    ts.bucket = ts_bucket;
    const int ipb = std::max(256, J->ctx->tune.sort_items_per_block);
    launch_bin_pairs(J->pair_key, J->pairs_ranked, J->pair_tri, J->n_pairs, J->bnd, J->wprefix, (uint32_t)J->B, ts,
                     ipb, J->boff, J->nm > 1 ? nullptr : &J->mctr[0].max_per_voxel, J->list, J->tris, J->bidx,
                     J->g[0], J->view(), J->pmask, s);
    J->pairs_ranked = true;
    if (J->nm > 1) launch_mesh_bins(J->boff, J->B, J->view(), s);
    if (!J->cols_side) {
        release(J, J->coff);
        release(J, J->crec);
        TRY(alloc(J, &J->coff, (size_t)J->ncol + 1));
        TRY(alloc(J, &J->crec, (size_t)std::max<uint64_t>(J->n_cols, 1) * 3));
        launch_bin_cols(J->tris, J->col_key, J->col_z, J->n_cols, (uint32_t)J->ncol, ts, ipb, J->coff, J->crec, s);
    }
    release(J, ts_cnt);
    release(J, ts_bucket);
    HIP_TRY(hipGetLastError());
    if (J->n2max > 0) {
        release(J, J->sat2);
        release(J, J->par2);
        if (J->out_bidx) {   // W2 = B * w2 fits: B <= nb_cap was checked above
            J->sat2 = J->out_sat2;
            J->par2 = J->out_par2;
        } else {
            TRY(alloc(J, &J->sat2, (size_t)std::max<int64_t>(J->W2, 1)));
            TRY(alloc(J, &J->par2, (size_t)std::max<int64_t>(J->W2, 1)));
        }
    }
    J->compact_done = true;
    return GPV_OK;
}

// Level2 parity waves of the job; in a batch each mesh's first wave goes to
// MeshOff::pw0 on the device (main stream, before any parity launch)
int l2_parity_plan(Job* J)
{
    int64_t w = 0;
    for (int m = 0; m < J->nm; ++m) {
        J->off[m].pw0 = (uint32_t)w;
        w += l2_parity_waves(J->g[m].n2, J->nb(m));
    }
    J->off[J->nm].pw0 = (uint32_t)w;
    if (w >= (int64_t(1) << 32)) return fail(GPV_ERR_INVALID, "too many Level2 parity waves");
    J->l2p_waves = w;
    if (J->nm > 1) {
        hipStream_t s = J->ctx->stream;
        // the MeshOff staging region was last filled by the size read-back,
        // which the host has already waited for; a later job refills it only
        // after its own read-back, which this stream orders after this copy
        MeshOff* h = reinterpret_cast<MeshOff*>(J->ctx->h_buf + kHOff);
        memcpy(h, J->off.data(), sizeof(MeshOff) * ((size_t)J->nm + 1));
        HIP_TRY(hipMemcpyAsync(J->d_off, h, sizeof(MeshOff) * ((size_t)J->nm + 1), hipMemcpyHostToDevice, s));
    }
    return GPV_OK;
}

// cpass words per column entry: the pass bits of its sub-columns and a header
int cpass_stride(const Job* J) { return (J->n2max * J->n2max + 63) / 64 + 1; }

// Level2 parity kernels on stream s (cpass allocated, waves planned)
int l2_parity_kernels(Job* J, hipStream_t s)
{
    const int stride = cpass_stride(J);
    const gpv_tune& tu = J->ctx->tune;
    launch_col_pass(J->crec, J->g[0], J->view(), J->n_cols, stride, tu.l2_parity_waves, J->cpass, s);
    launch_l2_parity(J->bidx, J->B, J->l2p_waves, J->coff, J->crec, J->cpass, stride, J->n2max, tu.l2_parity_waves,
                     J->g[0], J->view(), J->par2, s);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(J->ev[11], s));
    return GPV_OK;
}

int alloc_cpass(Job* J)
{
    const int stride = cpass_stride(J);
    release(J, J->cpass);
    return alloc(J, &J->cpass, (size_t)std::max<uint64_t>(J->n_cols, 1) * stride);
}

int run_l2_parity(Job* J)
{
    if (J->n2max <= 0) return GPV_OK;
    hipStream_t s = J->ctx->stream;
    HIP_TRY(hipEventRecord(J->ev[5], s));
    TRY(alloc_cpass(J));
    TRY(l2_parity_plan(J));
    TRY(l2_parity_kernels(J, s));
    release(J, J->cpass);
    J->l2p_done = true;
    return GPV_OK;
}

// group arrays of the Level2 SAT (allocated outside any two-stream region)
int alloc_l2_groups(Job* J)
{
    const size_t n = (size_t)std::max<int64_t>(J->B, 1) + 1;
    release(J, J->gflag);
    release(J, J->gidx);
    release(J, J->gstart);
    TRY(alloc(J, &J->gflag, n));
    TRY(alloc(J, &J->gidx, n));
    TRY(alloc(J, &J->gstart, n));
    return ensure_scan_tmp(J, J->B);
}

void release_l2_groups(Job* J)
{
    release(J, J->gflag);
    release(J, J->gidx);
    release(J, J->gstart);
}

// whether the Level2 SAT shares waves among voxels: few bin entries per voxel
bool l2_grouped(const Job* J)
{
    if (J->B <= 0 || (int64_t)J->n_pairs >= (int64_t)J->ctx->tune.l2_group_below * J->B) return false;
    // a group of one voxel (n2 = 16: its 64 words fill the wave's LDS) is
    // the per-voxel kernel at a lower occupancy
    for (int m = 0; m < J->nm; ++m)
        if (J->g[m].n2 > 0 && l2_group_voxels(J->g[m].w2, J->g[m].n2) > 1) return true;
    return false;
}

// group formation of the Level2 SAT (main stream, group arrays allocated)
void l2_sat_groups(Job* J)
{
    if (l2_grouped(J))
        launch_l2_groups(J->boff, J->B, J->g[0], J->view(), J->gflag, J->gidx, J->scan32, J->gstart,
                         J->ctx->stream);
}

// the Level2 SAT kernel on the main stream (groups formed)
int l2_sat_kernels(Job* J)
{
    hipStream_t s = J->ctx->stream;
    HIP_TRY(hipEventRecord(J->ev[6], s));
    // voxels with few bin entries share waves (k_l2_sat_grp); with many one
    // voxel per wave is faster
    const bool grouped = l2_grouped(J);
    // one wave per group, for an upper bound on the group count computed
    // here (a group starts at a mesh start, at a new 64-entry chunk of first
    // bin entries, or after a full group), so the grid is sized without
    // reading the count back; waves past the real count exit
    int64_t gmax = J->B;
    if (grouped) {
        gmax = (int64_t)(J->n_pairs / 64) + 1;
        for (int m = 0; m < J->nm; ++m) {
            const int64_t nb = J->nb(m);
            const int w2 = J->g[m].w2;
            const int64_t cap = l2_group_voxels(w2, J->g[m].n2);
            gmax += (nb + cap - 1) / cap + 1;
        }
        gmax = std::min<int64_t>(gmax, J->B);
    }
    launch_l2_sat(J->tris, J->bidx, J->B, J->boff, J->list, J->pmask, J->slab2, grouped ? J->gstart : nullptr,
                  J->gidx + J->B, gmax, J->ctx->tune.l2_sat_waves, J->g[0], J->view(), J->sat2, s);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(J->ev[7], s));
    return GPV_OK;
}

int run_l2_sat(Job* J)
{
    if (J->n2max <= 0) return GPV_OK;
    TRY(alloc_l2_groups(J));
    l2_sat_groups(J);
    TRY(l2_sat_kernels(J));
    release_l2_groups(J);
    J->l2s_done = true;
    return GPV_OK;
}

// Both Level2 stages of the fused pipeline: the parity on the side stream
// beside the SAT on the main one (they read the same inputs and write
// different planes).  Nothing is allocated or released between the fork and
// the join, so the pool's blocks stay in one stream order.
int run_level2(Job* J)
{
    if (J->n2max <= 0) return GPV_OK;
    gpv_ctx* c = J->ctx;
    hipStream_t s = c->stream;
    TRY(alloc_cpass(J));
    TRY(alloc_l2_groups(J));
    TRY(l2_parity_plan(J));
    // the SAT's small group-formation kernels go first: launched beside the
    // parity they would wait for its waves to drain
    l2_sat_groups(J);
    HIP_TRY(hipEventRecord(J->ev[5], s));
    HIP_TRY(hipStreamWaitEvent(c->side, J->ev[5], 0));
    TRY(l2_parity_kernels(J, c->side));
    TRY(l2_sat_kernels(J));
    HIP_TRY(hipStreamWaitEvent(s, J->ev[11], 0));
    release(J, J->cpass);
    release(J, J->cs_cnt);
    release(J, J->cs_bucket);
    release_l2_groups(J);
    J->l2p_done = true;
    J->l2s_done = true;
    return GPV_OK;
}

// Level1 and Level2 normals of the meshes that ask for them
int run_normals(Job* J)
{
    if (!J->any_normals) return GPV_OK;
    hipStream_t s = J->ctx->stream;
    HIP_TRY(hipEventRecord(J->ev[9], s));
    release(J, J->fnorm);
    release(J, J->slist);
    release(J, J->l1n);
    release(J, J->l2n);
    TRY(alloc(J, &J->fnorm, (size_t)std::max<int64_t>(J->T, 1) * 3));
    TRY(alloc(J, &J->slist, (size_t)std::max<uint64_t>(J->n_pairs, 1)));
    TRY(alloc(J, &J->l1n, (size_t)std::max<int64_t>(J->B, 1) * 3));
    TRY(alloc(J, &J->l2n, (size_t)std::max<int64_t>(J->NL2, 1) * 3));
    launch_face_normals(J->tris, J->T, J->fnorm, s);
    launch_sort_bins(J->list, J->boff, J->B, J->g[0], J->view(), J->slist, s);
    launch_l1_normals(J->slist, J->boff, J->B, J->fnorm, J->g[0], J->view(), J->l1n, s);
    if (J->NL2 > 0)
        launch_l2_normals(J->tris, J->bidx, J->B, J->boff, J->slist, J->fnorm, J->g[0], J->view(), J->l2n, s);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(J->ev[10], s));
    return GPV_OK;
}

// counts, per-mesh info, and the release of everything but the results
int job_finish(Job* J)
{
    gpv_ctx* c = J->ctx;
    hipStream_t s = c->stream;
    launch_count_states(J->bnd, J->ins, J->W, J->sat2, J->par2, J->W2, J->view(), s);
    HIP_TRY(hipEventRecord(J->ev[8], s));
    const size_t at = hctr_at(J->nm);
    HIP_TRY(hipMemcpyAsync(c->h_buf + at, J->mctr, sizeof(MeshCtr) * (size_t)J->nm, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const MeshCtr* mc = reinterpret_cast<const MeshCtr*>(c->h_buf + at);
    auto el = [&](int a, int b) -> float {
        float v = 0.0f;
        if (hipEventElapsedTime(&v, J->ev[a], J->ev[b]) != hipSuccess) {
            (void)hipGetLastError();
            v = 0.0f;
        }
        return v;
    };
    // setup, L1 SAT (+ boundary scan), L1 parity (ray expansion + parity
    // fill), compaction + bins, L2 parity, L2 SAT, each on its own stream;
    // stages on the side stream overlap the main one's.  For a batch these
    // are the batch's stage times
    float ms[6] = {el(0, 1), el(1, 12), el(2, 13) + el(3, 14), 0.0f, 0.0f, 0.0f};
    if (J->n2max > 0) {
        ms[3] = el(4, 5);
        ms[4] = el(5, 11);   // in the fused pipeline the two Level2 stages overlap
        ms[5] = el(6, 7);
    } else {
        ms[3] = el(4, 8);
    }
    const float total = el(0, 8);
    const float ms_nrm = J->any_normals ? el(9, 10) : 0.0f;
    for (int m = 0; m < J->nm; ++m) {
        gpv_grid_info& I = J->info[m];
        const int64_t n23 = (int64_t)I.n2 * I.n2 * I.n2;
        I.nb = J->nb(m);
        I.l1_boundary = I.nb;
        I.l1_inside = (int64_t)mc[m].l1_inside;
        I.l2_inside = (int64_t)mc[m].l2_inside;
        I.l2_boundary = (int64_t)mc[m].l2_boundary;
        I.tests_l1 = (int64_t)mc[m].tests_l1;
        I.pairs = J->nm > 1 ? (int64_t)mc[m].pairs : (int64_t)J->n_pairs;
        I.tests_l2 = I.n2 > 0 ? I.pairs * n23 : 0;
        I.max_tris_per_voxel = (int32_t)mc[m].max_per_voxel;
        I.regrows = J->regrows;
        for (int k = 0; k < 6; ++k) I.ms_stage[k] = ms[k];
        I.ms_total = total;
        I.ms_normals = ms_nrm;
    }
    // only the result planes stay resident
    release(J, J->box0); release(J, J->box1); release(J, J->rng); release(J, J->slab1); release(J, J->slab2);
    release(J, J->pmask);
    release(J, J->sat_items); release(J, J->ray_items); release(J, J->sat_off); release(J, J->ray_off);
    release(J, J->scan64); J->scan32 = nullptr; J->scan_cap = 0; release(J, J->scan_ray);
    release(J, J->pair_key); release(J, J->pair_tri); release(J, J->col_key); release(J, J->col_z);
    release(J, J->hit_key); release(J, J->wcnt); release(J, J->boff);
    release(J, J->list); release(J, J->coff); release(J, J->crec); release(J, J->hcnt);
    release(J, J->cs_cnt); release(J, J->cs_bucket);
    release(J, J->hoff); release(J, J->hrec); release(J, J->scan_side); release(J, J->tri_mesh);
    release(J, J->st_cnt); release(J, J->st_off); release(J, J->st_scan); release(J, J->st_tog);
    release(J, J->bmesh);
    release(J, J->d_grids); release(J, J->d_off); release(J, J->own_soup); release(J, J->mctr);
    release(J, J->soup_pieces);
    release(J, J->fnorm); release(J, J->slist);
    J->tris = nullptr;
    J->finished = true;
    return GPV_OK;
}

// the fused pipeline: every stage back to back, one size read-back
int job_run(Job* J)
{
    gpv_ctx* c = J->ctx;
    // independent stages go beside each other on the side stream: the ray
    // expansion beside the Level1 SAT, the parity fill and the column-entry
    // sort beside the bins (and, in run_level2, the Level2 parity beside the
    // Level2 SAT).  Side-stream work allocates nothing and its buffers are
    // allocated before the fork and live until the join, so pool blocks keep
    // one stream order.
    const bool side = c->tune.l2_overlap != 0;
    if (side) {
        TRY(stage_setup_split(J));
        TRY(launch_ray_half(J, c->side));
        TRY(launch_sat_half(J));
        // the count read-back waits for the ray expansion's hits; in a batch
        // that mixes the two fills the stencil follows it on the side stream,
        // after that point (with only stencil meshes it has run in its place)
        if (!J->all_stencil) HIP_TRY(hipStreamWaitEvent(c->stream, J->ev[13], 0));
        TRY(launch_stencil(J, c->side));
    } else {
        TRY(stage_setup(J));
        TRY(launch_sat_half(J));
        TRY(launch_ray_half(J));
    }
    TRY(read_counts(J, true));
    int64_t needed = 0;
    bool ray_ok = false;
    TRY(complete_sat(J, &needed));
    TRY(complete_ray(J, &ray_ok));
    for (int attempt = 0; attempt < 2 && (needed > 0 || !ray_ok); ++attempt) {
        if (needed > 0) TRY(launch_sat_half(J));
        if (!ray_ok) TRY(launch_ray_half(J));
        TRY(read_counts(J, true));
        if (needed > 0) TRY(complete_sat(J, &needed));
        if (!ray_ok) TRY(complete_ray(J, &ray_ok));
    }
    if (needed > 0 || !ray_ok) return fail(GPV_ERR_CAPACITY, "buffer capacity did not converge");
    if (side) {
        TRY(alloc_parity(J));
        // the column entries (read only by the Level2 parity, which runs on
        // the side stream too) sort there after the fill, beside the bins
        J->cols_side = J->n2max > 0;
        if (J->cols_side) TRY(alloc_cols(J));
        // the main stream's compaction (the critical path) is enqueued first;
        // the side stream waits for the point before it
        HIP_TRY(hipEventRecord(J->ev[3], c->stream));
        TRY(launch_compact(J));
        HIP_TRY(hipStreamWaitEvent(c->side, J->ev[3], 0));
        TRY(parity_kernels(J, c->side));
        if (J->cols_side) TRY(col_sort_kernels(J, c->side, J->scan_side));
        J->inout_done = true;
        TRY(run_level2(J));
        HIP_TRY(hipStreamWaitEvent(c->stream, J->ev[14], 0));
    } else {
        TRY(launch_parity(J));
        TRY(launch_compact(J));
        TRY(run_l2_parity(J));
        TRY(run_l2_sat(J));
    }
    TRY(run_normals(J));
    return job_finish(J);
}

// Upper bound of the bytes a one-mesh run requests from the pool (every
// buffer it allocates, as if none were returned before the run ends), with
// the run's own capacities for the bin and hit buffers and an estimate of
// the boundary voxels: twice the voxels on the faces of the grid's box, to
// allow for folds.  Exceeded only when the bins regrow or the surface holds
// more boundary voxels than that.
int64_t predict_run_bytes(const gpv_mesh_desc& d, const gpv_grid_info& I)
{
    const int64_t nx = I.n[0], ny = I.n[1], nz = I.n[2];
    const int64_t W = nz * ny * ((nx + 63) / 64);
    const int64_t n1 = nx * ny * nz, ncol = nx * ny;
    const int64_t n2 = I.n2 > 0 ? I.n2 : 0;
    const int64_t w2 = n2 > 0 ? (n2 * n2 * n2 + 63) / 64 : 0;
    const int64_t T = d.ntris;
    const int64_t nmax = std::max(nx, std::max(ny, nz));
    const int64_t P = d.params.initial_capacity > 0 ? d.params.initial_capacity
                                                    : std::max<int64_t>(int64_t(1) << 16, 16 * T + 8 * nmax * nmax);
    const int64_t H = std::max<int64_t>(8 * T + 4 * ncol, int64_t(1) << 16);
    const int64_t B = std::min(n1, 4 * (nx * ny + ny * nz + nz * nx));
    auto scan = [](int64_t n) { return 8 * (int64_t)scan_tmp_elems(n); };
    int64_t b = 0;
    b += 4 * (2 * (nx + ny + nz) * (1 + n2) + 1);               // center tables
    b += 36 * T + 4 * T + 4 * B + 4096;                          // soup copy, batch mesh ids, descriptors
    b += 120 * T + 16 + 2 * scan(std::max(T, W));               // per-triangle set-up, item scans
    b += 24 * W + 64 * W;                                        // planes, word counts and prefix, flag bytes
    b += 24 * P + 8 * H;                                         // bin pairs, column entries, ray hits
    b += 8 * ncol + 8 + 48 * H + scan(ncol);                     // parity fill: hits by column, records
    b += 4 * (ncol + 1) + 48 * P + 16 * P + 64 * (ncol / 64 + 1);   // column entries by column, records
    b += 8 * B + 4 + 4 * P + 8 * P + 16 * P + 64 * (B / 64 + 1) + scan(B);   // boundary list, bins, masks
    b += 16 * B * w2;                                            // Level2 planes
    b += n2 > 0 ? 8 * P * ((n2 * n2 + 63) / 64 + 1) + 12 * (B + 1) : 0;   // Level2 parity bits, SAT groups
    if (d.params.normals) b += 12 * T + 4 * P + 12 * B + 12 * B * n2 * n2 * n2;
    if (d.params.l1_stencil) b += 12 * T + 8 + scan(T) + 8 * nz * 2 * ny * ((2 * nx + 63) / 64);   // stencil fill
    return b;
}

gpv_grid* make_view(Job* J, int m)
{
    gpv_grid* G = new gpv_grid();
    G->ctx = J->ctx;
    G->job = J;
    G->m = m;
    J->refs++;
    return G;
}

int check_grid(gpv_ctx* c, const gpv_grid* G)
{
    if (!c || !G || G->ctx != c || !G->job) return fail(GPV_ERR_INVALID, "bad context/grid");
    return GPV_OK;
}

// staged entry points drive single-mesh jobs only
int check_staged(gpv_ctx* c, const gpv_grid* G)
{
    TRY(check_grid(c, G));
    if (G->job->nm != 1) return fail(GPV_ERR_STATE, "stage entry points take grids from gpv_grid_create");
    if (G->job->finished) return fail(GPV_ERR_STATE, "grid is already finished");
    return GPV_OK;
}

// Meshes [lo, hi) as one run.  A run that fails on a per-run limit found
// after Level1 (bin entries, boundary voxels, ray hits) or on device memory
// is split in halves and each half runs on its own, down to single meshes;
// a single mesh's failure is returned.  On failure no grid of the range is
// left behind.
int run_batch_range(gpv_ctx* c, const gpv_mesh_desc* d, int lo, int hi, int on_device, gpv_grid** out)
{
    Job* J = nullptr;
    int r = job_create(c, d + lo, hi - lo, on_device, &J);
    if (r == GPV_OK) {
        r = job_run(J);
        if (r == GPV_OK) {
            for (int m = lo; m < hi; ++m) out[m] = make_view(J, m - lo);
            return GPV_OK;
        }
        std::string keep = g_err;
        job_free(J);
        g_err = keep;
    }
    const bool split = r == GPV_ERR_OOM || r == GPV_ERR_CAPACITY || r == GPV_ERR_INVALID;
    if (hi - lo < 2 || !split) return r;   // (a failed run's own soup copy is redone by its halves)
    const int mid = lo + (hi - lo) / 2;
    TRY(run_batch_range(c, d, lo, mid, on_device, out));
    r = run_batch_range(c, d, mid, hi, on_device, out);
    if (r != GPV_OK) {
        std::string keep = g_err;
        for (int m = lo; m < mid; ++m) {
            gpv_grid_release(c, out[m]);
            out[m] = nullptr;
        }
        g_err = keep;
    }
    return r;
}

// batch uploads ahead of their runs: host soups of runs [lo, hi) copied in
// mesh order into one device buffer per run on the context's copy stream,
// an event after each run's copies, and the descriptors pointing there
constexpr int kStageFirst = 4;   // the first run takes 1/kStageFirst of the soup
constexpr int64_t kStageMinBytes = int64_t(2) << 20;

struct StagedSoups {
    std::vector<void*> buf;
    std::vector<hipEvent_t> ev;
    std::vector<gpv_mesh_desc> desc;
    std::vector<std::vector<SoupPiece>> pieces;   // gather tables, read by pending copies
};

constexpr int kMappedGatherBlocks = 16;   // gather workgroups per mesh for mapped host soups
constexpr int64_t kMappedMaxBytes = int64_t(256) << 10;   // larger soups go by DMA copy

// The device address of host soup [p, p + bytes) when it lies in one
// page-locked allocation mapped into the device's address space at the same
// address (hipHostMalloc memory on this platform), so that a kernel can read
// it across the bus; null for pageable or any other memory.
const float* mapped_host_soup(const float* p, int64_t bytes)
{
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    if (a.type != hipMemoryTypeHost || a.devicePointer != (const void*)p || a.hostPointer != (const void*)p)
        return nullptr;
    hipDeviceptr_t start = nullptr;
    size_t size = 0;
    if (hipPointerGetAttribute(&start, HIP_POINTER_ATTRIBUTE_RANGE_START_ADDR, (hipDeviceptr_t)p) != hipSuccess ||
        hipPointerGetAttribute(&size, HIP_POINTER_ATTRIBUTE_RANGE_SIZE, (hipDeviceptr_t)p) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    const uintptr_t lo = reinterpret_cast<uintptr_t>(start), q = reinterpret_cast<uintptr_t>(p);
    if (q < lo || q - lo > size || (uint64_t)(size - (q - lo)) < (uint64_t)bytes) return nullptr;
    return p;
}

void unstage_soups(gpv_ctx* c, StagedSoups* st)
{
    if (c->copy && !st->buf.empty()) (void)hipStreamSynchronize(c->copy);
    // the runs that read the buffers have finished (every run ends with a
    // host synchronisation), so the blocks go back to the pool
    for (void* p : st->buf) c->pool.put(p);
    for (hipEvent_t e : st->ev) (void)hipEventDestroy(e);
    st->buf.clear();
    st->ev.clear();
    st->desc.clear();
    st->pieces.clear();
}

int stage_soups(gpv_ctx* c, const gpv_mesh_desc* d, int nm, const std::vector<std::pair<int, int>>& runs,
                StagedSoups* st)
{
    if (!c->copy && hipStreamCreateWithFlags(&c->copy, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        c->copy = nullptr;
        return GPV_ERR_HIP;
    }
    // pool blocks are reused in the context stream's order; the copy stream
    // writes its blocks out of that order, so nothing may be pending there
    if (hipStreamSynchronize(c->stream) != hipSuccess || hipStreamSynchronize(c->side) != hipSuccess) {
        (void)hipGetLastError();
        return GPV_ERR_HIP;
    }
    st->desc.assign(d, d + nm);
    st->pieces.reserve(runs.size());
    for (const auto& rg : runs) {
        // small soups in mapped page-locked memory are gathered by one kernel
        // reading across the bus (a copy per mesh costs more than the bytes
        // of a small soup); every other soup is one DMA copy
        std::vector<SoupPiece> pc;
        std::vector<int64_t> at((size_t)(rg.second - rg.first));
        std::vector<char> gathered((size_t)(rg.second - rg.first), 0);
        int64_t bytes = 0;
        for (int m = rg.first; m < rg.second; ++m) {
            const int64_t b = d[m].ntris * 36;
            const float* src = b > 0 && b <= kMappedMaxBytes ? mapped_host_soup(d[m].tris, b) : nullptr;
            if (src) pc.push_back(SoupPiece{src, (uint64_t)bytes / 4, (uint64_t)b / 4});
            gathered[(size_t)(m - rg.first)] = src != nullptr;
            at[(size_t)(m - rg.first)] = bytes;
            bytes += b;
        }
        const size_t tab_at = ((size_t)bytes + 255) & ~size_t(255);
        const size_t alloc = pc.empty() ? (size_t)std::max<int64_t>(bytes, 1) : tab_at + sizeof(SoupPiece) * pc.size();
        void* p = nullptr;
        hipEvent_t e = nullptr;
        if (c->pool.get(alloc, &p) != hipSuccess) {
            (void)hipGetLastError();
            unstage_soups(c, st);
            return GPV_ERR_OOM;
        }
        st->buf.push_back(p);
        char* const dst = static_cast<char*>(p);
        for (int m = rg.first; m < rg.second; ++m) {
            const int64_t b = d[m].ntris * 36, o = at[(size_t)(m - rg.first)];
            if (b > 0 && !gathered[(size_t)(m - rg.first)] &&
                hipMemcpyAsync(dst + o, d[m].tris, (size_t)b, hipMemcpyHostToDevice, c->copy) != hipSuccess) {
                (void)hipGetLastError();
                unstage_soups(c, st);
                return GPV_ERR_HIP;
            }
            st->desc[(size_t)m].tris = reinterpret_cast<const float*>(dst + o);
        }
        if (!pc.empty()) {
            const int n = (int)pc.size();
            st->pieces.push_back(std::move(pc));
            SoupPiece* tab = reinterpret_cast<SoupPiece*>(dst + tab_at);
            if (hipMemcpyAsync(tab, st->pieces.back().data(), sizeof(SoupPiece) * (size_t)n, hipMemcpyHostToDevice,
                               c->copy) != hipSuccess) {
                (void)hipGetLastError();
                unstage_soups(c, st);
                return GPV_ERR_HIP;
            }
            launch_gather_soups(tab, n, reinterpret_cast<float*>(dst), kMappedGatherBlocks, true, c->copy);
            if (hipGetLastError() != hipSuccess) {
                unstage_soups(c, st);
                return GPV_ERR_HIP;
            }
        }
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess ||
            hipEventRecord(e, c->copy) != hipSuccess) {
            (void)hipGetLastError();
            if (e) (void)hipEventDestroy(e);
            unstage_soups(c, st);
            return GPV_ERR_HIP;
        }
        st->ev.push_back(e);
    }
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
    case GPV_ERR_TOO_LARGE: return "input exceeds the fixed scratch or result buffers";
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
    c->pool.set_cap(prop.totalGlobalMem / 8);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(GPV_ERR_HIP, "stream creation failed");
    }
    if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return fail(GPV_ERR_HIP, "stream creation failed");
    }
    if (ensure_hbuf(c, 4096) != GPV_OK) {
        (void)hipStreamDestroy(c->stream);
        (void)hipStreamDestroy(c->side);
        delete c;
        return fail(GPV_ERR_HIP, "pinned staging buffer allocation failed");
    }
    for (hipEvent_t& e : c->na_ev) {
        if (hipEventCreate(&e) != hipSuccess) {
            e = nullptr;
            gpv_ctx_destroy(c);
            return fail(GPV_ERR_HIP, "event creation failed");
        }
    }
    scratch_setup(device);
    *out = c;
    return GPV_OK;
}

int gpv_ctx_destroy(gpv_ctx* c)
{
    if (!c) return GPV_OK;
    {
        std::lock_guard<std::recursive_mutex> lk(c->mu);
        (void)hipSetDevice(c->device);
        (void)hipStreamSynchronize(c->stream);
        (void)hipStreamSynchronize(c->side);
        if (c->copy) (void)hipStreamSynchronize(c->copy);
        for (hipEvent_t e : c->events) (void)hipEventDestroy(e);
        c->events.clear();
        for (hipEvent_t e : c->na_ev)
            if (e) (void)hipEventDestroy(e);
        c->pool.trim();
        if (c->h_buf) (void)hipHostFree(c->h_buf);
        if (c->tab_buf) (void)hipHostFree(c->tab_buf);
        (void)hipStreamDestroy(c->stream);
        (void)hipStreamDestroy(c->side);
        if (c->copy) (void)hipStreamDestroy(c->copy);
    }
    delete c;
    return GPV_OK;
}

int gpv_device_upload(gpv_ctx* c, const void* host, int64_t bytes, void** dev)
{
    std::unique_lock<std::recursive_mutex> lk_;
    if (c) lk_ = std::unique_lock<std::recursive_mutex>(c->mu);
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

int gpv_host_alloc(int64_t bytes, void** host)
{
    if (!host || bytes < 0) return fail(GPV_ERR_INVALID, "bad host allocation arguments");
    *host = nullptr;
    const hipError_t e = hipHostMalloc(host, (size_t)std::max<int64_t>(bytes, 1), hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *host = nullptr;
        return fail(GPV_ERR_OOM, "hipHostMalloc of %lld bytes: %s", (long long)bytes, hipGetErrorString(e));
    }
    return GPV_OK;
}

int gpv_host_free(void* host)
{
    if (host) HIP_TRY(hipHostFree(host));
    return GPV_OK;
}

int gpv_ctx_get_tune(gpv_ctx* c, gpv_tune* out)
{
    if (!c || !out) return fail(GPV_ERR_INVALID, "NULL argument");
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    *out = c->tune;
    return GPV_OK;
}

int gpv_ctx_set_tune(gpv_ctx* c, const gpv_tune* t)
{
    if (!c || !t) return fail(GPV_ERR_INVALID, "NULL argument");
    auto wv = [](int w) { return w == 1 || w == 2 || w == 4; };
    if (t->expand_blocks_per_cu < 1 || t->expand_blocks_per_cu > 64 || t->l2_group_below < 0 ||
        t->sort_items_per_block < 256 || t->sort_items_per_block > (1 << 20) || !wv(t->l2_sat_waves) ||
        !wv(t->l2_parity_waves) || (t->expand_stage_tris != 128 && t->expand_stage_tris != 256))
        return fail(GPV_ERR_INVALID, "launch shape out of range");
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    c->tune = *t;
    return GPV_OK;
}

// device time of one batch run (ms_total of the batch), grids released
static int tune_trial(gpv_ctx* c, const gpv_mesh_desc* d, int nm, int on_device, int repeats, float* ms)
{
    std::vector<gpv_grid*> g((size_t)nm, nullptr);
    float best = 0.0f;
    for (int r = 0; r < repeats; ++r) {
        TRY(gpv_voxelize_batch(c, d, nm, on_device, g.data()));
        gpv_grid_info I;
        gpv_grid_get_info(g[0], &I);
        for (gpv_grid* x : g) gpv_grid_release(c, x);
        if (r == 0 || I.ms_total < best) best = I.ms_total;
    }
    *ms = best;
    return GPV_OK;
}

int gpv_ctx_autotune(gpv_ctx* c, const gpv_mesh_desc* d, int nm, int on_device, int repeats, gpv_tune* best,
                     float* best_ms)
{
    if (!c || !d || nm < 1) return fail(GPV_ERR_INVALID, "bad autotune arguments");
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    repeats = std::max(1, repeats);
    const gpv_tune start = c->tune;
    gpv_tune cur = start;
    float base = 0.0f;
    // one untimed run so that allocation and first-launch costs stay out
    float warm = 0.0f;
    int st = tune_trial(c, d, nm, on_device, 1, &warm);
    if (st == GPV_OK) st = tune_trial(c, d, nm, on_device, repeats, &base);
    const std::vector<int> expand = {4, 8, 16, 32}, group = {0, 4, 8, 16, 1 << 20}, items = {1024, 2048, 4096, 8192},
                           overlap = {0, 1}, waves = {1, 2, 4}, stage = {128, 256};
    auto sweep = [&](int gpv_tune::*field, const std::vector<int>& vals) -> int {
        for (int v : vals) {
            if (v == cur.*field) continue;
            gpv_tune t = cur;
            t.*field = v;
            c->tune = t;
            float ms = 0.0f;
            TRY(tune_trial(c, d, nm, on_device, repeats, &ms));
            if (ms < base * 0.99f) {
                base = ms;
                cur = t;
            }
        }
        c->tune = cur;
        return GPV_OK;
    };
    if (st == GPV_OK) st = sweep(&gpv_tune::expand_blocks_per_cu, expand);
    if (st == GPV_OK) st = sweep(&gpv_tune::l2_group_below, group);
    if (st == GPV_OK) st = sweep(&gpv_tune::sort_items_per_block, items);
    if (st == GPV_OK) st = sweep(&gpv_tune::l2_overlap, overlap);
    if (st == GPV_OK) st = sweep(&gpv_tune::l2_sat_waves, waves);
    if (st == GPV_OK) st = sweep(&gpv_tune::l2_parity_waves, waves);
    if (st == GPV_OK) st = sweep(&gpv_tune::expand_stage_tris, stage);
    if (st != GPV_OK) {
        c->tune = start;
        return st;
    }
    c->tune = cur;
    if (best) *best = cur;
    if (best_ms) *best_ms = base;
    return GPV_OK;
}

int gpv_ctx_pool_stats(gpv_ctx* c, gpv_pool_stats* out)
{
    if (!c || !out) return fail(GPV_ERR_INVALID, "NULL argument");
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    c->pool.stats(out);
    return GPV_OK;
}

int gpv_predict_footprint(const gpv_mesh_desc* d, int64_t* bytes)
{
    if (!d || !bytes || d->ntris < 0) return fail(GPV_ERR_INVALID, "bad mesh description");
    gpv_grid_info I;
    TRY(derive_grid(d->bmin, d->bmax, d->max_model_size, &d->params, &I));
    *bytes = predict_run_bytes(*d, I);
    return GPV_OK;
}

int gpv_ctx_set_mem_limit(gpv_ctx* c, int64_t bytes)
{
    if (!c || bytes < 0) return fail(GPV_ERR_INVALID, "bad memory limit");
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    c->pool.set_limit((size_t)bytes);
    return GPV_OK;
}

int gpv_device_free(gpv_ctx* c, void* dev)
{
    std::unique_lock<std::recursive_mutex> lk_;
    if (c) lk_ = std::unique_lock<std::recursive_mutex>(c->mu);
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

int gpv_grid_centers(const gpv_grid_info* info, int cuda_level2, int which, int axis, float* out)
{
    if (!info || !out || which < 0 || which > 3 || axis < 0 || axis > 2)
        return fail(GPV_ERR_INVALID, "gpv_grid_centers: bad argument");
    if (which >= 2 && info->n2 <= 0) return fail(GPV_ERR_INVALID, "gpv_grid_centers: Level2 is off");
    if (info->n[0] < 1 || info->n[1] < 1 || info->n[2] < 1 || info->n2 > 16)
        return fail(GPV_ERR_INVALID, "gpv_grid_centers: not a derived grid");
    std::vector<float> tab;
    size_t off[12];
    build_tables(*info, cuda_level2 != 0, tab, off);
    const size_t n = (size_t)info->n[axis] * (which >= 2 ? (size_t)info->n2 : 1);
    memcpy(out, tab.data() + off[3 * which + axis], n * sizeof(float));
    return GPV_OK;
}

int gpv_grid_create(gpv_ctx* c, const float* tris, int64_t T, int on_device, const float bmin[3],
                    const float bmax[3], float mms, const gpv_params* params, gpv_grid** out)
{
    std::unique_lock<std::recursive_mutex> lk_;
    if (c) lk_ = std::unique_lock<std::recursive_mutex>(c->mu);
    if (!c || !out || !bmin || !bmax || !params) return fail(GPV_ERR_INVALID, "NULL argument");
    if (T < 0 || (T > 0 && !tris)) return fail(GPV_ERR_INVALID, "bad triangle soup");
    if (T >= (int64_t(1) << 31)) return fail(GPV_ERR_INVALID, "at most 2^31-1 triangles");
    *out = nullptr;
    gpv_mesh_desc d{};
    d.tris = tris;
    d.ntris = T;
    for (int q = 0; q < 3; ++q) {
        d.bmin[q] = bmin[q];
        d.bmax[q] = bmax[q];
    }
    d.max_model_size = mms;
    d.params = *params;
    Job* J = nullptr;
    TRY(job_create(c, &d, 1, on_device, &J));
    *out = make_view(J, 0);
    return GPV_OK;
}

int gpv_classify_tessellation(gpv_ctx* c, gpv_grid* G, int64_t* needed)
{
    std::unique_lock<std::recursive_mutex> lk_;
    if (c) lk_ = std::unique_lock<std::recursive_mutex>(c->mu);
    TRY(check_staged(c, G));
    Job* J = G->job;
    int64_t nd = 0;
    HIP_TRY(hipSetDevice(c->device));
    TRY(stage_setup(J));
    TRY(launch_sat_half(J));
    TRY(read_counts(J, true));
    TRY(complete_sat(J, &nd));
    if (needed) *needed = nd;
    return GPV_OK;
}

int gpv_classify_inout(gpv_ctx* c, gpv_grid* G)
{
    std::unique_lock<std::recursive_mutex> lk_;
    if (c) lk_ = std::unique_lock<std::recursive_mutex>(c->mu);
    TRY(check_staged(c, G));
    Job* J = G->job;
    HIP_TRY(hipSetDevice(c->device));
    TRY(stage_setup(J));
    bool ok = false;
    for (int attempt = 0; attempt < 2 && !ok; ++attempt) {
        TRY(launch_ray_half(J));
        TRY(read_counts(J, false));
        TRY(complete_ray(J, &ok));
    }
    if (!ok) return fail(GPV_ERR_CAPACITY, "ray hit capacity did not converge");
    return launch_parity(J);
}

int gpv_compact_boundary(gpv_ctx* c, gpv_grid* G)
{
    std::unique_lock<std::recursive_mutex> lk_;
    if (c) lk_ = std::unique_lock<std::recursive_mutex>(c->mu);
    TRY(check_staged(c, G));
    if (!G->job->sat_done) return fail(GPV_ERR_STATE, "gpv_classify_tessellation has not completed");
    HIP_TRY(hipSetDevice(c->device));
    return launch_compact(G->job);
}

int gpv_classify_inout_level2(gpv_ctx* c, gpv_grid* G)
{
    std::unique_lock<std::recursive_mutex> lk_;
    if (c) lk_ = std::unique_lock<std::recursive_mutex>(c->mu);
    TRY(check_staged(c, G));
    if (!G->job->compact_done) return fail(GPV_ERR_STATE, "gpv_compact_boundary has not completed");
    HIP_TRY(hipSetDevice(c->device));
    return run_l2_parity(G->job);
}

int gpv_classify_tessellation_level2(gpv_ctx* c, gpv_grid* G)
{
    std::unique_lock<std::recursive_mutex> lk_;
    if (c) lk_ = std::unique_lock<std::recursive_mutex>(c->mu);
    TRY(check_staged(c, G));
    if (!G->job->compact_done) return fail(GPV_ERR_STATE, "gpv_compact_boundary has not completed");
    HIP_TRY(hipSetDevice(c->device));
    return run_l2_sat(G->job);
}

int gpv_grid_finish(gpv_ctx* c, gpv_grid* G)
{
    std::unique_lock<std::recursive_mutex> lk_;
    if (c) lk_ = std::unique_lock<std::recursive_mutex>(c->mu);
    TRY(check_staged(c, G));
    Job* J = G->job;
    if (!J->compact_done || !J->inout_done) return fail(GPV_ERR_STATE, "Level1 stages have not completed");
    if (J->n2max > 0 && (!J->l2p_done || !J->l2s_done)) return fail(GPV_ERR_STATE, "Level2 stages have not completed");
    HIP_TRY(hipSetDevice(c->device));
    TRY(run_normals(J));
    return job_finish(J);
}

int gpv_voxelize(gpv_ctx* c, const float* tris, int64_t T, int on_device, const float bmin[3],
                 const float bmax[3], float mms, const gpv_params* params, gpv_grid** out)
{
    std::unique_lock<std::recursive_mutex> lk_;
    if (c) lk_ = std::unique_lock<std::recursive_mutex>(c->mu);
    gpv_grid* G = nullptr;
    TRY(gpv_grid_create(c, tris, T, on_device, bmin, bmax, mms, params, &G));
    const int r = job_run(G->job);
    if (r != GPV_OK) {
        std::string keep = g_err;
        gpv_grid_release(c, G);
        g_err = keep;
        return r;
    }
    *out = G;
    return GPV_OK;
}

int gpv_voxelize_noalloc(gpv_ctx* c, const float* tris, int64_t T, const float bmin[3], const float bmax[3],
                         float mms, const gpv_params* params, uint64_t* l1_boundary_bits, uint64_t* l1_inside_bits,
                         int64_t l1_words, int32_t* bidx, uint64_t* l2_boundary_bits, uint64_t* l2_inside_bits,
                         int64_t nb_cap, gpv_grid_info* info)
{
    std::unique_lock<std::recursive_mutex> lk_;
    if (c) lk_ = std::unique_lock<std::recursive_mutex>(c->mu);
    if (!c || !bmin || !bmax || !params || !info || !l1_boundary_bits || !l1_inside_bits || l1_words < 0 ||
        nb_cap < 0 || (nb_cap > 0 && !bidx))
        return fail(GPV_ERR_INVALID, "gpv_voxelize_noalloc: NULL or negative argument");
    if (T < 0 || (T > 0 && !tris)) return fail(GPV_ERR_INVALID, "bad triangle soup");
    if (params->normals) return fail(GPV_ERR_INVALID, "gpv_voxelize_noalloc does not compute normals");
    if (params->level2 && nb_cap > 0 && (!l2_boundary_bits || !l2_inside_bits))
        return fail(GPV_ERR_INVALID, "gpv_voxelize_noalloc: Level2 planes are NULL");
    if (c->device < 0 || c->device >= kMaxDevices) return fail(GPV_ERR_INVALID, "device index out of range");
    char* arena = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (!g_scratch_ptr[c->device])
            return fail(GPV_ERR_HIP, "scratch address of device %d: %s", c->device,
                        hipGetErrorString(g_scratch_err[c->device]));
        arena = static_cast<char*>(g_scratch_ptr[c->device]);
    }
    // the size-independent part of the working set, checked before any
    // launch: boundary flag bytes and word counts per Level1 word, the
    // per-triangle set-up records
    gpv_grid_info gi;
    TRY(derive_grid(bmin, bmax, mms, params, &gi));
    const int64_t W = (int64_t)gi.n[2] * gi.n[1] * ((gi.n[0] + 63) / 64);
    const double base = 72.0 * (double)W + 140.0 * (double)T;
    if (base > (double)kScratchBytes)
        return fail(GPV_ERR_TOO_LARGE,
                    "gpv_voxelize_noalloc: %lld triangles on a %dx%dx%d grid need more than the %zu bytes of "
                    "scratch; use gpv_voxelize",
                    (long long)T, gi.n[0], gi.n[1], gi.n[2], kScratchBytes);
    std::lock_guard<std::mutex> call_lk(g_scratch_call_mu[c->device]);
    HIP_TRY(hipSetDevice(c->device));
    gpv_mesh_desc d{};
    d.tris = tris;
    d.ntris = T;
    for (int q = 0; q < 3; ++q) {
        d.bmin[q] = bmin[q];
        d.bmax[q] = bmax[q];
    }
    d.max_model_size = mms;
    d.params = *params;
    const NoAlloc na{arena, kScratchBytes, l1_boundary_bits, l1_inside_bits, l1_words,
                     bidx, l2_boundary_bits, l2_inside_bits, nb_cap};
    Job* J = nullptr;
    TRY(job_create(c, &d, 1, 1, &J, &na));
    int r = job_run(J);
    if (r == GPV_OK) *info = J->info[0];
    std::string keep = g_err;
    job_free(J);   // waits for both streams: the scratch is free for the next call
    g_err = keep;
    return r;
}

int gpv_voxelize_batch(gpv_ctx* c, const gpv_mesh_desc* meshes, int nmeshes, int on_device, gpv_grid** out)
{
    std::unique_lock<std::recursive_mutex> lk_;
    if (c) lk_ = std::unique_lock<std::recursive_mutex>(c->mu);
    if (!c || !meshes || !out || nmeshes < 1) return fail(GPV_ERR_INVALID, "bad batch arguments");
    for (int m = 0; m < nmeshes; ++m) out[m] = nullptr;
    // every mesh is checked first, so that a failure below is a batch limit
    // or memory, which splitting resolves
    std::vector<int64_t> vox((size_t)nmeshes), cols((size_t)nmeshes), foot((size_t)nmeshes);
    for (int m = 0; m < nmeshes; ++m) {
        const gpv_mesh_desc& d = meshes[m];
        if (d.ntris < 0 || (d.ntris > 0 && !d.tris)) return fail(GPV_ERR_INVALID, "mesh %d: bad triangle soup", m);
        gpv_grid_info I;
        TRY(derive_grid(d.bmin, d.bmax, d.max_model_size, &d.params, &I));
        const int64_t nx = I.n[0], ny = I.n[1], nz = I.n[2];
        if (d.ntris >= (int64_t(1) << 31) || nx * ny > (int64_t(1) << 27) ||
            nz * ny * ((nx + 63) / 64) * 64 >= (int64_t(1) << 32))
            return fail(GPV_ERR_INVALID, "mesh %d exceeds the limits of one run (2^31-1 triangles, 2^27 columns, "
                        "2^32-1 padded voxels)", m);
        vox[(size_t)m] = nz * ny * ((nx + 63) / 64) * 64;
        cols[(size_t)m] = nx * ny;
        TRY(gpv_predict_footprint(&d, &foot[(size_t)m]));
    }
    HIP_TRY(hipSetDevice(c->device));
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
        (void)hipGetLastError();
        free_b = size_t(1) << 36;
    }
    gpv_pool_stats ps;
    c->pool.stats(&ps);
    int64_t room = (int64_t)free_b + ps.cached_bytes;
    if (ps.limit_bytes > 0) room = std::min<int64_t>(room, ps.limit_bytes - ps.live_bytes);
    const int64_t chunk_bytes = std::max<int64_t>(room / 2, 1);
    // Host soups of a batch above 2 MiB go up in two runs: both runs' soups are
    // uploaded ahead on the copy stream and each run waits only for its own,
    // so the second upload proceeds under the first run's kernels.
    int64_t soup_bytes = 0;
    for (int m = 0; m < nmeshes; ++m) soup_bytes += meshes[m].ntris * 36;
    const int pieces = !on_device && nmeshes >= 2 && soup_bytes >= kStageMinBytes ? 2 : 1;
    // the piece of a mesh: where the middle of its soup falls; the first
    // piece is the first kStageFirst of the soup, so its upload (the one
    // nothing overlaps) is short and its run outlasts the rest's upload
    std::vector<int> piece_of((size_t)nmeshes, 0);
    {
        const int64_t first = std::max<int64_t>(1, soup_bytes / kStageFirst);
        int64_t at = 0;
        for (int m = 0; m < nmeshes && pieces > 1; ++m) {
            const int64_t b = meshes[m].ntris * 36;
            piece_of[(size_t)m] = at + b / 2 < first ? 0 : 1;
            at += b;
        }
    }
    // consecutive runs within the per-run limits and half the free memory
    std::vector<std::pair<int, int>> runs;
    for (int lo = 0; lo < nmeshes;) {
        int hi = lo;
        int64_t T = 0, V = 0, C = 0, F = 0;
        while (hi < nmeshes) {
            const size_t k = (size_t)hi;
            const bool first = hi == lo;
            if (!first && (T + meshes[hi].ntris >= (int64_t(1) << 31) || V + vox[k] >= (int64_t(1) << 32) ||
                           C + cols[k] > (int64_t(1) << 27) || F + foot[k] > chunk_bytes ||
                           piece_of[k] != piece_of[(size_t)lo]))
                break;
            T += meshes[hi].ntris;
            V += vox[k];
            C += cols[k];
            F += foot[k];
            ++hi;
        }
        runs.emplace_back(lo, hi);
        lo = hi;
    }
    StagedSoups st;
    const gpv_mesh_desc* src = meshes;
    int src_on_device = on_device;
    if (pieces > 1 && runs.size() > 1 && stage_soups(c, meshes, nmeshes, runs, &st) == GPV_OK) {
        src = st.desc.data();
        src_on_device = 1;
    }
    int r = GPV_OK;
    for (size_t q = 0; q < runs.size() && r == GPV_OK; ++q) {
        if (!st.ev.empty() && hipStreamWaitEvent(c->stream, st.ev[q], 0) != hipSuccess) {
            r = fail(GPV_ERR_HIP, "stream wait on a staged upload failed");
            break;
        }
        r = run_batch_range(c, src, runs[q].first, runs[q].second, src_on_device, out);
        if (r != GPV_OK) {
            std::string keep = g_err;
            for (int m = 0; m < runs[q].first; ++m) {
                gpv_grid_release(c, out[m]);
                out[m] = nullptr;
            }
            g_err = keep;
        }
    }
    unstage_soups(c, &st);
    return r;
}

int gpv_grid_get_info(const gpv_grid* G, gpv_grid_info* info)
{
    if (!G || !G->job || !info) return fail(GPV_ERR_INVALID, "NULL argument");
    *info = G->job->info[G->m];
    return GPV_OK;
}

int gpv_grid_device_views(const gpv_grid* G, const uint64_t** l1b, const uint64_t** l1i, const uint64_t** l2b,
                          const uint64_t** l2i, const int32_t** bidx)
{
    if (!G || !G->job) return fail(GPV_ERR_INVALID, "NULL grid");
    const Job* J = G->job;
    const MeshOff& o = J->off[G->m];
    if (l1b) *l1b = J->bnd ? J->bnd + o.word0 : nullptr;
    if (l1i) *l1i = J->ins ? J->ins + o.word0 : nullptr;
    if (l2b) *l2b = J->sat2 && J->g[G->m].n2 > 0 ? J->sat2 + o.l2w0 : nullptr;
    if (l2i) *l2i = J->par2 && J->g[G->m].n2 > 0 ? J->par2 + o.l2w0 : nullptr;
    if (bidx) *bidx = J->bidx ? J->bidx + o.b0 : nullptr;
    return GPV_OK;
}

int gpv_grid_download_bits(gpv_ctx* c, const gpv_grid* G, uint64_t* l1b, uint64_t* l1i, uint64_t* l2b,
                           uint64_t* l2i)
{
    std::unique_lock<std::recursive_mutex> lk_;
    if (c) lk_ = std::unique_lock<std::recursive_mutex>(c->mu);
    TRY(check_grid(c, G));
    const Job* J = G->job;
    if (!J->finished) return fail(GPV_ERR_STATE, "grid is not finished");
    HIP_TRY(hipSetDevice(c->device));
    const GridDev& g = J->g[G->m];
    const MeshOff& o = J->off[G->m];
    const size_t w1 = (size_t)g.nz * g.ny * g.nxw * 8;
    const size_t w2 = g.n2 > 0 ? (size_t)J->nb(G->m) * g.w2 * 8 : 0;
    hipStream_t s = c->stream;
    if (l1b && w1) HIP_TRY(hipMemcpyAsync(l1b, J->bnd + o.word0, w1, hipMemcpyDeviceToHost, s));
    if (l1i && w1) HIP_TRY(hipMemcpyAsync(l1i, J->ins + o.word0, w1, hipMemcpyDeviceToHost, s));
    if (l2b && w2) HIP_TRY(hipMemcpyAsync(l2b, J->sat2 + o.l2w0, w2, hipMemcpyDeviceToHost, s));
    if (l2i && w2) HIP_TRY(hipMemcpyAsync(l2i, J->par2 + o.l2w0, w2, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return GPV_OK;
}

static int download_impl(gpv_ctx* c, const gpv_grid* G, uint8_t unit, uint8_t* l1, int32_t* prefix,
                         int32_t* bidx, uint8_t* l2)
{
    std::unique_lock<std::recursive_mutex> lk_;
    if (c) lk_ = std::unique_lock<std::recursive_mutex>(c->mu);
    TRY(check_grid(c, G));
    Job* J = G->job;
    if (!J->finished) return fail(GPV_ERR_STATE, "grid is not finished");
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int m = G->m;
    const GridDev& g = J->g[m];
    const MeshOff& o = J->off[m];
    const int64_t n1 = J->info[m].n1;
    const int64_t nb = J->nb(m);
    const int64_t n23 = (int64_t)g.n2 * g.n2 * g.n2;
    if (l1) {
        uint8_t* d = nullptr;
        TRY(alloc(J, &d, (size_t)n1));
        launch_export_l1(J->bnd + o.word0, J->ins + o.word0, g, unit, d, s);
        HIP_TRY(hipMemcpyAsync(l1, d, (size_t)n1, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        release(J, d);
    }
    if (prefix) {
        int32_t* d = nullptr;
        TRY(alloc(J, &d, (size_t)n1));
        launch_export_prefix(J->bnd + o.word0, J->wprefix + o.word0, o.b0, g, d, s);
        HIP_TRY(hipMemcpyAsync(prefix, d, (size_t)n1 * 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        release(J, d);
    }
    if (bidx && nb > 0) {
        HIP_TRY(hipMemcpyAsync(bidx, J->bidx + o.b0, (size_t)nb * 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    if (l2 && g.n2 > 0 && nb > 0) {
        uint8_t* d = nullptr;
        TRY(alloc(J, &d, (size_t)(nb * n23)));
        launch_export_l2(J->sat2 + o.l2w0, J->par2 + o.l2w0, nb, g, unit, d, s);
        HIP_TRY(hipMemcpyAsync(l2, d, (size_t)(nb * n23), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        release(J, d);
    }
    return GPV_OK;
}

// states 0 outside / 1 inside / 2 boundary
int gpv_grid_download(gpv_ctx* c, const gpv_grid* G, uint8_t* l1, int32_t* prefix, int32_t* bidx, uint8_t* l2)
{
    return download_impl(c, G, 1, l1, prefix, bidx, l2);
}

int gpv_grid_download_normals(gpv_ctx* c, const gpv_grid* G, float* l1n, float* l2n)
{
    std::unique_lock<std::recursive_mutex> lk_;
    if (c) lk_ = std::unique_lock<std::recursive_mutex>(c->mu);
    TRY(check_grid(c, G));
    Job* J = G->job;
    if (!J->finished) return fail(GPV_ERR_STATE, "grid is not finished");
    if (!J->g[G->m].normals) return fail(GPV_ERR_STATE, "grid was made without params.normals");
    HIP_TRY(hipSetDevice(c->device));
    const MeshOff& o = J->off[G->m];
    const int64_t nb = J->nb(G->m);
    const int n2 = J->g[G->m].n2;
    const int64_t n23 = (int64_t)n2 * n2 * n2;
    if (l1n && nb > 0)
        HIP_TRY(hipMemcpyAsync(l1n, J->l1n + 3 * (size_t)o.b0, (size_t)nb * 3 * sizeof(float), hipMemcpyDeviceToHost,
                               c->stream));
    if (l2n && nb > 0 && n23 > 0)
        HIP_TRY(hipMemcpyAsync(l2n, J->l2n + 3 * (size_t)o.l2n0, (size_t)(nb * n23) * 3 * sizeof(float),
                               hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return GPV_OK;
}

// quantised normal files of SaveVoxelization (src/Object.cpp:3024-3050)
static int normal_bytes(gpv_ctx* c, const gpv_grid* G, std::vector<uint8_t>& l1b, std::vector<uint8_t>& l2b)
{
    Job* J = G->job;
    hipStream_t s = c->stream;
    const int m = G->m;
    const GridDev& g = J->g[m];
    const MeshOff& o = J->off[m];
    const int64_t n1 = J->info[m].n1, nb = J->nb(m);
    const int64_t n23 = (int64_t)g.n2 * g.n2 * g.n2;
    l1b.assign((size_t)n1 * 3, 0);
    uint8_t* d = nullptr;
    TRY(alloc(J, &d, (size_t)n1 * 3));
    launch_export_l1_normals(J->bnd + o.word0, J->wprefix + o.word0, o.b0, J->l1n + 3 * (size_t)o.b0, g, d, s);
    HIP_TRY(hipMemcpyAsync(l1b.data(), d, (size_t)n1 * 3, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    release(J, d);
    l2b.assign(g.n2 > 0 ? (size_t)(nb * n23 * 3) : 0, 0);
    if (!l2b.empty()) {
        TRY(alloc(J, &d, l2b.size()));
        launch_export_l2_normals(J->l2n + 3 * (size_t)o.l2n0, (int64_t)l2b.size(), d, s);
        HIP_TRY(hipMemcpyAsync(l2b.data(), d, l2b.size(), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        release(J, d);
    }
    return GPV_OK;
}

// Object::SaveVoxelization, src/Object.cpp:2934-3075 (in/out, prefix and
// config files; the config uses the ostream default of 6 significant digits).
int gpv_save_raw(gpv_ctx* c, const gpv_grid* G, int obj_id, const char* dir)
{
    std::unique_lock<std::recursive_mutex> lk_;
    if (c) lk_ = std::unique_lock<std::recursive_mutex>(c->mu);
    TRY(check_grid(c, G));
    const gpv_grid_info& I = G->job->info[G->m];
    const bool l2 = I.n2 > 0;
    const int64_t n23 = (int64_t)I.n2 * I.n2 * I.n2;
    std::vector<uint8_t> l1v((size_t)I.n1);
    std::vector<int32_t> pre(l2 ? (size_t)I.n1 : 0);
    std::vector<uint8_t> l2v(l2 ? (size_t)(I.nb * n23) : 0);
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
    if (G->job->g[G->m].normals) {
        std::vector<uint8_t> n1b, n2b;
        TRY(normal_bytes(c, G, n1b, n2b));
        TRY(put("Level1Normal.raw", n1b.data(), n1b.size()));
        if (l2) TRY(put("Level2Normal.raw", n2b.data(), n2b.size()));
    }
    return GPV_OK;
}

int gpv_grid_release(gpv_ctx* c, gpv_grid* G)
{
    std::unique_lock<std::recursive_mutex> lk_;
    if (c) lk_ = std::unique_lock<std::recursive_mutex>(c->mu);
    if (!G) return GPV_OK;
    if (!c || G->ctx != c) return fail(GPV_ERR_INVALID, "grid belongs to another context");
    Job* J = G->job;
    delete G;
    if (J && --J->refs == 0) job_free(J);
    return GPV_OK;
}

}  // extern "C"
