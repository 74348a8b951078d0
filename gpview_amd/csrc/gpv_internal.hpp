// gpv_internal.hpp -- shared host/device declarations of libgpv (not public).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

struct gpv_ctx;

namespace gpv {

// Grid constants and per-axis center tables of one mesh.
// Bit layout of every Level1 bitplane: word w = (z*ny + y)*nxw + x/64,
// bit x%64 (rows padded to whole 64-bit words), so words are in the
// reference's linear order z*ny*nx + y*nx + x and a boundary rank is a
// popcount prefix.  Level2: w2 = ceil(n2^3/64) words per boundary voxel,
// bit s = (r*n2 + q)*n2 + p inside the block ([b][r][q][p] order).
struct GridDev {
    int nx, ny, nz, nxw;
    int n2, w2;
    int normals;           // this mesh asks for Level1/Level2 normals
    int l2cuda;            // Level2 in the reference's CUDA arithmetic (gpv_params::cuda_level2)
    int r2mono;            // each Level2 ray-origin table non-decreasing over its whole length
    int l1st;              // Level1 inside/outside by the GL-stencil-equivalent fill (gpv_params::l1_stencil)
    float bminx, bminy, bminz, bmaxx, bmaxy, bmaxz;
    float h1x, h1y, h1z;   // Level1 half extents
    float h2x, h2y, h2z;   // Level2 half extents
    const float* c1x; const float* c1y; const float* c1z;   // [n]    Level1 SAT centers
    const float* r1x; const float* r1y; const float* r1z;   // [n]    Level1 ray origins
    const float* c2x; const float* c2y; const float* c2z;   // [n*n2] Level2 SAT centers
    const float* r2x; const float* r2y; const float* r2z;   // [n*n2] Level2 ray origins
};

// Where one mesh of a batch lives in the batch-wide arrays.  A batch
// concatenates the meshes' soups (tri0), Level1 bitplanes (word0), xy
// columns (col0), boundary voxels (b0) and Level2 words (l2w0); entry nm is
// the sentinel holding the totals.  b0 and l2w0 are filled on the device
// once the boundary scan is known.
struct MeshOff {
    uint64_t word0;
    uint64_t l2w0;
    uint64_t l2n0;   // first Level2 normal (one per Level2 voxel of meshes with normals)
    uint32_t col0;
    uint32_t tri0;
    uint32_t b0;
    uint32_t ptask0;   // first Level1 parity task (16 layers x 64 columns of one row)
    uint32_t pw0;      // first Level2 parity wave (set on the host once boundary counts are known)
    uint32_t pad;
};

// Per-mesh results accumulated on the device.
struct MeshCtr {
    unsigned long long tests_l1;      // sum of clipped index-AABB volumes
    unsigned long long l1_inside;
    unsigned long long l2_inside;
    unsigned long long l2_boundary;
    unsigned long long pairs;         // Level1 (voxel, triangle) bin entries
    unsigned int max_per_voxel;
    unsigned int pad;
};

// A batch as kernels see it.  With nm == 1 kernels run their single-mesh
// instantiation: the grid comes from the by-value GridDev, every offset is
// 0 and only ctr is used.
struct Meshes {
    const GridDev* grids;       // [nm]
    MeshOff* off;               // [nm + 1]
    const uint32_t* tri_mesh;   // [T]  mesh of each triangle (non-decreasing)
    uint32_t* bmesh;            // [B]  mesh of each boundary voxel (non-decreasing)
    MeshCtr* ctr;               // [nm]
    int nm;
};

// Append-buffer descriptor: a device counter and a capacity.  Writers past
// the capacity are dropped; the host sees count > cap and regrows.
struct Append64 {
    unsigned long long* count;
    uint64_t* key;
    uint32_t* val;     // optional payload (nullptr if none)
    uint64_t cap;
};

// Device-side counters of one pipeline run (batch-wide).
struct Counters {
    unsigned long long pairs;       // Level1 SAT (voxel, triangle) pairs
    unsigned long long colents;     // (column, triangle, zmax) entries
    unsigned long long hits;        // (column, triangle) ray hits, Level1
    unsigned long long bad_input;   // nonzero: a vertex coordinate is not finite or beyond +-2^40
};

// sets the calling thread's gpv_last_error text and returns code
int set_error(int code, const char* msg);
// restarts the context pool's peak of bytes in use from the bytes in use now
void pool_reset_peak(gpv_ctx* c);

// ---- launch wrappers (gpv_kernels.hip) ------------------------------------
// Launchers that take the batch view M and a grid g0 run the single-mesh
// instantiation (grid g0) when M.nm == 1, the multi-mesh one otherwise.
void launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s);
void launch_fill_u64(uint64_t* p, uint64_t v, int64_t n, hipStream_t s);

void launch_mesh_ids(const MeshOff* off, int nm, int64_t T, uint32_t* tri_mesh, hipStream_t s);
// part 1 (and 3) also raise *bad_input for a coordinate outside the domain
void launch_tri_setup(const float* tris, int64_t T, const GridDev& g0, const Meshes& M, int4* box0, int4* box1,
                      uint32_t* sat_items, int4* ray_rng, uint32_t* ray_items, double4* slab1, float4* slab2, int part,
                      unsigned long long* bad_input, hipStream_t s);

// exclusive scan of n u32 into out[0..n] (out[n] = total); tmp sized by
// scan_tmp_elems(n).  If max_out != nullptr the max input is atomically
// folded into it.
size_t scan_tmp_elems(int64_t n);
void launch_scan_u32_u64(const uint32_t* in, uint64_t* out, int64_t n, uint64_t* tmp, hipStream_t s);
void launch_scan_u32_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* tmp, unsigned int* max_out,
                         hipStream_t s);

// stage: triangles staged per LDS pass (128 or 256)
void launch_sat_expand(const float* tris, int64_t T, const int4* box0, const int4* box1, const double4* slab1,
                       const uint64_t* off, const GridDev& g0, const Meshes& M, uint8_t* bbytes, Append64 pairs,
                       Append64 cols, int blocks, int stage, hipStream_t s);
// boundary bytes (W * 64) -> boundary bitplane (W words)
void launch_pack_bytes(const uint8_t* by, int64_t W, uint64_t* bits, hipStream_t s);
void launch_ray_expand(const float* tris, int64_t T, const int4* ray_rng, const uint64_t* off, const GridDev& g0,
                       const Meshes& M, Append64 hits, int blocks, int stage, hipStream_t s);

// counting sort helpers (keys hold a batch-wide column index in the high word)
void launch_count_keys_hi(const uint64_t* key, uint64_t n, uint32_t* cnt, hipStream_t s);
void launch_scatter_hits(const float* tris, const uint64_t* key, uint64_t n, const GridDev& g0, const Meshes& M,
                         const uint32_t* off, uint32_t* cnt, float4* rec, hipStream_t s);
// scratch of a two-level counting sort (ts_tiles(nkeys) tiles): total and
// fill counters [2*tiles + 1], tile offsets [tiles + 1], scan temporary
// [scan_tmp_elems(tiles)], bucket [items]
struct TileSortTmp {
    uint32_t* total;
    uint32_t* off;
    uint32_t* scan;
    uint4* bucket;
};
size_t ts_tiles(uint64_t nkeys);

// Level2 SAT voxel groups (k_l2_sat_grp): at most kL2GroupMax voxels, their
// Level2 words within one wave's 64 and their 3 * n2 centers each within
// kL2GroupCenters floats of LDS
constexpr int kL2GroupMax = 16;
constexpr int kL2GroupCenters = 192;
__host__ __device__ inline int l2_group_voxels(int w2, int n2)
{
    int c = w2 > 0 ? 64 / w2 : 64;
    if (c > kL2GroupMax) c = kL2GroupMax;
    if (n2 > 0 && c > kL2GroupCenters / (3 * n2)) c = kL2GroupCenters / (3 * n2);
    return c;
}
// column entries (key = column << 32 | triangle) -> CSR by column with the MT
// records; ncol <= 2^27
void launch_bin_cols(const float* tris, const uint64_t* key, const uint32_t* zmax, uint64_t n, uint32_t ncol,
                     const TileSortTmp& tmp, int items_per_block, uint32_t* coff, float4* rec, hipStream_t s);

// Level1 parity tasks: L1P_Z consecutive layers of one 64-column row strip
constexpr int L1P_Z = 16;
inline int64_t l1_parity_tasks(const GridDev& g) { return (int64_t)((g.nz + L1P_Z - 1) / L1P_Z) * g.ny * g.nxw; }
// one mesh's soup (n floats, device memory or page-locked host memory mapped
// into the device's address space) and its place in the batch's soup
struct SoupPiece {
    const float* src;
    uint64_t at;
    uint64_t n;
};
void launch_gather_soups(const SoupPiece* pc, int nm, float* dst, int blocks, bool host, hipStream_t s);

// GL-stencil-equivalent Level1 fill of one mesh (soup tris[T], grid g):
// overwrites the mesh's Level1 inside bits ins[nz*ny*nxw].  Scratch: cnt[T],
// off[T+1], scan_tmp[scan_tmp_elems(T)], tog[l1_stencil_words(g)].
size_t l1_stencil_words(const GridDev& g);
void launch_l1_stencil(const float* tris, int64_t T, const GridDev& g, uint32_t* cnt, uint64_t* off,
                       uint64_t* scan_tmp, uint64_t* tog, uint64_t* ins, hipStream_t s);
void launch_l1_parity(const float4* rec, const uint32_t* hoff, const GridDev& g0, const Meshes& M, int64_t NT,
                      uint64_t* ins, hipStream_t s);

void launch_popc_words(const uint64_t* bits, int64_t W, uint32_t* cnt, hipStream_t s);
// b0 and l2w0 of every mesh (and the totals in entry nm) from the word prefix
void launch_mesh_bounds(const uint32_t* wprefix, const Meshes& M, hipStream_t s);
void launch_boundary_index(const uint64_t* bnd, const uint32_t* wprefix, const GridDev& g0, const Meshes& M,
                           int64_t W, int32_t* bidx, hipStream_t s);
// pairs: key = batch-wide boundary bit index, val = triangle
// bin pairs -> CSR by boundary rank (boff[B + 1], list, Level2 box-axis
// masks); B <= 2^27.  The pairs' voxel bits are replaced by their boundary
// ranks (ranked: they hold ranks already)
void launch_bin_pairs(uint64_t* bits, bool ranked, const uint32_t* tri, uint64_t n, const uint64_t* bnd,
                      const uint32_t* wprefix, uint32_t B, const TileSortTmp& tmp, int items_per_block,
                      uint32_t* boff, unsigned int* max_out, uint32_t* list, const float* tris,
                      const int32_t* bidx, const GridDev& g0, const Meshes& M, uint64_t* pmask, hipStream_t s);
// per-mesh bin totals and largest bin (multi-mesh runs)
void launch_mesh_bins(const uint32_t* boff, int64_t B, const Meshes& M, hipStream_t s);

// Level2 SAT work groups of consecutive boundary voxels: flag[B], gidx[B+1]
// (gidx[B] = group count), gstart[B+1]
void launch_l2_groups(const uint32_t* boff, int64_t B, const GridDev& g0, const Meshes& M, uint32_t* flag,
                      uint32_t* gidx, uint32_t* scan_tmp, uint32_t* gstart, hipStream_t s);
// gmax: an upper bound of the group count (grouped form); waves: per block (1, 2, 4)
void launch_l2_sat(const float* tris, const int32_t* bidx, int64_t B, const uint32_t* boff, const uint32_t* list,
                   const uint64_t* pmask, const float4* slab2, const uint32_t* gstart, const uint32_t* ngroups,
                   int64_t gmax, int waves, const GridDev& g0, const Meshes& M, uint64_t* sat2, hipStream_t s);
// det/u/v pass bits per (column entry, sub-column): stride words per entry
void launch_col_pass(const float4* rec, const GridDev& g0, const Meshes& M, uint64_t nent, int stride, int waves,
                     uint64_t* cpass, hipStream_t s);
// Level2 parity waves of one mesh: one per boundary voxel, one per four when
// n2 <= 4, none when Level2 is off
int64_t l2_parity_waves(int n2, int64_t nb);
// n2max: the largest Level2 resolution of the run
void launch_l2_parity(const int32_t* bidx, int64_t B, int64_t nwaves, const uint32_t* coff, const float4* rec,
                      const uint64_t* cpass, int stride, int n2max, int waves, const GridDev& g0, const Meshes& M,
                      uint64_t* par2, hipStream_t s);

// normals (src/Object.cpp:3219-3253 Level1, 2372-2421 + 2705-2720 Level2)
void launch_face_normals(const float* tris, int64_t T, float* fn, hipStream_t s);
void launch_sort_bins(const uint32_t* list, const uint32_t* boff, int64_t B, const GridDev& g0, const Meshes& M,
                      uint32_t* slist, hipStream_t s);
void launch_l1_normals(const uint32_t* slist, const uint32_t* boff, int64_t B, const float* fn, const GridDev& g0,
                       const Meshes& M, float* l1n, hipStream_t s);
void launch_l2_normals(const float* tris, const int32_t* bidx, int64_t B, const uint32_t* boff,
                       const uint32_t* slist, const float* fn, const GridDev& g0, const Meshes& M, float* l2n,
                       hipStream_t s);
void launch_export_l1_normals(const uint64_t* bnd, const uint32_t* wprefix, uint32_t base, const float* l1n,
                              const GridDev& g, uint8_t* out, hipStream_t s);
void launch_export_l2_normals(const float* l2n, int64_t n, uint8_t* out, hipStream_t s);

void launch_count_states(const uint64_t* bnd, const uint64_t* ins, int64_t W, const uint64_t* sat2,
                         const uint64_t* par2, int64_t W2, const Meshes& M, hipStream_t s);

// egress of one mesh (pointers already offset to that mesh's words)
void launch_export_l1(const uint64_t* bnd, const uint64_t* ins, const GridDev& g, uint8_t unit, uint8_t* out,
                      hipStream_t s);
void launch_export_prefix(const uint64_t* bnd, const uint32_t* wprefix, uint32_t base, const GridDev& g,
                          int32_t* out, hipStream_t s);
void launch_export_l2(const uint64_t* sat2, const uint64_t* par2, int64_t B, const GridDev& g, uint8_t unit,
                      uint8_t* out, hipStream_t s);

}  // namespace gpv
