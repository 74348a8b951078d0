/* The C ABI as a C99 caller uses it: the stage sequence, the batch call,
 * the no-allocation entry point and the serving loop of the integration
 * notes, written against include/gpv.h, gpv_mesh.h and gpv_serve.h.  It is
 * only compiled (and linked against libgpv), never run, by
 * tests/test_abi.py::test_c_headers_compile_as_c99. */
#include <stddef.h>
#include <stdint.h>

#include "gpv.h"
#include "gpv_mesh.h"
#include "gpv_serve.h"

int voxelize_staged(const float *soup, int64_t ntris, const float *verts, int64_t nverts)
{
    gpv_ctx *ctx = NULL;
    gpv_grid *grid = NULL;
    float bmin[3], bmax[3], mms = 0.0f;
    int64_t needed = 0;
    int st = gpv_ctx_create(-1, 0, &ctx);
    if (st != GPV_OK) return st;
    st = gpv_mesh_bbox(verts, nverts, bmin, bmax, &mms);
    gpv_params p = {256, 8, 1, 1, 0, 0, 0};
    if (st == GPV_OK) st = gpv_grid_create(ctx, soup, ntris, 0, bmin, bmax, mms, &p, &grid);
    if (st == GPV_OK) st = gpv_classify_inout(ctx, grid);
    do {
        if (st == GPV_OK) st = gpv_classify_tessellation(ctx, grid, &needed);
    } while (st == GPV_OK && needed > 0);
    if (st == GPV_OK) st = gpv_compact_boundary(ctx, grid);
    if (st == GPV_OK) st = gpv_classify_inout_level2(ctx, grid);
    if (st == GPV_OK) st = gpv_classify_tessellation_level2(ctx, grid);
    if (st == GPV_OK) st = gpv_grid_finish(ctx, grid);
    if (st == GPV_OK) st = gpv_save_raw(ctx, grid, -1, ".");
    if (grid) gpv_grid_release(ctx, grid);
    gpv_ctx_destroy(ctx);
    return st;
}

int voxelize_batch_and_serve(gpv_ctx *ctx, gpv_mesh_desc *d, int n, gpv_grid **out)
{
    int st = gpv_voxelize_batch(ctx, d, n, 1, out);
    if (st != GPV_OK) return st;
    gpv_server_opts o = {-1, 2, 200000000000LL, 0, 0, 0};
    gpv_server *srv = NULL;
    st = gpv_server_create(&o, &srv);
    if (st != GPV_OK) return st;
    uint64_t id = 0;
    gpv_grid *grid = NULL;
    gpv_ctx *lane = NULL;
    gpv_request_stats rs;
    st = gpv_server_submit(srv, &d[0], 1, &id);
    if (st == GPV_OK) st = gpv_server_wait(srv, id, -1.0, &grid, &lane, &rs);
    if (st == GPV_OK) st = gpv_server_release(srv, id);
    gpv_server_destroy(srv);
    return st;
}

int voxelize_without_allocation(gpv_ctx *ctx, const float *d_soup, int64_t ntris, const float bmin[3],
                                const float bmax[3], float mms, uint64_t *l1b, uint64_t *l1i, int32_t *bidx,
                                uint64_t *l2b, uint64_t *l2i, int64_t nb_cap)
{
    gpv_params p = {256, 8, 1, 0, 0, 0, 0};
    gpv_grid_info info;
    int st = gpv_grid_derive(bmin, bmax, mms, &p, &info);
    if (st != GPV_OK) return st;
    const int64_t w1 = (int64_t)info.n[2] * info.n[1] * ((info.n[0] + 63) / 64);
    return gpv_voxelize_noalloc(ctx, d_soup, ntris, bmin, bmax, mms, &p, l1b, l1i, w1, bidx, l2b, l2i, nb_cap,
                                &info);
}
