/*
 * gpv_serve.h -- serving loop of libgpv: voxelization requests accepted as
 * they arrive, batched continuously onto the device, results kept resident.
 *
 * The reference voxelizes on a keypress, one Object at a time
 * (src/GPView.cpp:1378-1380, Object::PerformVoxelization src/Object.cpp:3077);
 * this layer has no counterpart there.  It is built on gpv_voxelize_batch
 * (include/gpv.h): each request's grid is exactly what gpv_voxelize of that
 * mesh alone produces, whatever it was batched with.
 *
 * Model:
 *  - `lanes` host threads, each owning a gpv_ctx (stream + pool).  A lane
 *    that is free takes the waiting requests in arrival order, up to the
 *    batch limits and the memory budget, and runs them as one batch, so
 *    requests that arrive while other batches are on the device are admitted
 *    as lanes and memory free up, and their work overlaps on the device.
 *  - Admission charges each request its predicted HBM footprint
 *    (gpv_predict_footprint: Level1 bit-grid plus the estimated Level2
 *    boundary count, plus transient bins) against hbm_budget; once the batch
 *    has run, the charge drops to the bytes the resident result really holds,
 *    and it is returned by gpv_server_release.  A request larger than the
 *    whole budget is admitted alone when nothing else is charged.
 *  - A batch that fails (device memory short of the prediction, a per-run
 *    limit) is not lost: its requests go back to the front of the queue and
 *    each runs alone, when no other batch is on the device; only a request
 *    that fails alone is reported failed.  This is the serving analogue of
 *    the reference's rerun with a larger bin buffer (src/Object.cpp:3198-3214).
 *  - The caller keeps each request's soup valid until the request is done.
 */
#ifndef GPV_SERVE_H
#define GPV_SERVE_H

#include <stdint.h>

#include "gpv.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gpv_server gpv_server;

typedef struct gpv_server_opts {
    int device;               /* -1: current device */
    int lanes;                /* batches in flight at once; <= 0 picks 2 */
    int64_t hbm_budget;       /* admission budget in bytes; <= 0 picks 200e9 */
    int max_batch_meshes;     /* <= 0 picks 256 */
    int64_t max_batch_tris;   /* <= 0 picks 2^27 */
    int64_t lane_mem_limit;   /* > 0: each lane's context may have at most this many
                                 device bytes in use (gpv_ctx_set_mem_limit) */
} gpv_server_opts;

typedef struct gpv_request_stats {
    double t_submit, t_admit, t_done;   /* seconds on the server's clock */
    int64_t predicted_bytes;            /* admission charge while running */
    int64_t resident_bytes;             /* device bytes the result holds */
    int32_t batch_id, batch_size, lane;
    int32_t status;                     /* GPV_OK or the batch's error */
    int32_t retries;                    /* 1 if it was re-admitted after its batch failed */
    int64_t batch_predicted_bytes;      /* sum of the predicted bytes of its batch */
    int64_t batch_peak_bytes;           /* measured: the most bytes its batch had requested
                                           from the lane's pool at once, above what the lane
                                           held when it started */
} gpv_request_stats;

/* Predicted device bytes of voxelizing one mesh: every buffer its run
 * requests, transient ones included, with the run's own bin and hit
 * capacities and a boundary estimate of twice the voxels on the grid box's
 * faces.  An upper bound of the run's peak of requested bytes unless the
 * bins regrow or the surface has more boundary voxels than that. */
int gpv_predict_footprint(const gpv_mesh_desc *mesh, int64_t *bytes);

int gpv_server_create(const gpv_server_opts *opts, gpv_server **out);
/* Stops accepting, finishes queued work, releases every result. */
int gpv_server_destroy(gpv_server *srv);
/* Queues one request; *id identifies it.  tris_on_device as in gpv.h. */
int gpv_server_submit(gpv_server *srv, const gpv_mesh_desc *mesh, int tris_on_device, uint64_t *id);
/* Waits up to timeout_s (< 0: no limit) for request id.  On success *grid
 * and *ctx are the resident result and the context to pass with it to the
 * result calls of gpv.h (get_info, download, save_raw, device_views);
 * do not release it with gpv_grid_release.  Returns GPV_ERR_STATE on time-out
 * and the batch's status if it failed. */
int gpv_server_wait(gpv_server *srv, uint64_t id, double timeout_s, gpv_grid **grid, gpv_ctx **ctx,
                    gpv_request_stats *stats);
/* Frees request id's result and returns its charge to the budget. */
int gpv_server_release(gpv_server *srv, uint64_t id);
/* Totals: requests completed, batches run, bytes charged now, peak charge. */
int gpv_server_counters(gpv_server *srv, int64_t *completed, int64_t *batches, int64_t *charged,
                        int64_t *peak_charged);
/* Requests re-admitted to run alone after their batch failed. */
int gpv_server_readmitted(gpv_server *srv, int64_t *readmitted);
/* The lanes' device memory pools, summed (gpv_ctx_pool_stats per lane). */
int gpv_server_pool_stats(gpv_server *srv, gpv_pool_stats *out);

#ifdef __cplusplus
}
#endif
#endif /* GPV_SERVE_H */
