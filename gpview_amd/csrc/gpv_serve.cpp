// gpv_serve.cpp -- continuous-batching request loop over gpv_voxelize_batch
// (include/gpv_serve.h).  No counterpart in the reference, which voxelizes
// one Object per keypress (src/GPView.cpp:1378-1380).
#include "gpv_serve.h"

#include "gpv_internal.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

enum ReqState { kPending = 0, kRunning = 1, kDone = 2, kFailed = 3 };

struct Req {
    uint64_t id = 0;
    gpv_mesh_desc d{};
    int on_device = 0;
    int64_t predicted = 0;
    int64_t voxels = 0;         // padded Level1 voxels (batch limit)
    int64_t resident = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    int state = kPending;
    int status = GPV_OK;
    std::string err;
    gpv_grid* grid = nullptr;
    int lane = -1, batch_id = -1, batch_size = 0;
    int retries = 0;                 // re-admissions after a failed batch
    int64_t batch_predicted = 0;     // predicted bytes of its batch
    int64_t batch_peak = 0;          // measured peak pool bytes of its batch
};

int err(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    return gpv::set_error(code, buf);
}

// device bytes a finished grid keeps: Level1 boundary + inside planes and the
// word prefix, the boundary list and both Level2 planes
int64_t resident_bytes(const gpv_grid_info& I)
{
    const int64_t nxw = (I.n[0] + 63) / 64;
    const int64_t W = (int64_t)I.n[2] * I.n[1] * nxw;
    const int64_t w2 = I.n2 > 0 ? ((int64_t)I.n2 * I.n2 * I.n2 + 63) / 64 : 0;
    const int64_t tables = 4 * (2 * ((int64_t)I.n[0] + I.n[1] + I.n[2]) * (1 + (I.n2 > 0 ? I.n2 : 0)));
    return W * 20 + I.nb * (4 + 16 * w2) + tables;
}

}  // namespace

struct gpv_server {
    gpv_server_opts o{};
    std::vector<gpv_ctx*> ctx;
    std::vector<std::thread> th;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req*> pending;
    std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;
    uint64_t next_id = 1;
    int64_t charged = 0, peak = 0, completed = 0, batches = 0, readmitted = 0;
    int running = 0;   // batches on the device now
    bool stop = false;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    void lane_main(int lane);
};

void gpv_server::lane_main(int lane)
{
    gpv_ctx* c = ctx[(size_t)lane];
    std::vector<Req*> batch;
    std::vector<gpv_mesh_desc> descs;
    std::vector<gpv_grid*> grids;
    for (;;) {
        int flag = 0, bid = 0;
        {
            std::unique_lock<std::mutex> lk(mu);
            for (;;) {
                if (stop && pending.empty()) return;
                batch.clear();
                if (!pending.empty() && pending.front()->retries > 0) {
                    // a request of a failed batch runs alone, with no other
                    // batch on the device, so it has the most memory there is
                    if (running == 0) {
                        Req* r = pending.front();
                        flag = r->on_device;
                        charged += r->predicted;
                        peak = std::max(peak, charged);
                        r->t_admit = now();
                        r->state = kRunning;
                        r->lane = lane;
                        batch.push_back(r);
                        pending.pop_front();
                    }
                } else if (!pending.empty()) {
                    // arrival order; a request that does not fit ends the batch
                    flag = pending.front()->on_device;
                    int64_t tris = 0, vox = 0;
                    while (!pending.empty()) {
                        Req* r = pending.front();
                        if (r->on_device != flag || (int)batch.size() >= o.max_batch_meshes || r->retries > 0) break;
                        if (!batch.empty() &&
                            (tris + r->d.ntris > o.max_batch_tris || vox + r->voxels >= (int64_t(1) << 32)))
                            break;
                        const bool fits = charged + r->predicted <= o.hbm_budget || (charged == 0 && batch.empty());
                        if (!fits) break;
                        charged += r->predicted;
                        peak = std::max(peak, charged);
                        tris += r->d.ntris;
                        vox += r->voxels;
                        r->t_admit = now();
                        r->state = kRunning;
                        r->lane = lane;
                        batch.push_back(r);
                        pending.pop_front();
                    }
                }
                if (!batch.empty()) break;
                cv_work.wait(lk);
            }
            bid = (int)batches++;
            ++running;
        }
        const int n = (int)batch.size();
        descs.resize((size_t)n);
        grids.assign((size_t)n, nullptr);
        int64_t predicted = 0;
        for (int i = 0; i < n; ++i) {
            descs[(size_t)i] = batch[(size_t)i]->d;
            predicted += batch[(size_t)i]->predicted;
        }
        // the batch's own peak: bytes requested from the pool above what the
        // lane held when it started (earlier results stay resident until
        // released), the quantity gpv_predict_footprint bounds
        gpv_pool_stats p0{}, p1{};
        gpv::pool_reset_peak(c);
        gpv_ctx_pool_stats(c, &p0);
        const int st = gpv_voxelize_batch(c, descs.data(), n, flag, grids.data());
        gpv_ctx_pool_stats(c, &p1);
        const int64_t measured = p1.peak_requested_bytes - p0.requested_bytes;
        const std::string msg = st == GPV_OK ? std::string() : std::string(gpv_last_error());
        std::vector<int64_t> res((size_t)n, 0);
        if (st == GPV_OK)
            for (int i = 0; i < n; ++i) {
                gpv_grid_info I;
                gpv_grid_get_info(grids[(size_t)i], &I);
                res[(size_t)i] = resident_bytes(I);
            }
        const double t = now();
        {
            std::lock_guard<std::mutex> lk(mu);
            --running;
            for (int i = n - 1; i >= 0; --i) {
                Req* r = batch[(size_t)i];
                r->batch_id = bid;
                r->batch_size = n;
                r->batch_predicted = predicted;
                r->batch_peak = measured;
                charged -= r->predicted;
                if (st == GPV_OK) {
                    r->t_done = t;
                    r->status = st;
                    r->grid = grids[(size_t)i];
                    r->resident = res[(size_t)i];
                    charged += r->resident;
                    r->state = kDone;
                    ++completed;
                } else if (r->retries == 0) {
                    // re-admitted at the front, in arrival order, one at a time
                    r->retries = 1;
                    r->state = kPending;
                    r->lane = -1;
                    pending.push_front(r);
                    ++readmitted;
                } else {
                    r->t_done = t;
                    r->status = st;
                    r->err = msg;
                    r->state = kFailed;
                    ++completed;
                }
            }
        }
        cv_done.notify_all();
        cv_work.notify_all();
    }
}

extern "C" {

int gpv_server_create(const gpv_server_opts* opts, gpv_server** out)
{
    if (!out) return err(GPV_ERR_INVALID, "out is NULL");
    *out = nullptr;
    gpv_server* S = new gpv_server();
    if (opts) S->o = *opts;
    else S->o.device = -1;
    if (S->o.lanes <= 0) S->o.lanes = 2;
    if (S->o.hbm_budget <= 0) S->o.hbm_budget = 200000000000LL;
    if (S->o.max_batch_meshes <= 0) S->o.max_batch_meshes = 256;
    if (S->o.max_batch_tris <= 0) S->o.max_batch_tris = int64_t(1) << 27;
    for (int i = 0; i < S->o.lanes; ++i) {
        gpv_ctx* c = nullptr;
        int st = gpv_ctx_create(S->o.device, 0, &c);
        if (st == GPV_OK && S->o.lane_mem_limit > 0) {
            st = gpv_ctx_set_mem_limit(c, S->o.lane_mem_limit);
            if (st != GPV_OK) gpv_ctx_destroy(c);
        }
        if (st != GPV_OK) {
            for (gpv_ctx* p : S->ctx) gpv_ctx_destroy(p);
            delete S;
            return st;
        }
        S->ctx.push_back(c);
    }
    for (int i = 0; i < S->o.lanes; ++i) S->th.emplace_back([S, i] { S->lane_main(i); });
    *out = S;
    return GPV_OK;
}

int gpv_server_destroy(gpv_server* S)
{
    if (!S) return GPV_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    for (auto& t : S->th) t.join();
    for (auto& kv : S->reqs)
        if (kv.second->grid) gpv_grid_release(S->ctx[(size_t)kv.second->lane], kv.second->grid);
    for (gpv_ctx* c : S->ctx) gpv_ctx_destroy(c);
    delete S;
    return GPV_OK;
}

int gpv_server_submit(gpv_server* S, const gpv_mesh_desc* d, int on_device, uint64_t* id)
{
    if (!S || !d || !id) return err(GPV_ERR_INVALID, "NULL argument");
    if (d->ntris < 0 || (d->ntris > 0 && !d->tris)) return err(GPV_ERR_INVALID, "bad triangle soup");
    std::unique_ptr<Req> r(new Req());
    r->d = *d;
    r->on_device = on_device ? 1 : 0;
    const int st = gpv_predict_footprint(d, &r->predicted);
    if (st != GPV_OK) return st;
    gpv_grid_info I;
    gpv_grid_derive(d->bmin, d->bmax, d->max_model_size, &d->params, &I);
    r->voxels = (int64_t)I.n[2] * I.n[1] * ((I.n[0] + 63) / 64) * 64;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stop) return err(GPV_ERR_STATE, "server is shutting down");
        r->id = S->next_id++;
        r->t_submit = S->now();
        *id = r->id;
        S->pending.push_back(r.get());
        S->reqs.emplace(r->id, std::move(r));
    }
    S->cv_work.notify_all();
    return GPV_OK;
}

int gpv_server_wait(gpv_server* S, uint64_t id, double timeout_s, gpv_grid** grid, gpv_ctx** ctx,
                    gpv_request_stats* stats)
{
    if (!S) return err(GPV_ERR_INVALID, "NULL server");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return err(GPV_ERR_INVALID, "unknown request %llu", (unsigned long long)id);
    Req* r = it->second.get();
    auto done = [r] { return r->state >= kDone; };
    if (timeout_s < 0) {
        S->cv_done.wait(lk, done);
    } else if (!S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done)) {
        return err(GPV_ERR_STATE, "request %llu not done after %g s", (unsigned long long)id, timeout_s);
    }
    if (stats) {
        stats->t_submit = r->t_submit;
        stats->t_admit = r->t_admit;
        stats->t_done = r->t_done;
        stats->predicted_bytes = r->predicted;
        stats->resident_bytes = r->resident;
        stats->batch_id = r->batch_id;
        stats->batch_size = r->batch_size;
        stats->lane = r->lane;
        stats->status = r->status;
        stats->retries = r->retries;
        stats->batch_predicted_bytes = r->batch_predicted;
        stats->batch_peak_bytes = r->batch_peak;
    }
    if (r->state == kFailed) return err(r->status, "request %llu: %s", (unsigned long long)id, r->err.c_str());
    if (grid) *grid = r->grid;
    if (ctx) *ctx = S->ctx[(size_t)r->lane];
    return GPV_OK;
}

int gpv_server_release(gpv_server* S, uint64_t id)
{
    if (!S) return err(GPV_ERR_INVALID, "NULL server");
    gpv_grid* g = nullptr;
    gpv_ctx* c = nullptr;
    int64_t charge = 0;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        auto it = S->reqs.find(id);
        if (it == S->reqs.end()) return err(GPV_ERR_INVALID, "unknown request %llu", (unsigned long long)id);
        Req* r = it->second.get();
        if (r->state < kDone) return err(GPV_ERR_STATE, "request %llu is still queued or running", (unsigned long long)id);
        g = r->grid;
        c = g ? S->ctx[(size_t)r->lane] : nullptr;
        charge = r->state == kDone ? r->resident : 0;
        S->reqs.erase(it);
    }
    const int st = g ? gpv_grid_release(c, g) : GPV_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->charged -= charge;
    }
    S->cv_work.notify_all();
    return st;
}

int gpv_server_counters(gpv_server* S, int64_t* completed, int64_t* batches, int64_t* charged, int64_t* peak)
{
    if (!S) return err(GPV_ERR_INVALID, "NULL server");
    std::lock_guard<std::mutex> lk(S->mu);
    if (completed) *completed = S->completed;
    if (batches) *batches = S->batches;
    if (charged) *charged = S->charged;
    if (peak) *peak = S->peak;
    return GPV_OK;
}

int gpv_server_readmitted(gpv_server* S, int64_t* readmitted)
{
    if (!S || !readmitted) return err(GPV_ERR_INVALID, "NULL argument");
    std::lock_guard<std::mutex> lk(S->mu);
    *readmitted = S->readmitted;
    return GPV_OK;
}

int gpv_server_pool_stats(gpv_server* S, gpv_pool_stats* out)
{
    if (!S || !out) return err(GPV_ERR_INVALID, "NULL argument");
    gpv_pool_stats sum{};
    for (gpv_ctx* c : S->ctx) {
        gpv_pool_stats p{};
        const int st = gpv_ctx_pool_stats(c, &p);
        if (st != GPV_OK) return st;
        sum.held_bytes += p.held_bytes;
        sum.cached_bytes += p.cached_bytes;
        sum.cap_bytes += p.cap_bytes;
        sum.allocations += p.allocations;
        sum.reuses += p.reuses;
        sum.evictions += p.evictions;
        sum.live_bytes += p.live_bytes;
        sum.peak_live_bytes += p.peak_live_bytes;
        sum.limit_bytes += p.limit_bytes;
        sum.requested_bytes += p.requested_bytes;
        sum.peak_requested_bytes += p.peak_requested_bytes;
    }
    *out = sum;
    return GPV_OK;
}

}  // extern "C"
