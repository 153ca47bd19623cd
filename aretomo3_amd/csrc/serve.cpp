// serve.cpp -- the serving loop: a device-resident request scheduler that
// reconstructs tilt series as they arrive.
//
// A request is one tilt series (SART, optionally also WBP).  Requests wait in
// arrival order and are admitted while their device bytes fit the HBM budget
// (atr_estimate_device_bytes of their plans + input/output buffers).  Every
// admitted series keeps its projections, sinogram and volume on the GPU from
// pre-stage to output.  The scheduler thread advances all admitted SART series
// together in rounds: one persistent forward-projection launch and one
// persistent back-projection launch per round carry one SART step of every
// series (continuous batching: a series admitted during the run joins at the
// next round with its initial back-projection, a finished one leaves).  A
// series computes exactly what it computes alone: the same kernels, the same
// operation order, nothing shared across series.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "atr_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

constexpr int kRing = 4;   // rounds whose item arrays may be in flight at once

struct Request {
    uint64_t id = 0;
    atr_series_desc desc{};
    std::vector<float> tilts, axes, shifts, local;
    std::vector<unsigned char> excl;
    int with_wbp = 0;
    const float* proj_host = nullptr;
    float* vol_host = nullptr;
    float* wbp_host = nullptr;
    size_t in_bytes = 0, out_bytes = 0, wbp_bytes = 0, need = 0;
    Clock::time_point submit, admit, done;
    atr_status status = ATR_OK;
    std::string error;
    bool finished = false;
    // device state while admitted
    atr_plan* sart = nullptr;
    atr_plan* wbp = nullptr;
    float* d_in = nullptr;
    float* d_out = nullptr;
    float* d_wbp = nullptr;
    int step = 0, last = 0;      // SART steps: 0 initial BP, 1..last iterations
    hipEvent_t ev_done = nullptr;
};

}  // namespace

struct atr_server {
    atr_server_config cfg{};
    size_t budget = 0, used = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> pending;
    std::map<uint64_t, std::shared_ptr<Request>> all;
    std::vector<std::shared_ptr<Request>> active;   // scheduler thread only
    std::vector<std::shared_ptr<Request>> draining; // scheduler thread only
    uint64_t next_id = 1;
    bool stop = false;
    std::thread worker;
    Clock::time_point t0;
    int device = 0;
    hipStream_t stream = nullptr;
    // per ring slot: item arrays (device + pinned host) and queue counters
    int cap = 0;
    atr::BpItem* d_bp[kRing] = {};
    atr::FpItem* d_fp[kRing] = {};
    atr::BpItem* h_bp[kRing] = {};
    atr::FpItem* h_fp[kRing] = {};
    int* d_cnt[kRing] = {};
    hipEvent_t ev_slot[kRing] = {};
    bool slot_used[kRing] = {};
    int slot = 0;
    int bp_grid = 1, fp_grid = 1, storage = 0;
    size_t fp_shm_max = 0;
    // statistics
    uint64_t rounds = 0, max_batch = 0;
};

namespace atr {

static double ms_since(const atr_server* s, Clock::time_point t)
{
    return std::chrono::duration<double, std::milli>(t - s->t0).count();
}

static void release(atr_server* s, Request& r)
{
    if (r.sart) atr_plan_destroy(r.sart);
    if (r.wbp) atr_plan_destroy(r.wbp);
    if (r.d_in) (void)hipFree(r.d_in);
    if (r.d_out) (void)hipFree(r.d_out);
    if (r.d_wbp) (void)hipFree(r.d_wbp);
    if (r.ev_done) (void)hipEventDestroy(r.ev_done);
    r.sart = r.wbp = nullptr;
    r.d_in = r.d_out = r.d_wbp = nullptr;
    r.ev_done = nullptr;
    s->used -= r.need;
}

static void finish_request(atr_server* s, Request& r, atr_status st, const std::string& err)
{
    std::lock_guard<std::mutex> lk(s->mu);
    r.status = st;
    r.error = err;
    r.done = Clock::now();
    r.finished = true;
    s->cv_done.notify_all();
}

// Create the plans and buffers of a request and enqueue its pre-stage,
// weighting and (if asked) its whole WBP reconstruction.
static atr_status admit(atr_server* s, Request& r)
{
    atr_status st = atr_plan_create(&r.desc, &r.sart);
    if (st != ATR_OK) return st;
    if (r.with_wbp) {
        atr_series_desc w = r.desc;
        w.method = ATR_METHOD_WBP;
        if ((st = atr_plan_create(&w, &r.wbp)) != ATR_OK) return st;
    }
    ATR_HIP(hipMalloc((void**)&r.d_in, r.in_bytes));
    ATR_HIP(hipMalloc((void**)&r.d_out, r.out_bytes));
    if (r.with_wbp) ATR_HIP(hipMalloc((void**)&r.d_wbp, r.wbp_bytes));
    ATR_HIP(hipEventCreateWithFlags(&r.ev_done, hipEventDisableTiming));
    ATR_HIP(hipMemcpyAsync(r.d_in, r.proj_host, r.in_bytes, hipMemcpyHostToDevice, s->stream));
    Plan& p = r.sart->p;
    sart_prepare(p, r.d_in, r.d_out, 0, s->stream);
    if (r.with_wbp && (st = atr_execute(r.wbp, r.d_in, r.d_wbp, 0, s->stream)) != ATR_OK) return st;
    r.step = p.all_tilts.empty() ? 1 : 0;
    r.last = p.iters * p.n_subsets;
    return ATR_OK;
}

static atr_status grow_slots(atr_server* s, int need)
{
    if (need <= s->cap) return ATR_OK;
    int cap = std::max(need, 2 * s->cap);
    (void)hipStreamSynchronize(s->stream);
    for (int k = 0; k < kRing; k++) {
        if (s->d_bp[k]) (void)hipFree(s->d_bp[k]);
        if (s->d_fp[k]) (void)hipFree(s->d_fp[k]);
        if (s->h_bp[k]) (void)hipHostFree(s->h_bp[k]);
        if (s->h_fp[k]) (void)hipHostFree(s->h_fp[k]);
        ATR_HIP(hipMalloc((void**)&s->d_bp[k], sizeof(BpItem) * cap));
        ATR_HIP(hipMalloc((void**)&s->d_fp[k], sizeof(FpItem) * cap));
        ATR_HIP(hipHostMalloc((void**)&s->h_bp[k], sizeof(BpItem) * cap, hipHostMallocDefault));
        ATR_HIP(hipHostMalloc((void**)&s->h_fp[k], sizeof(FpItem) * cap, hipHostMallocDefault));
        s->slot_used[k] = false;
    }
    s->cap = cap;
    return ATR_OK;
}

// One round: one SART step of every admitted series.
static atr_status run_round(atr_server* s)
{
    auto& act = s->active;
    ATR_HIP(hipSetDevice(s->device));
    atr_status st = grow_slots(s, (int)act.size());
    if (st != ATR_OK) return st;
    const int k = s->slot;
    s->slot = (s->slot + 1) % kRing;
    if (s->slot_used[k]) ATR_HIP(hipEventSynchronize(s->ev_slot[k]));
    int n_fp = 0, n_bp = 0, tot_fp = 0, tot_bp = 0;
    size_t shm = 0;
    for (auto& rp : act) {
        Request& r = *rp;
        const Plan& p = r.sart->p;
        const int n_all = (int)p.all_tilts.size();
        if (r.step == 0) {
            BpItem e{};
            e.a = bp_args(p, p.d_sino, p.n, p.d_lists, n_all, 1.0f, true);
            e.item0 = tot_bp;
            e.n_items_y = e.a.n_ztiles * e.a.n_xtiles;
            e.shape = bp_shape(p.bp_tz, p.bp_chunk);
            tot_bp += p.vol_y * e.n_items_y;
            s->h_bp[k][n_bp++] = e;
        } else if (r.step <= r.last) {
            const int sb = (r.step - 1) % p.n_subsets;
            const int f0 = p.subset_fp_off[sb], f1 = p.subset_fp_off[sb + 1];
            const int b0 = p.subset_off[sb], b1 = p.subset_off[sb + 1];
            FpItem f{};
            f.a = fp_args(p, p.d_fp_lists + f0, f1 - f0, p.d_fp_pairs + p.subset_pair_off[sb]);
            f.item0 = tot_fp;
            f.n_groups = (f1 - f0 + f.a.group - 1) / f.a.group;
            tot_fp += p.vol_y * f.n_groups;
            s->h_fp[k][n_fp++] = f;
            shm = std::max(shm, fp_lds_bytes(p.geo.fp_bz, p.vol_x));
            if (b1 > b0) {
                BpItem e{};
                e.a = bp_args(p, p.d_diff, p.max_subset, p.d_lists + n_all + b0, b1 - b0, p.relax,
                              false);
                e.item0 = tot_bp;
                e.n_items_y = e.a.n_ztiles * e.a.n_xtiles;
                e.shape = bp_shape(p.bp_tz, p.bp_chunk);
                tot_bp += p.vol_y * e.n_items_y;
                s->h_bp[k][n_bp++] = e;
            }
        }
    }
    hipStream_t q = s->stream;
    launch_zero_int(s->d_cnt[k], 3, q);
    if (n_fp) {
        ATR_HIP(hipMemcpyAsync(s->d_fp[k], s->h_fp[k], sizeof(FpItem) * n_fp,
                               hipMemcpyHostToDevice, q));
        launch_fp_batch(s->storage, s->d_fp[k], n_fp, tot_fp, s->d_cnt[k], shm,
                        std::min(s->fp_grid, tot_fp), q);
    }
    if (n_bp) {
        ATR_HIP(hipMemcpyAsync(s->d_bp[k], s->h_bp[k], sizeof(BpItem) * n_bp,
                               hipMemcpyHostToDevice, q));
        launch_bp_batch(s->storage, s->d_bp[k], n_bp, tot_bp, s->d_cnt[k] + 2,
                        std::min(s->bp_grid, tot_bp), q);
    }
    ATR_HIP(hipEventRecord(s->ev_slot[k], q));
    s->slot_used[k] = true;
    s->rounds++;
    s->max_batch = std::max<uint64_t>(s->max_batch, act.size());
    // advance; series past their last step are finished and copied out
    std::vector<std::shared_ptr<Request>> keep;
    for (auto& rp : act) {
        Request& r = *rp;
        r.step++;
        if (r.step > r.last) {
            sart_finish(r.sart->p, r.d_out, q);
            ATR_HIP(hipMemcpyAsync(r.vol_host, r.d_out, r.out_bytes, hipMemcpyDeviceToHost, q));
            if (r.with_wbp)
                ATR_HIP(hipMemcpyAsync(r.wbp_host, r.d_wbp, r.wbp_bytes, hipMemcpyDeviceToHost, q));
            ATR_HIP(hipEventRecord(r.ev_done, q));
            s->draining.push_back(rp);
        } else {
            keep.push_back(rp);
        }
    }
    act.swap(keep);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "serving round");
    return ATR_OK;
}

static void poll_done(atr_server* s, bool block)
{
    std::vector<std::shared_ptr<Request>> keep;
    for (auto& rp : s->draining) {
        Request& r = *rp;
        hipError_t e = block ? hipEventSynchronize(r.ev_done) : hipEventQuery(r.ev_done);
        if (e == hipErrorNotReady) {
            keep.push_back(rp);
            continue;
        }
        finish_request(s, r, e == hipSuccess ? ATR_OK : ATR_ERR_HIP,
                       e == hipSuccess ? "" : hipGetErrorString(e));
        std::lock_guard<std::mutex> lk(s->mu);
        release(s, r);
        s->cv_work.notify_all();
    }
    s->draining.swap(keep);
}

static void worker_main(atr_server* s)
{
    (void)hipSetDevice(s->device);
    for (;;) {
        // admission in arrival order while the budget allows
        std::vector<std::shared_ptr<Request>> admitted;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] {
                return s->stop || !s->active.empty() || !s->draining.empty() ||
                       (!s->pending.empty() && s->used + s->pending.front()->need <= s->budget);
            });
            if (s->stop && s->pending.empty() && s->active.empty() && s->draining.empty()) return;
            while (!s->pending.empty() && s->used + s->pending.front()->need <= s->budget &&
                   (s->cfg.max_active <= 0 || (int)(s->active.size() + admitted.size()) < s->cfg.max_active)) {
                auto rp = s->pending.front();
                s->pending.pop_front();
                s->used += rp->need;
                admitted.push_back(rp);
            }
        }
        for (auto& rp : admitted) {
            rp->admit = Clock::now();
            atr_status st = admit(s, *rp);
            if (st != ATR_OK) {
                std::string err = atr_last_error_string();
                (void)hipStreamSynchronize(s->stream);
                finish_request(s, *rp, st, err);
                std::lock_guard<std::mutex> lk(s->mu);
                release(s, *rp);
                continue;
            }
            s->active.push_back(rp);
        }
        if (!s->active.empty()) {
            atr_status st = run_round(s);
            if (st != ATR_OK) {
                std::string err = atr_last_error_string();
                (void)hipStreamSynchronize(s->stream);
                for (auto& rp : s->active) {
                    finish_request(s, *rp, st, err);
                    std::lock_guard<std::mutex> lk(s->mu);
                    release(s, *rp);
                }
                s->active.clear();
            }
            poll_done(s, false);
        } else {
            poll_done(s, true);
        }
    }
}

}  // namespace atr

using namespace atr;

extern "C" {

void atr_server_config_init(atr_server_config* c)
{
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->max_active = 4;   // a few series overlap; later arrivals queue (lower p50 than all-at-once)
    c->storage = ATR_STORAGE_F16;
}

atr_status atr_server_create(const atr_server_config* cfg, atr_server** out)
{
    if (!cfg || !out) return fail(ATR_ERR_INVALID_ARG, "atr_server_create: NULL argument");
    if (cfg->storage < ATR_STORAGE_F32 || cfg->storage > ATR_STORAGE_BF16)
        return fail(ATR_ERR_INVALID_ARG, "atr_server_create: unknown storage");
    *out = nullptr;
    auto* s = new atr_server;
    s->cfg = *cfg;
    s->storage = cfg->storage;
    s->t0 = Clock::now();
    auto bail = [&](atr_status st) {
        if (s->stream) (void)hipStreamDestroy(s->stream);
        for (int k = 0; k < kRing; k++) {
            if (s->d_cnt[k]) (void)hipFree(s->d_cnt[k]);
            if (s->ev_slot[k]) (void)hipEventDestroy(s->ev_slot[k]);
        }
        delete s;
        return st;
    };
    if (hipGetDevice(&s->device) != hipSuccess) return bail(hip_fail(hipErrorNoDevice, "hipGetDevice"));
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return bail(fail(ATR_ERR_HIP, "hipMemGetInfo"));
    const size_t cap = free_b - free_b / 10;
    s->budget = cfg->hbm_budget_bytes ? std::min(cfg->hbm_budget_bytes, cap) : cap;
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(ATR_ERR_HIP, "hipStreamCreate"));
    for (int k = 0; k < kRing; k++) {
        if (hipMalloc((void**)&s->d_cnt[k], 4 * sizeof(int)) != hipSuccess ||
            hipEventCreateWithFlags(&s->ev_slot[k], hipEventDisableTiming) != hipSuccess)
            return bail(fail(ATR_ERR_OUT_OF_MEMORY, "atr_server_create: ring"));
    }
    int cus = 1;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->device);
    s->bp_grid = cus * batch_blocks_per_cu(s->storage, false, 0);
    s->fp_grid = cus * batch_blocks_per_cu(s->storage, true, FP_LDS_BUDGET);
    s->worker = std::thread(worker_main, s);
    *out = s;
    return ATR_OK;
}

atr_status atr_server_submit(atr_server* s, const atr_series_desc* desc, int with_wbp,
                             const float* proj_host, float* vol_host, float* wbp_host,
                             uint64_t* id)
{
    if (!s || !desc || !proj_host || !vol_host || !id || (with_wbp && !wbp_host))
        return fail(ATR_ERR_INVALID_ARG, "atr_server_submit: NULL argument");
    if (desc->method != ATR_METHOD_SART)
        return fail(ATR_ERR_INVALID_ARG, "atr_server_submit: requests are SART series (WBP on request)");
    auto r = std::make_shared<Request>();
    r->desc = *desc;
    r->desc.storage = s->storage;
    const int n = desc->n_tilts;
    if (n < 1 || !desc->tilt_deg) return fail(ATR_ERR_INVALID_ARG, "atr_server_submit: no tilts");
    // the descriptor's arrays are copied: the caller may free them after submit
    r->tilts.assign(desc->tilt_deg, desc->tilt_deg + n);
    r->desc.tilt_deg = r->tilts.data();
    if (desc->tilt_axis_deg) {
        r->axes.assign(desc->tilt_axis_deg, desc->tilt_axis_deg + n);
        r->desc.tilt_axis_deg = r->axes.data();
    }
    if (desc->shift_xy) {
        r->shifts.assign(desc->shift_xy, desc->shift_xy + 2 * n);
        r->desc.shift_xy = r->shifts.data();
    }
    if (desc->exclude) {
        r->excl.assign(desc->exclude, desc->exclude + n);
        r->desc.exclude = r->excl.data();
    }
    if (desc->n_patches > 0 && desc->local_params) {
        r->local.assign(desc->local_params, desc->local_params + (size_t)n * 5 * desc->n_patches);
        r->desc.local_params = r->local.data();
    }
    size_t bytes_sart = 0, bytes_wbp = 0;
    atr_status st = atr_estimate_device_bytes(&r->desc, &bytes_sart);
    if (st != ATR_OK) return st;
    atr_series_desc w = r->desc;
    w.method = ATR_METHOD_WBP;
    if (with_wbp && (st = atr_estimate_device_bytes(&w, &bytes_wbp)) != ATR_OK) return st;
    // volume size from the same rules as the plan
    const size_t row = desc->row_stride ? desc->row_stride : (size_t)desc->proj_x;
    const size_t tstride = desc->tilt_stride ? desc->tilt_stride : row * desc->proj_y;
    r->in_bytes = sizeof(float) * (tstride * (n - 1) + row * (desc->proj_y - 1) + desc->proj_x);
    int vx = 0, vy = 0, vz = 0;
    if ((st = desc_volume(&r->desc, &vx, &vy, &vz)) != ATR_OK) return st;
    r->out_bytes = sizeof(float) * (size_t)vx * vy * vz;
    r->wbp_bytes = with_wbp ? r->out_bytes : 0;
    r->need = bytes_sart + bytes_wbp + r->in_bytes + r->out_bytes + r->wbp_bytes;
    r->with_wbp = with_wbp;
    r->proj_host = proj_host;
    r->vol_host = vol_host;
    r->wbp_host = wbp_host;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (r->need > s->budget)
            return fail(ATR_ERR_OUT_OF_MEMORY, "atr_server_submit: request larger than the HBM budget");
        r->id = s->next_id++;
        r->submit = Clock::now();
        s->pending.push_back(r);
        s->all[r->id] = r;
        *id = r->id;
    }
    s->cv_work.notify_all();
    return ATR_OK;
}

atr_status atr_server_wait(atr_server* s, uint64_t id, atr_request_info* info)
{
    if (!s) return fail(ATR_ERR_INVALID_ARG, "server is NULL");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) return fail(ATR_ERR_INVALID_ARG, "atr_server_wait: unknown request id");
    auto r = it->second;
    s->cv_done.wait(lk, [&] { return r->finished; });
    if (info) {
        info->submit_ms = ms_since(s, r->submit);
        info->admit_ms = ms_since(s, r->admit);
        info->done_ms = ms_since(s, r->done);
        info->latency_ms = info->done_ms - info->submit_ms;
        info->status = r->status;
    }
    s->all.erase(it);
    if (r->status != ATR_OK) return fail(r->status, "request failed: " + r->error);
    return ATR_OK;
}

atr_status atr_server_stats(atr_server* s, uint64_t* rounds, uint64_t* max_batch)
{
    if (!s) return fail(ATR_ERR_INVALID_ARG, "server is NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    if (rounds) *rounds = s->rounds;
    if (max_batch) *max_batch = s->max_batch;
    return ATR_OK;
}

void atr_server_destroy(atr_server* s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    (void)hipStreamSynchronize(s->stream);
    for (int k = 0; k < kRing; k++) {
        if (s->d_bp[k]) (void)hipFree(s->d_bp[k]);
        if (s->d_fp[k]) (void)hipFree(s->d_fp[k]);
        if (s->h_bp[k]) (void)hipHostFree(s->h_bp[k]);
        if (s->h_fp[k]) (void)hipHostFree(s->h_fp[k]);
        if (s->d_cnt[k]) (void)hipFree(s->d_cnt[k]);
        if (s->ev_slot[k]) (void)hipEventDestroy(s->ev_slot[k]);
    }
    (void)hipStreamDestroy(s->stream);
    delete s;
}

atr_status atr_host_alloc(void** ptr, size_t bytes)
{
    if (!ptr) return fail(ATR_ERR_INVALID_ARG, "ptr is NULL");
    ATR_HIP(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
    return ATR_OK;
}

atr_status atr_host_free(void* ptr)
{
    ATR_HIP(hipHostFree(ptr));
    return ATR_OK;
}

}  // extern "C"
