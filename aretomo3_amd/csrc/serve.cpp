// serve.cpp -- the serving loop: a device-resident request scheduler that
// reconstructs tilt series as they arrive.
//
// A request is one tilt series (SART, optionally also WBP).  Four host
// threads share the work:
//   * the submitting thread copies the descriptor and queues the request;
//   * the preparer builds each request's plans on the host (FP64 geometry,
//     fill tables, wave schedules: tens to hundreds of ms for large series)
//     in arrival order, ahead of admission;
//   * the scheduler admits prepared requests in arrival order while their
//     device bytes fit the HBM budget, and advances every admitted SART
//     series in rounds: one persistent forward-projection launch and one
//     persistent back-projection launch per round carry one SART step of each
//     (continuous batching: a series admitted mid-run joins at the next
//     round, a finished one leaves);
//   * the finisher copies finished volumes out and returns their memory.
// Device work runs on four streams: rounds (compute), WBP reconstructions
// (compute, each followed by the copy-out of its volume), host-to-device and
// device-to-host copies, ordered by events, so a request's input upload, its
// WBP and the copy-out of earlier volumes overlap the rounds.  Device memory comes from an arena that keeps what it
// took from HIP until the server is destroyed: admitting and releasing a
// request synchronises nothing.  A series computes exactly what it computes
// alone: the same kernels, the same operation order, nothing shared across
// series.
#include <algorithm>
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
// queue counters of a round: one per forward-projection variant, then one
// per back-projection tile shape
constexpr int kCounters = atr::FP_VARIANTS + atr::BP_SHAPES;
constexpr int kStarveRounds = 16;   // shortest-first: longest wait of a series, in rounds

// Device arena: chunks from hipMalloc, kept until the arena is destroyed;
// blocks handed out first-fit in address order and coalesced with their
// free neighbours (within one chunk) when returned.
class Arena : public atr::DevAlloc {
public:
    explicit Arena(size_t chunk) : chunk_(chunk) {}
    ~Arena() override
    {
        for (auto& c : chunks_) (void)hipFree(c.first);
    }

    hipError_t alloc(void** ptr, size_t bytes) override
    {
        const size_t b = round_up(std::max<size_t>(bytes, 1), kAlign);
        std::lock_guard<std::mutex> lk(mu_);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second.size >= b) {
                *ptr = take(it, b);
                return hipSuccess;
            }
        }
        const size_t size = std::max(chunk_, round_up(b, kChunkAlign));
        char* base = nullptr;
        hipError_t e = hipMalloc((void**)&base, size);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return e;
        }
        const int id = (int)chunks_.size();
        chunks_.push_back({base, size});
        reserved_ += size;
        auto it = free_.emplace(base, Block{size, id}).first;
        *ptr = take(it, b);
        return hipSuccess;
    }

    void release(void* ptr) override
    {
        std::lock_guard<std::mutex> lk(mu_);
        auto u = used_.find((char*)ptr);
        if (u == used_.end()) return;
        char* start = u->first;
        Block blk = u->second;
        used_.erase(u);
        auto next = free_.lower_bound(start);
        if (next != free_.end() && next->first == start + blk.size && next->second.chunk == blk.chunk) {
            blk.size += next->second.size;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second.size == start && prev->second.chunk == blk.chunk) {
                prev->second.size += blk.size;
                return;
            }
        }
        free_.emplace(start, blk);
    }

    size_t reserved()
    {
        std::lock_guard<std::mutex> lk(mu_);
        return reserved_;
    }

private:
    struct Block {
        size_t size;
        int chunk;
    };
    static constexpr size_t kAlign = 256;
    static constexpr size_t kChunkAlign = size_t(2) << 20;
    static size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

    char* take(std::map<char*, Block>::iterator it, size_t b)
    {
        char* start = it->first;
        Block blk = it->second;
        free_.erase(it);
        if (blk.size > b) free_.emplace(start + b, Block{blk.size - b, blk.chunk});
        used_.emplace(start, Block{b, blk.chunk});
        return start;
    }

    std::mutex mu_;
    size_t chunk_;
    size_t reserved_ = 0;
    std::vector<std::pair<char*, size_t>> chunks_;
    std::map<char*, Block> free_, used_;
};

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
    Clock::time_point submit, prepared, admit, compute_done, done;
    atr_status status = ATR_OK;
    std::string error;
    bool ready = false;      // plans prepared on the host (or failed: status)
    bool finished = false;
    // plans (host tables from the preparer, device memory from the arena)
    std::unique_ptr<atr::Plan> sart, wbp;
    float* d_in = nullptr;
    float* d_out = nullptr;
    float* d_wbp = nullptr;
    hipEvent_t ev_in = nullptr, ev_sart = nullptr, ev_wbp = nullptr, ev_done = nullptr;
    int step = 0, last = 0;  // SART steps: 0 initial BP, 1..last iterations
    double cost = 0.0;       // voxel-tilt updates of one SART step (fair-share rounds)
    double credit = 0.0;     // fair-share credit, in the same unit
    int skipped = 0;         // consecutive rounds without a step (shortest-first)
};

}  // namespace

struct atr_server {
    atr_server_config cfg{};
    size_t budget = 0, used = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_prep, cv_fin;
    std::deque<std::shared_ptr<Request>> pending;    // arrival order, until admitted
    std::deque<std::shared_ptr<Request>> to_prep;    // preparer queue
    std::deque<std::shared_ptr<Request>> to_finish;  // finisher queue
    std::map<uint64_t, std::shared_ptr<Request>> all;
    std::vector<std::shared_ptr<Request>> active;    // scheduler thread only
    int in_flight = 0;          // admitted, not yet released
    bool stop = false, sched_done = false;
    uint64_t next_id = 1;
    std::thread th_prep, th_sched, th_fin;
    Clock::time_point t0;
    int device = 0;
    hipStream_t s_comp = nullptr, s_wbp = nullptr, s_h2d = nullptr, s_d2h = nullptr;
    std::unique_ptr<Arena> arena;
    std::vector<hipEvent_t> ev_pool;                 // under mu
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
    int storage = 0;
    // statistics
    uint64_t rounds = 0, max_batch = 0;
};

namespace atr {

static double ms_since(const atr_server* s, Clock::time_point t)
{
    return std::chrono::duration<double, std::milli>(t - s->t0).count();
}

static hipEvent_t take_event(atr_server* s)
{
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (!s->ev_pool.empty()) {
            hipEvent_t e = s->ev_pool.back();
            s->ev_pool.pop_back();
            return e;
        }
    }
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
    return e;
}

// Returns the request's device memory and events.  Called only once every
// GPU operation of the request has completed (or none was issued), so
// nothing waits here.
static void release(atr_server* s, Request& r)
{
    if (r.sart) plan_release(*r.sart);
    if (r.wbp) plan_release(*r.wbp);
    r.sart.reset();
    r.wbp.reset();
    for (float* q : {r.d_in, r.d_out, r.d_wbp})
        if (q) s->arena->release(q);
    r.d_in = r.d_out = r.d_wbp = nullptr;
    std::lock_guard<std::mutex> lk(s->mu);
    for (hipEvent_t* e : {&r.ev_in, &r.ev_sart, &r.ev_wbp, &r.ev_done}) {
        if (*e) s->ev_pool.push_back(*e);
        *e = nullptr;
    }
}

// Marks the request done (waiters wake) and gives its budget back.
static void finish_request(atr_server* s, Request& r, bool was_admitted)
{
    std::lock_guard<std::mutex> lk(s->mu);
    r.done = Clock::now();
    r.finished = true;
    if (was_admitted) {
        s->used -= r.need;
        s->in_flight -= 1;
    }
    s->cv_done.notify_all();
    s->cv_work.notify_all();
    s->cv_fin.notify_all();
}

// ---- preparer: host tables of the request's plans ---------------------------
static void prep_main(atr_server* s)
{
    for (;;) {
        std::shared_ptr<Request> rp;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_prep.wait(lk, [&] { return s->sched_done || !s->to_prep.empty(); });
            if (s->to_prep.empty()) return;   // the scheduler has finished
            rp = s->to_prep.front();
            s->to_prep.pop_front();
        }
        Request& r = *rp;
        r.sart.reset(new Plan());
        atr_status st = plan_prepare(&r.desc, *r.sart);
        if (st == ATR_OK && r.with_wbp) {
            atr_series_desc w = r.desc;
            w.method = ATR_METHOD_WBP;
            r.wbp.reset(new Plan());
            st = plan_prepare(&w, *r.wbp);
        }
        std::lock_guard<std::mutex> lk(s->mu);
        if (st != ATR_OK) {
            r.status = st;
            r.error = atr_last_error_string();
            r.sart.reset();
            r.wbp.reset();
        }
        r.prepared = Clock::now();
        r.ready = true;
        s->cv_work.notify_all();
    }
}

// ---- scheduler -----------------------------------------------------------------
// Device half of admission: the plans' memory from the arena and their
// uploads, input/output buffers, the input upload on the H2D stream and the
// whole WBP on the WBP stream after it.  The SART series joins the rounds
// once the upload has landed (start_sart), so the round stream never waits
// on a copy.
static atr_status admit(atr_server* s, Request& r)
{
    atr_status st;
    r.sart->alloc = s->arena.get();
    if ((st = plan_upload(*r.sart)) != ATR_OK) return st;
    if (r.wbp) {
        r.wbp->alloc = s->arena.get();
        if ((st = plan_upload(*r.wbp)) != ATR_OK) return st;
    }
    ATR_HIP(s->arena->alloc((void**)&r.d_in, r.in_bytes));
    ATR_HIP(s->arena->alloc((void**)&r.d_out, r.out_bytes));
    if (r.wbp) ATR_HIP(s->arena->alloc((void**)&r.d_wbp, r.wbp_bytes));
    for (hipEvent_t* e : {&r.ev_in, &r.ev_sart, &r.ev_wbp, &r.ev_done})
        if ((*e = take_event(s)) == nullptr) return fail(ATR_ERR_HIP, "serving: event creation failed");
    ATR_HIP(hipMemcpyAsync(r.d_in, r.proj_host, r.in_bytes, hipMemcpyHostToDevice, s->s_h2d));
    ATR_HIP(hipEventRecord(r.ev_in, s->s_h2d));
    if (r.wbp) {
        // the WBP volume is final as soon as its kernels end: it is copied
        // out right behind them on the same stream, while the SART rounds
        // of the series still run (ev_wbp: WBP volume on the host)
        ATR_HIP(hipStreamWaitEvent(s->s_wbp, r.ev_in, 0));
        if ((st = plan_execute(*r.wbp, r.d_in, r.d_wbp, 0, s->s_wbp)) != ATR_OK) return st;
        ATR_HIP(hipMemcpyAsync(r.wbp_host, r.d_wbp, r.wbp_bytes, hipMemcpyDeviceToHost, s->s_wbp));
        ATR_HIP(hipEventRecord(r.ev_wbp, s->s_wbp));
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "serving admission");
    return ATR_OK;
}

// Pre-stage, weighting and the volume state of a series whose input is on
// the device, on the round stream; the series takes part from the next round.
static atr_status start_sart(atr_server* s, Request& r)
{
    ATR_HIP(hipStreamWaitEvent(s->s_comp, r.ev_in, 0));
    Plan& p = *r.sart;
    sart_prepare(p, r.d_in, r.d_out, 0, s->s_comp);
    r.step = p.all_tilts.empty() ? 1 : 0;
    r.last = p.iters * p.n_subsets;
    r.cost = (double)p.vol_x * p.vol_y * p.vol_z * ((double)p.n / std::max(1, p.n_subsets));
    r.credit = 0.0;
    r.skipped = 0;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "serving start");
    return ATR_OK;
}

static atr_status grow_slots(atr_server* s, int need)
{
    if (need <= s->cap) return ATR_OK;
    int cap = std::max(need, 2 * s->cap);
    ATR_HIP(hipStreamSynchronize(s->s_comp));
    for (int k = 0; k < kRing; k++) {
        if (s->d_bp[k]) (void)hipFree(s->d_bp[k]);
        if (s->d_fp[k]) (void)hipFree(s->d_fp[k]);
        if (s->h_bp[k]) (void)hipHostFree(s->h_bp[k]);
        if (s->h_fp[k]) (void)hipHostFree(s->h_fp[k]);
        s->d_bp[k] = nullptr;
        s->d_fp[k] = nullptr;
        s->h_bp[k] = nullptr;
        s->h_fp[k] = nullptr;
        ATR_HIP(hipMalloc((void**)&s->d_bp[k], sizeof(BpItem) * cap));
        // forward projection: each entry in its variant's queue and in the
        // empty-voxel queue
        ATR_HIP(hipMalloc((void**)&s->d_fp[k], sizeof(FpItem) * 2 * cap));
        ATR_HIP(hipHostMalloc((void**)&s->h_bp[k], sizeof(BpItem) * cap, hipHostMallocDefault));
        ATR_HIP(hipHostMalloc((void**)&s->h_fp[k], sizeof(FpItem) * 2 * cap, hipHostMallocDefault));
        s->slot_used[k] = false;
    }
    s->cap = cap;
    return ATR_OK;
}

// One round: one SART step of every admitted series.  A series past its last
// step gets its fill + layout on the round stream and goes to the finisher.
//
// Which series take a step.  Fair share: deficit round robin
// on the voxel-tilt updates of a step.  Every round each series earns the
// step cost of the cheapest one and takes a step when its credit covers its
// own cost, so the cheapest series steps every round and one whose steps
// cost m times as much every m-th round: every series gets the same GPU
// work per round, and a small series admitted beside a large one finishes in
// about its own share of the time instead of at the large one's step rate.
// ATR_SCHEDULE_EVERY_STEP: every series takes one step every round.
// ATR_SCHEDULE_SHORTEST_FIRST (the default): the series closest to done
// (least remaining work, and those within twice it) step; the rest wait, at
// most kStarveRounds rounds in a row.  Small requests then pass large ones
// instead of sharing their rounds (lowest median latency), and a large one
// still advances at least every 17th round.
static atr_status run_round(atr_server* s)
{
    auto& act = s->active;
    atr_status st = grow_slots(s, (int)act.size());
    if (st != ATR_OK) return st;
    const int k = s->slot;
    s->slot = (s->slot + 1) % kRing;
    if (s->slot_used[k]) ATR_HIP(hipEventSynchronize(s->ev_slot[k]));
    std::vector<char> take(act.size(), 1);
    if (s->cfg.schedule == ATR_SCHEDULE_SHORTEST_FIRST) {
        // the series with the least remaining work, and every series within
        // twice that; a series passed over for kStarveRounds rounds in a row
        // is taken regardless, so every series keeps advancing
        double least = -1.0;
        for (auto& rp : act) {
            const double rem = (double)(rp->last - rp->step + 1) * rp->cost;
            if (least < 0.0 || rem < least) least = rem;
        }
        for (size_t i = 0; i < act.size(); i++) {
            Request& r = *act[i];
            const double rem = (double)(r.last - r.step + 1) * r.cost;
            take[i] = rem <= 2.0 * least || r.skipped >= kStarveRounds;
            r.skipped = take[i] ? 0 : r.skipped + 1;
        }
    } else if (s->cfg.schedule == ATR_SCHEDULE_FAIR_SHARE) {
        double quantum = 0.0;
        for (auto& rp : act) quantum = quantum > 0.0 ? std::min(quantum, rp->cost) : rp->cost;
        for (size_t i = 0; i < act.size(); i++) {
            Request& r = *act[i];
            r.credit += quantum;
            take[i] = r.credit >= r.cost * (1.0 - 1e-9);
            if (take[i]) r.credit -= r.cost;
        }
    }
    int n_take = 0;
    std::vector<QueueEntry<FpItem>> fent;
    std::vector<QueueEntry<BpItem>> bent;
    for (size_t i = 0; i < act.size(); i++) {
        if (!take[i]) continue;
        n_take++;
        Request& r = *act[i];
        const Plan& p = *r.sart;
        const int n_all = (int)p.all_tilts.size();
        if (r.step == 0) {
            BpItem e{};
            e.a = bp_args(p, p.d_sino, p.n, p.d_lists, n_all, 1.0f, true);
            e.n_items_y = bp_items_per_slice(e.a);
            e.shape = bp_shape(p.bp_tz, p.bp_chunk);
            bent.push_back({e, p.vol_y * e.n_items_y, (double)n_all * p.bp_tz});
        } else if (r.step <= r.last) {
            const int sb = (r.step - 1) % p.n_subsets;
            const int f0 = p.subset_fp_off[sb], f1 = p.subset_fp_off[sb + 1];
            const int b0 = p.subset_off[sb], b1 = p.subset_off[sb + 1];
            FpItem f{};
            f.a = fp_args(p, p.d_fp_lists + f0, f1 - f0, p.d_fp_pairs + p.subset_pair_off[sb]);
            f.n_groups = (f1 - f0 + f.a.group - 1) / f.a.group;
            fent.push_back({f, p.vol_y * f.n_groups,
                            (double)p.vol_x * p.vol_z * std::min(f.a.group, f1 - f0)});
            if (b1 > b0) {
                BpItem e{};
                e.a = bp_args(p, p.d_diff, p.max_subset, p.d_lists + n_all + b0, b1 - b0, p.relax,
                              false);
                e.n_items_y = bp_items_per_slice(e.a);
                e.shape = bp_shape(p.bp_tz, p.bp_chunk);
                bent.push_back({e, p.vol_y * e.n_items_y, (double)(b1 - b0) * p.bp_tz});
            }
        }
    }
    // one forward-projection launch per staging variant present, one
    // back-projection launch per tile shape
    const std::vector<FpLaunch> fl = lay_out_fp(fent, s->h_fp[k]);
    const int n_fp = fl.empty() ? 0 : fl.back().first + fl.back().n_entries;
    const std::vector<BpLaunch> bl = lay_out_bp(bent, s->h_bp[k]);
    const int n_bp = (int)bent.size();
    hipStream_t q = s->s_comp;
    launch_zero_int(s->d_cnt[k], kCounters, q);
    if (n_fp) {
        ATR_HIP(hipMemcpyAsync(s->d_fp[k], s->h_fp[k], sizeof(FpItem) * n_fp,
                               hipMemcpyHostToDevice, q));
        for (size_t i = 0; i < fl.size(); i++)
            launch_fp_batch(s->storage, fl[i], s->d_fp[k] + fl[i].first, s->d_cnt[k] + i,
                            fp_batch_grid(s->storage, fl[i]), q);
    }
    if (n_bp) {
        ATR_HIP(hipMemcpyAsync(s->d_bp[k], s->h_bp[k], sizeof(BpItem) * n_bp,
                               hipMemcpyHostToDevice, q));
        for (size_t i = 0; i < bl.size(); i++)
            launch_bp_batch(s->storage, bl[i], s->d_bp[k] + bl[i].first,
                            s->d_cnt[k] + FP_VARIANTS + i, bp_batch_grid(s->storage, bl[i]), q);
    }
    ATR_HIP(hipEventRecord(s->ev_slot[k], q));
    s->slot_used[k] = true;
    s->rounds++;
    s->max_batch = std::max<uint64_t>(s->max_batch, (uint64_t)n_take);
    std::vector<std::shared_ptr<Request>> keep;
    for (size_t i = 0; i < act.size(); i++) {
        auto& rp = act[i];
        Request& r = *rp;
        if (!take[i]) {
            keep.push_back(rp);
            continue;
        }
        r.step++;
        if (r.step > r.last) {
            sart_finish(*r.sart, r.d_out, q);
            ATR_HIP(hipEventRecord(r.ev_sart, q));
            std::lock_guard<std::mutex> lk(s->mu);
            s->to_finish.push_back(rp);
            s->cv_fin.notify_all();
        } else {
            keep.push_back(rp);
        }
    }
    act.swap(keep);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "serving round");
    return ATR_OK;
}

// A request that failed on the device side: wait for whatever of it was
// issued (the streams it used), then release it.
static void fail_admitted(atr_server* s, const std::shared_ptr<Request>& rp, atr_status st,
                          const std::string& err)
{
    (void)hipStreamSynchronize(s->s_h2d);
    (void)hipStreamSynchronize(s->s_comp);
    (void)hipStreamSynchronize(s->s_wbp);
    rp->status = st;
    rp->error = err;
    release(s, *rp);
    finish_request(s, *rp, true);
}

// The head of the queue did not fit the arena (it could not grow): its plans
// are released, and it goes back to the preparer and waits, with everything
// behind it, for an earlier request to return memory.
static void requeue_after_release(atr_server* s, std::vector<std::shared_ptr<Request>>& admitted,
                                  size_t i)
{
    for (size_t j = admitted.size(); j-- > i;) {
        auto& rp = admitted[j];
        s->used -= rp->need;
        s->in_flight -= 1;
        s->pending.push_front(rp);
    }
    auto& head = admitted[i];
    head->ready = false;
    s->to_prep.push_front(head);
    s->cv_prep.notify_all();
    admitted.resize(i);
}

static void sched_main(atr_server* s)
{
    (void)hipSetDevice(s->device);
    std::vector<std::shared_ptr<Request>> staging;   // admitted, input upload in flight
    int wait_below = -1;   // after an arena failure: admit again once in_flight < this
    for (;;) {
        std::vector<std::shared_ptr<Request>> admitted, failed;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            auto admissible = [&] {
                if (s->pending.empty() || !s->pending.front()->ready) return false;
                const auto& h = s->pending.front();
                if (h->status != ATR_OK) return true;   // failed on the host: finish it
                if (wait_below >= 0 && s->in_flight >= wait_below) return false;
                const int busy = (int)(s->active.size() + staging.size() + admitted.size());
                return s->used + h->need <= s->budget &&
                       (s->cfg.max_active <= 0 || busy < s->cfg.max_active);
            };
            s->cv_work.wait(lk, [&] {
                return (s->stop && s->pending.empty() && s->active.empty() && staging.empty()) ||
                       !s->active.empty() || !staging.empty() || admissible();
            });
            if (s->stop && s->pending.empty() && s->active.empty() && staging.empty()) break;
            if (wait_below >= 0 && s->in_flight < wait_below) wait_below = -1;
            while (admissible()) {
                auto rp = s->pending.front();
                s->pending.pop_front();
                if (rp->status != ATR_OK) {
                    failed.push_back(rp);
                    continue;
                }
                s->used += rp->need;
                s->in_flight += 1;
                admitted.push_back(rp);
            }
        }
        for (auto& rp : failed) finish_request(s, *rp, false);
        for (size_t i = 0; i < admitted.size(); i++) {
            auto& rp = admitted[i];
            rp->admit = Clock::now();
            atr_status st = admit(s, *rp);
            if (st == ATR_OK) {
                staging.push_back(rp);
                continue;
            }
            std::string err = atr_last_error_string();
            (void)hipStreamSynchronize(s->s_h2d);
            (void)hipStreamSynchronize(s->s_wbp);
            release(s, *rp);
            std::unique_lock<std::mutex> lk(s->mu);
            // requests admitted before this round of admissions still hold memory
            const int others = s->in_flight - (int)(admitted.size() - i);
            if (st == ATR_ERR_OUT_OF_MEMORY && others > 0) {
                requeue_after_release(s, admitted, i);
                wait_below = s->in_flight;
                break;
            }
            rp->status = st;
            rp->error = err;
            lk.unlock();
            finish_request(s, *rp, true);
        }
        // series whose input has landed join the rounds; with nothing else to
        // run, wait for the oldest upload
        if (s->active.empty() && !staging.empty()) (void)hipEventSynchronize(staging.front()->ev_in);
        std::vector<std::shared_ptr<Request>> still;
        for (auto& rp : staging) {
            if (hipEventQuery(rp->ev_in) == hipErrorNotReady) {
                still.push_back(rp);
                continue;
            }
            atr_status st = start_sart(s, *rp);
            if (st == ATR_OK) s->active.push_back(rp);
            else fail_admitted(s, rp, st, atr_last_error_string());
        }
        staging.swap(still);
        if (!s->active.empty()) {
            atr_status st = run_round(s);
            if (st != ATR_OK) {
                std::string err = atr_last_error_string();
                auto act = s->active;
                s->active.clear();
                for (auto& rp : act) fail_admitted(s, rp, st, err);
            }
        }
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->sched_done = true;
    s->cv_fin.notify_all();
    s->cv_prep.notify_all();
}

// ---- finisher: copy-out and release, in the order series finish ------------------
static void fin_main(atr_server* s)
{
    (void)hipSetDevice(s->device);
    for (;;) {
        std::shared_ptr<Request> rp;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_fin.wait(lk, [&] {
                return !s->to_finish.empty() || (s->sched_done && s->in_flight == 0);
            });
            if (s->to_finish.empty()) return;
            rp = s->to_finish.front();
            s->to_finish.pop_front();
        }
        Request& r = *rp;
        hipError_t e = hipEventSynchronize(r.ev_sart);
        r.compute_done = Clock::now();
        if (e == hipSuccess)
            e = hipMemcpyAsync(r.vol_host, r.d_out, r.out_bytes, hipMemcpyDeviceToHost, s->s_d2h);
        if (e == hipSuccess) e = hipEventRecord(r.ev_done, s->s_d2h);
        if (e == hipSuccess) e = hipEventSynchronize(r.ev_done);
        // the WBP volume went out behind its kernels (usually long done)
        if (e == hipSuccess && r.wbp) e = hipEventSynchronize(r.ev_wbp);
        if (e != hipSuccess) {
            (void)hipStreamSynchronize(s->s_d2h);
            r.status = ATR_ERR_HIP;
            r.error = std::string("serving copy-out: ") + hipGetErrorString(e);
        }
        release(s, r);
        finish_request(s, r, true);
    }
}

}  // namespace atr

using namespace atr;

extern "C" {

void atr_server_config_init(atr_server_config* c)
{
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->max_active = 8;   // several series overlap; later arrivals queue (lower p50 than all-at-once)
    c->schedule = ATR_SCHEDULE_SHORTEST_FIRST;
    c->storage = ATR_STORAGE_F16;
}

static void server_free(atr_server* s)
{
    for (hipStream_t q : {s->s_comp, s->s_wbp, s->s_h2d, s->s_d2h})
        if (q) (void)hipStreamSynchronize(q);
    for (int k = 0; k < kRing; k++) {
        if (s->d_bp[k]) (void)hipFree(s->d_bp[k]);
        if (s->d_fp[k]) (void)hipFree(s->d_fp[k]);
        if (s->h_bp[k]) (void)hipHostFree(s->h_bp[k]);
        if (s->h_fp[k]) (void)hipHostFree(s->h_fp[k]);
        if (s->d_cnt[k]) (void)hipFree(s->d_cnt[k]);
        if (s->ev_slot[k]) (void)hipEventDestroy(s->ev_slot[k]);
    }
    for (hipEvent_t e : s->ev_pool) (void)hipEventDestroy(e);
    s->arena.reset();
    for (hipStream_t q : {s->s_comp, s->s_wbp, s->s_h2d, s->s_d2h})
        if (q) (void)hipStreamDestroy(q);
    delete s;
}

atr_status atr_server_create(const atr_server_config* cfg, atr_server** out)
{
    if (!cfg || !out) return fail(ATR_ERR_INVALID_ARG, "atr_server_create: NULL argument");
    if (cfg->storage < ATR_STORAGE_F32 || cfg->storage > ATR_STORAGE_BF16)
        return fail(ATR_ERR_INVALID_ARG, "atr_server_create: unknown storage");
    if (cfg->schedule != ATR_SCHEDULE_FAIR_SHARE && cfg->schedule != ATR_SCHEDULE_EVERY_STEP &&
        cfg->schedule != ATR_SCHEDULE_SHORTEST_FIRST)
        return fail(ATR_ERR_INVALID_ARG, "atr_server_create: unknown schedule");
    *out = nullptr;
    auto* s = new atr_server;
    s->cfg = *cfg;
    s->storage = cfg->storage;
    s->t0 = Clock::now();
    auto bail = [&](atr_status st) {
        server_free(s);
        return st;
    };
    if (hipGetDevice(&s->device) != hipSuccess) return bail(hip_fail(hipErrorNoDevice, "hipGetDevice"));
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return bail(fail(ATR_ERR_HIP, "hipMemGetInfo"));
    const size_t cap = free_b - free_b / 10;
    s->budget = cfg->hbm_budget_bytes ? std::min(cfg->hbm_budget_bytes, cap) : cap;
    // arena chunks of 1/32 of the budget, at least 256 MiB and at most 8 GiB
    // (a larger single allocation gets a chunk of its own size)
    const size_t chunk = std::min<size_t>(size_t(8) << 30,
                                          std::max<size_t>(size_t(256) << 20, s->budget / 32));
    s->arena.reset(new Arena(chunk));
    for (hipStream_t* q : {&s->s_comp, &s->s_wbp, &s->s_h2d, &s->s_d2h})
        if (hipStreamCreateWithFlags(q, hipStreamNonBlocking) != hipSuccess)
            return bail(fail(ATR_ERR_HIP, "hipStreamCreate"));
    for (int k = 0; k < kRing; k++) {
        if (hipMalloc((void**)&s->d_cnt[k], kCounters * sizeof(int)) != hipSuccess ||
            hipEventCreateWithFlags(&s->ev_slot[k], hipEventDisableTiming) != hipSuccess)
            return bail(fail(ATR_ERR_OUT_OF_MEMORY, "atr_server_create: ring"));
    }
    if (grow_slots(s, std::max(8, cfg->max_active)) != ATR_OK) return bail(ATR_ERR_OUT_OF_MEMORY);
    s->th_prep = std::thread(prep_main, s);
    s->th_sched = std::thread(sched_main, s);
    s->th_fin = std::thread(fin_main, s);
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
        s->to_prep.push_back(r);
        s->all[r->id] = r;
        *id = r->id;
    }
    s->cv_prep.notify_all();
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
        info->admit_ms = r->admit == Clock::time_point() ? info->submit_ms : ms_since(s, r->admit);
        info->done_ms = ms_since(s, r->done);
        info->latency_ms = info->done_ms - info->submit_ms;
        info->status = r->status;
        info->prepared_ms = ms_since(s, r->prepared);
        info->compute_done_ms =
            r->compute_done == Clock::time_point() ? info->done_ms : ms_since(s, r->compute_done);
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
    s->cv_prep.notify_all();
    s->cv_work.notify_all();
    s->cv_fin.notify_all();
    if (s->th_prep.joinable()) s->th_prep.join();
    if (s->th_sched.joinable()) s->th_sched.join();
    if (s->th_fin.joinable()) s->th_fin.join();
    server_free(s);
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
