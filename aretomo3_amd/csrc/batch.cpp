// batch.cpp -- SART reconstruction of many tilt series at once: every
// back-projection and forward-projection step of all series is one persistent
// work-queue launch (k_back_proj_batch / k_for_proj_diff_batch), and the
// whole iteration sequence can be captured once into a HIP graph and
// replayed.
//
// The series may differ in size, tilt count, thickness, iterations and
// subsets; step k of the batch holds every series that has a step k.  Each
// series keeps its own plan (tables, sinogram, volume), so a series
// reconstructs exactly as it does alone: same kernels, same per-voxel
// operation order, no arithmetic shared across series.
#include <algorithm>
#include <vector>

#include "atr_internal.h"

struct atr_batch {
    std::vector<atr_plan*> plans;
    int storage = ATR_STORAGE_F32;
    struct Step {
        bool fp;        // forward projection + difference, else back-projection
        int first;      // first entry in the item array of its kind
        int n_entries;
        int total;      // queue items
        int counter;    // queue counter slot
        int grid;       // persistent grid
        atr::FpLaunch fl;   // forward projection: the launch (variant, LDS)
        atr::BpLaunch bl;   // back-projection: the launch (tile shape)
    };
    std::vector<Step> steps;
    atr::BpItem* d_bp = nullptr;
    atr::FpItem* d_fp = nullptr;
    int* d_counters = nullptr;
    int n_counters = 0;
    bool graph = false;
    hipStream_t stream = nullptr;   // the batch's own stream (graph capture needs one)
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    hipGraph_t g = nullptr;
    hipGraphExec_t gx = nullptr;
    bool profiling = false;
    int launches[atr::K_COUNT] = {0};
    double total_ms[atr::K_COUNT] = {0};
};

namespace atr {

static void batch_free(atr_batch* b)
{
    if (b->gx) (void)hipGraphExecDestroy(b->gx);
    if (b->g) (void)hipGraphDestroy(b->g);
    if (b->d_bp) (void)hipFree(b->d_bp);
    if (b->d_fp) (void)hipFree(b->d_fp);
    if (b->d_counters) (void)hipFree(b->d_counters);
    if (b->ev_in) (void)hipEventDestroy(b->ev_in);
    if (b->ev_out) (void)hipEventDestroy(b->ev_out);
    if (b->stream) (void)hipStreamDestroy(b->stream);
}

static atr_status batch_build(atr_batch* b)
{
    std::vector<BpItem> bp;
    std::vector<FpItem> fp;
    int max_iters = 0, max_sub = 0;
    for (atr_plan* h : b->plans) {
        max_iters = std::max(max_iters, h->p.iters);
        max_sub = std::max(max_sub, h->p.n_subsets);
    }
    // one back-projection launch per tile shape present
    auto push_bp = [&](std::vector<QueueEntry<BpItem>>& ent) {
        const int b_first = (int)bp.size();
        bp.resize(b_first + ent.size());
        for (const BpLaunch& l : lay_out_bp(ent, bp.data() + b_first)) {
            atr_batch::Step st{false, b_first + l.first, l.n_entries, l.total, b->n_counters++,
                               0, {}, l};
            st.grid = std::min(bp_batch_grid(b->storage, l), std::max(l.total, 1));
            b->steps.push_back(st);
        }
    };
    // initial back-projection of every series (relax 1, CTomoSart.cpp:68-69)
    {
        std::vector<QueueEntry<BpItem>> ent;
        for (atr_plan* h : b->plans) {
            const Plan& p = h->p;
            const int n_all = (int)p.all_tilts.size();
            if (n_all == 0) continue;
            BpItem e{};
            e.a = bp_args(p, p.d_sino, p.n, p.d_lists, n_all, 1.0f, true);
            e.n_items_y = bp_items_per_slice(e.a);
            e.shape = bp_shape(p.bp_tz, p.bp_chunk);
            ent.push_back({e, p.vol_y * e.n_items_y, (double)n_all * p.bp_tz});
        }
        push_bp(ent);
    }
    for (int it = 0; it < max_iters; it++)
        for (int sb = 0; sb < max_sub; sb++) {
            std::vector<QueueEntry<FpItem>> fent;
            std::vector<QueueEntry<BpItem>> bent;
            for (atr_plan* h : b->plans) {
                const Plan& p = h->p;
                if (it >= p.iters || sb >= p.n_subsets) continue;
                const int f0 = p.subset_fp_off[sb], f1 = p.subset_fp_off[sb + 1];
                const int b0 = p.subset_off[sb], b1 = p.subset_off[sb + 1];
                const int n_all = (int)p.all_tilts.size();
                FpItem f{};
                f.a = fp_args(p, p.d_fp_lists + f0, f1 - f0, p.d_fp_pairs + p.subset_pair_off[sb]);
                f.n_groups = (f1 - f0 + f.a.group - 1) / f.a.group;
                fent.push_back({f, p.vol_y * f.n_groups,
                                (double)p.vol_x * p.vol_z * std::min(f.a.group, f1 - f0)});
                if (b1 > b0) {
                    BpItem e{};
                    e.a = bp_args(p, p.d_diff, p.max_subset, p.d_lists + n_all + b0, b1 - b0,
                                  p.relax, false);
                    e.n_items_y = bp_items_per_slice(e.a);
                    e.shape = bp_shape(p.bp_tz, p.bp_chunk);
                    bent.push_back({e, p.vol_y * e.n_items_y, (double)(b1 - b0) * p.bp_tz});
                }
            }
            // one forward-projection launch per staging variant present
            const int f_first = (int)fp.size();
            fp.resize(f_first + 2 * fent.size());
            for (const FpLaunch& l : lay_out_fp(fent, fp.data() + f_first)) {
                atr_batch::Step sf{true, f_first + l.first, l.n_entries, l.total,
                                   b->n_counters++, 0, l, {}};
                sf.grid = std::min(fp_batch_grid(b->storage, l), std::max(l.total, 1));
                b->steps.push_back(sf);
            }
            push_bp(bent);
        }
    if (!bp.empty()) ATR_HIP(hipMalloc((void**)&b->d_bp, bp.size() * sizeof(BpItem)));
    if (!fp.empty()) ATR_HIP(hipMalloc((void**)&b->d_fp, fp.size() * sizeof(FpItem)));
    ATR_HIP(hipMalloc((void**)&b->d_counters, sizeof(int) * std::max(1, b->n_counters)));
    if (!bp.empty())
        ATR_HIP(hipMemcpy(b->d_bp, bp.data(), bp.size() * sizeof(BpItem), hipMemcpyHostToDevice));
    if (!fp.empty())
        ATR_HIP(hipMemcpy(b->d_fp, fp.data(), fp.size() * sizeof(FpItem), hipMemcpyHostToDevice));
    ATR_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    ATR_HIP(hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming));
    ATR_HIP(hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming));
    return ATR_OK;
}

// The iteration sequence: reset the queue counters (a kernel, so the
// sequence can be captured), then every step in order.
static void batch_steps(atr_batch* b, hipStream_t s, bool timing)
{
    launch_zero_int(b->d_counters, b->n_counters, s);
    for (const auto& st : b->steps) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (timing) {
            (void)hipEventCreate(&e0);
            (void)hipEventCreate(&e1);
            (void)hipEventRecord(e0, s);
        }
        if (st.fp)
            launch_fp_batch(b->storage, st.fl, b->d_fp + st.first, b->d_counters + st.counter,
                            st.grid, s);
        else
            launch_bp_batch(b->storage, st.bl, b->d_bp + st.first, b->d_counters + st.counter,
                            st.grid, s);
        if (timing) {
            (void)hipEventRecord(e1, s);
            float ms = 0.0f;
            if (hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) {
                const int k = st.fp ? K_FP : K_BP;
                b->launches[k] += 1;
                b->total_ms[k] += ms;
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
        }
    }
}

static atr_status batch_execute(atr_batch* b, const float* const* in, float* const* out,
                                hipStream_t caller)
{
    (void)hipGetLastError();
    hipStream_t s = b->stream;
    ATR_HIP(hipEventRecord(b->ev_in, caller));
    ATR_HIP(hipStreamWaitEvent(s, b->ev_in, 0));
    for (size_t i = 0; i < b->plans.size(); i++) {
        Plan& p = b->plans[i]->p;
        const bool prof = p.profiling;
        p.profiling = false;
        sart_prepare(p, in[i], out[i], 0, s);
        p.profiling = prof;
    }
    if (b->graph && !b->profiling) {
        if (b->gx == nullptr) {
            ATR_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            batch_steps(b, s, false);
            hipError_t e = hipStreamEndCapture(s, &b->g);
            if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
            e = hipGraphInstantiate(&b->gx, b->g, nullptr, nullptr, 0);
            if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
        }
        ATR_HIP(hipGraphLaunch(b->gx, s));
    } else {
        batch_steps(b, s, b->profiling);
    }
    for (size_t i = 0; i < b->plans.size(); i++) {
        Plan& p = b->plans[i]->p;
        const bool prof = p.profiling;
        p.profiling = false;
        sart_finish(p, out[i], s);
        p.profiling = prof;
    }
    ATR_HIP(hipEventRecord(b->ev_out, s));
    ATR_HIP(hipStreamWaitEvent(caller, b->ev_out, 0));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "batch launch");
    return ATR_OK;
}

}  // namespace atr

using namespace atr;

extern "C" {

atr_status atr_batch_create(atr_plan* const* plans, int n_plans, int flags, atr_batch** batch)
{
    if (batch == nullptr || plans == nullptr || n_plans < 1)
        return fail(ATR_ERR_INVALID_ARG, "atr_batch_create: need plans and an output pointer");
    *batch = nullptr;
    for (int i = 0; i < n_plans; i++) {
        if (plans[i] == nullptr) return fail(ATR_ERR_INVALID_ARG, "atr_batch_create: NULL plan");
        if (plans[i]->p.method != ATR_METHOD_SART)
            return fail(ATR_ERR_INVALID_ARG, "atr_batch_create: batches hold SART plans only");
        if (plans[i]->p.storage != plans[0]->p.storage)
            return fail(ATR_ERR_INVALID_ARG, "atr_batch_create: plans differ in storage type");
        for (int j = 0; j < i; j++)
            if (plans[j] == plans[i])
                return fail(ATR_ERR_INVALID_ARG, "atr_batch_create: a plan appears twice");
    }
    atr_batch* b = new atr_batch;
    b->plans.assign(plans, plans + n_plans);
    b->storage = plans[0]->p.storage;
    b->graph = (flags & ATR_BATCH_GRAPH) != 0;
    atr_status st = batch_build(b);
    if (st != ATR_OK) {
        batch_free(b);
        delete b;
        return st;
    }
    // the item arrays now hold the plans' table pointers: retuning a member
    // plan (which rebuilds them) is refused until the batch is destroyed
    for (atr_plan* h : b->plans) h->p.batch_refs += 1;
    *batch = b;
    return ATR_OK;
}

atr_status atr_batch_execute(atr_batch* batch, const float* const* proj_dev, float* const* vol_dev,
                             void* stream)
{
    if (!batch || !proj_dev || !vol_dev) return fail(ATR_ERR_INVALID_ARG, "atr_batch_execute: NULL argument");
    for (size_t i = 0; i < batch->plans.size(); i++)
        if (!proj_dev[i] || !vol_dev[i]) return fail(ATR_ERR_INVALID_ARG, "atr_batch_execute: NULL buffer");
    return batch_execute(batch, proj_dev, vol_dev, (hipStream_t)stream);
}

atr_status atr_batch_set_profiling(atr_batch* batch, int enable)
{
    if (!batch) return fail(ATR_ERR_INVALID_ARG, "batch is NULL");
    batch->profiling = enable != 0;
    return ATR_OK;
}

atr_status atr_batch_kernel_stats(const atr_batch* batch, int kernel, int* launches, double* total_ms)
{
    if (!batch || kernel < 0 || kernel >= K_COUNT) return fail(ATR_ERR_INVALID_ARG, "bad batch or kernel id");
    if (launches) *launches = batch->launches[kernel];
    if (total_ms) *total_ms = batch->total_ms[kernel];
    return ATR_OK;
}

atr_status atr_batch_reset_stats(atr_batch* batch)
{
    if (!batch) return fail(ATR_ERR_INVALID_ARG, "batch is NULL");
    for (int k = 0; k < K_COUNT; k++) {
        batch->launches[k] = 0;
        batch->total_ms[k] = 0.0;
    }
    return ATR_OK;
}

void atr_batch_destroy(atr_batch* batch)
{
    if (!batch) return;
    (void)hipStreamSynchronize(batch->stream);
    for (atr_plan* h : batch->plans) h->p.batch_refs -= 1;
    batch_free(batch);
    delete batch;
}

}  // extern "C"
