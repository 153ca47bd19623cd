// reproj.cpp -- C ABI of the ProjAlign re-projector (CCalcReproj): the
// reference-tilt range on the host, the FP64 geometry tables, and the slab
// buffer the kernels in kernels_reproj.hip share.
//
// Two ways to run it: through a handle (atr_reproj_create allocates the slab
// buffer once), or through atr_reproj_execute_noalloc, which allocates
// nothing and back-projects into `scratch`, a fixed device array of this
// file.
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>

#include "atr_internal.h"

using namespace atr;

struct atr_reproj {
    int proj_x = 0, proj_y = 0, n_tilts = 0, vol_z = 0, vol_x = 0;
    int slab_rows = 0;
    float* d_vol = nullptr;  // [slab_rows][vol_z][2 proj_x]
};

namespace {

// AreTomo/ProjAlign/CCalcReproj.cpp:73-124 (mFindProjRange), decisions in
// double from the float angles and the reference's 0.01745f.
bool in_range(float a, float t, double cos_a)
{
    const float d2r = 0.01745f;
    if (std::fabs((double)(a - t)) > 20.5) return false;
    return (float)(std::cos((double)(t * d2r)) / cos_a) <= 1.20f;
}

atr_status find_range(const float* tilts, const unsigned char* skip, int n, int idx, int* first,
                      int* last)
{
    const float d2r = 0.01745f;
    const float a = tilts[idx];
    const double cos_a = std::cos((double)(a * d2r));
    int start = -1, end = -1;
    for (int i = 0; i < n && start < 0; i++)
        if (!(skip && skip[i]) && in_range(a, tilts[i], cos_a)) start = i;
    for (int i = start < 0 ? n : start; i < n; i++)
        if (!(skip && skip[i]) && in_range(a, tilts[i], cos_a)) end = i;
    if (start < 0 || end < 0) {
        const int sign = a > 0 ? 1 : -1;
        start = end = idx - sign;
    }
    if (end - start > REPROJ_MAX_TILTS - 1) end = start + REPROJ_MAX_TILTS - 1;
    if (start < 0 || end >= n)
        return fail(ATR_ERR_INVALID_ARG, "re-projection: no reference tilt next to the projection");
    *first = start;
    *last = end;
    return ATR_OK;
}

// Scratch of the no-allocation entry point: the back-projected slab
// [rows][VolZ][2 ProjX] (FP32).  Every GPU holds its own copy, at its own
// address, until the process ends.  k_reproj_bp writes every element of the
// rows it covers before k_reproj_fp reads them, so nothing is cleared
// between calls.  Calls that overlap on one GPU share the array and race.
constexpr size_t REPROJ_SCRATCH_BYTES = size_t(64) << 20;
static __device__ __align__(256) float scratch[REPROJ_SCRATCH_BYTES / sizeof(float)];

// Device address of `scratch` on the current device, looked up once per
// device.  The host-side name `scratch` is not a device address; only this
// pointer is handed to kernels.
atr_status scratch_ptr(float** out)
{
    static std::mutex mu;
    static std::map<int, float*> addr;
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    std::lock_guard<std::mutex> lk(mu);
    auto it = addr.find(dev);
    if (it == addr.end()) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol,
                            "hipGetSymbolAddress (re-projector scratch)");
        it = addr.emplace(dev, (float*)p).first;
    }
    *out = it->second;
    return ATR_OK;
}

atr_status check_dims(int proj_x, int proj_y, int n_tilts, int vol_z)
{
    if (proj_x < 2 || proj_y < 1 || n_tilts < 2 || vol_z < 1)
        return fail(ATR_ERR_INVALID_ARG, "re-projector needs proj_x >= 2, n_tilts >= 2, vol_z >= 1");
    return ATR_OK;
}

// CCalcReproj::DoIt (CCalcReproj.cpp:47-65) for one projection: the
// reference range and FP64 tables on the host, then the slab kernels over
// `slab_rows` image rows at a time in `slab`.
atr_status reproj_run(int proj_x, int proj_y, int n_tilts, int vol_z, const float* proj_dev,
                      const float* tilts, const unsigned char* skip, int proj_idx,
                      float* reproj_dev, float* slab, int slab_rows, hipStream_t stream)
{
    if (proj_idx < 0 || proj_idx >= n_tilts)
        return fail(ATR_ERR_INVALID_ARG, "projection index outside the series");
    for (int t = 0; t < n_tilts; t++)
        if (!std::isfinite(tilts[t]) || std::fabs(tilts[t]) >= 89.0f)
            return fail(ATR_ERR_INVALID_ARG, "re-projection needs finite tilts below 89 deg");
    int first = 0, last = 0;
    atr_status st = find_range(tilts, skip, n_tilts, proj_idx, &first, &last);
    if (st != ATR_OK) return st;
    ReprojArgs a{};
    a.proj_x = proj_x;
    a.proj_y = proj_y;
    a.vol_x = 2 * proj_x;  // CCalcReproj::Setup (CCalcReproj.cpp:36-45)
    a.vol_z = vol_z;
    a.first = first;
    a.n_range = last - first + 1;
    const float d2r = 0.01745f;  // GReproj.cu:6
    const float ang = tilts[proj_idx];
    for (int k = 0; k < a.n_range; k++) {
        const double v = (double)tilts[first + k] * (double)d2r;
        a.cos_t[k] = std::cos(v);
        a.sin_t[k] = std::sin(v);
        a.weight[k] = std::cos(((double)ang - (double)tilts[first + k]) * (double)d2r);
    }
    // GReproj::mForwardProj (GReproj.cu:186-203): float cos / sin of the
    // angle, ray length and the 0.8 occupancy threshold in float
    const double rad = 4.0 * std::atan(1.0) / 180.0;
    const float fc = (float)std::cos(rad * ang), fs = (float)std::sin(rad * ang);
    a.ray = (int)((float)vol_z / fc + 0.5f);
    a.min_count = 0.8f * (float)a.ray;
    a.ray_cos = fc;
    a.ray_sin = fs;
    (void)hipGetLastError();
    launch_reproj(a, proj_dev, slab, slab_rows, reproj_dev, stream);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "re-projector launch");
    return ATR_OK;
}

}  // namespace

extern "C" {

atr_status atr_reproj_find_range(const float* tilts, const unsigned char* skip, int n_tilts,
                                 int proj_idx, int* first, int* last)
{
    if (!tilts || !first || !last) return fail(ATR_ERR_INVALID_ARG, "atr_reproj_find_range: NULL argument");
    if (proj_idx < 0 || proj_idx >= n_tilts) return fail(ATR_ERR_INVALID_ARG, "projection index outside the series");
    return find_range(tilts, skip, n_tilts, proj_idx, first, last);
}

atr_status atr_reproj_create(int proj_x, int proj_y, int n_tilts, int vol_z, atr_reproj** out)
{
    if (!out) return fail(ATR_ERR_INVALID_ARG, "atr_reproj_create: NULL handle");
    *out = nullptr;
    atr_status st = check_dims(proj_x, proj_y, n_tilts, vol_z);
    if (st != ATR_OK) return st;
    auto* r = new (std::nothrow) atr_reproj();
    if (!r) return fail(ATR_ERR_OUT_OF_MEMORY, "host allocation failed");
    r->proj_x = proj_x;
    r->proj_y = proj_y;
    r->n_tilts = n_tilts;
    r->vol_z = vol_z;
    r->vol_x = 2 * proj_x;  // CCalcReproj::Setup (CCalcReproj.cpp:36-45)
    const size_t slice = sizeof(float) * (size_t)r->vol_x * vol_z;
    const size_t budget = (size_t)512 << 20;
    r->slab_rows = (int)std::max<size_t>(1, std::min<size_t>(proj_y, budget / slice));
    hipError_t e = hipMalloc((void**)&r->d_vol, slice * r->slab_rows);
    if (e != hipSuccess) {
        delete r;
        return hip_fail(e, "hipMalloc (re-projector slab)");
    }
    *out = r;
    return ATR_OK;
}

atr_status atr_reproj_execute(atr_reproj* r, const float* proj_dev, const float* tilts,
                              const unsigned char* skip, int proj_idx, float* reproj_dev,
                              void* stream)
{
    if (!r || !proj_dev || !tilts || !reproj_dev)
        return fail(ATR_ERR_INVALID_ARG, "atr_reproj_execute: NULL argument");
    return reproj_run(r->proj_x, r->proj_y, r->n_tilts, r->vol_z, proj_dev, tilts, skip, proj_idx,
                      reproj_dev, r->d_vol, r->slab_rows, (hipStream_t)stream);
}

atr_status atr_reproj_execute_noalloc(int proj_x, int proj_y, int n_tilts, int vol_z,
                                      const float* proj_dev, const float* tilts,
                                      const unsigned char* skip, int proj_idx, float* reproj_dev,
                                      void* stream)
{
    if (!proj_dev || !tilts || !reproj_dev)
        return fail(ATR_ERR_INVALID_ARG, "atr_reproj_execute_noalloc: NULL argument");
    atr_status st = check_dims(proj_x, proj_y, n_tilts, vol_z);
    if (st != ATR_OK) return st;
    const size_t slice = sizeof(float) * 2 * (size_t)proj_x * (size_t)vol_z;
    if (slice > REPROJ_SCRATCH_BYTES)
        return fail(ATR_ERR_INVALID_ARG,
                    "atr_reproj_execute_noalloc: one slab row (2 ProjX x VolZ floats) exceeds the "
                    "64 MiB scratch (2 ProjX x VolZ <= 16777216)");
    float* slab = nullptr;
    if ((st = scratch_ptr(&slab)) != ATR_OK) return st;
    const int rows = (int)std::min<size_t>((size_t)proj_y, REPROJ_SCRATCH_BYTES / slice);
    return reproj_run(proj_x, proj_y, n_tilts, vol_z, proj_dev, tilts, skip, proj_idx, reproj_dev,
                      slab, rows, (hipStream_t)stream);
}

void atr_reproj_destroy(atr_reproj* r)
{
    if (!r) return;
    if (r->d_vol) (void)hipFree(r->d_vol);
    delete r;
}

}  // extern "C"
