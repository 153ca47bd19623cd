// thickness.hip -- sample thickness of a tilt series (Recon/CCalcVolThick.cpp):
// a heavily binned WBP volume, the per-section standard deviation of its
// central tile on the GPU, then the reference's smoothing and edge search on
// the host (a profile of at most a few hundred values).
//
// Built with -ffp-contract=off: the edge search compares float sums with
// the profile mean, so its arithmetic is kept exactly as written.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "atr_internal.h"

namespace atr {

// One workgroup per section z: FP64 sum, sum of squares and count of the
// non-empty samples of the central tile (CCalcVolThick.cpp:141-170); the
// tile is (int)(0.4 n)/2*2 wide in x and y.
__global__ __launch_bounds__(256) void k_tile_moments(const float* __restrict__ vol, int nx, int ny,
                                                      int tx, int ty, int x0, int y0,
                                                      double* __restrict__ moments)
{
    const int z = blockIdx.x;
    const float* sec = vol + (size_t)z * ny * nx;
    double s1 = 0.0, s2 = 0.0, c = 0.0;
    for (int i = threadIdx.x; i < tx * ty; i += blockDim.x) {
        const int j = i / tx, k = i - j * tx;
        const float v = sec[(size_t)(y0 + j) * nx + x0 + k];
        if (v < -1e20f) continue;
        s1 += v;
        s2 += (double)(v * v);
        c += 1.0;
    }
    __shared__ double r1[256], r2[256], rc[256];
    r1[threadIdx.x] = s1;
    r2[threadIdx.x] = s2;
    rc[threadIdx.x] = c;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            r1[threadIdx.x] += r1[threadIdx.x + o];
            r2[threadIdx.x] += r2[threadIdx.x + o];
            rc[threadIdx.x] += rc[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        moments[3 * z] = r1[0];
        moments[3 * z + 1] = r2[0];
        moments[3 * z + 2] = rc[0];
    }
}

}  // namespace atr

using namespace atr;

namespace {

// CCalcVolThick::mSmooth (CCalcVolThick.cpp:93-110)
void smooth(std::vector<float>& p)
{
    const int n = (int)p.size(), win = 11;
    std::vector<float> b(n);
    for (int i = 0; i < n; i++) {
        double sum = 0.0;
        for (int j = 0; j < win; j++) {
            int k = std::min(std::max(i - win / 2 + j, 0), n - 1);
            sum += p[k];
        }
        b[i] = (float)sum / win;
    }
    p.swap(b);
}

// CCalcVolThick::mDetectEdges (CCalcVolThick.cpp:185-255); a side without a
// crossing of the mean keeps its search bound (uninitialised in the reference)
void edges(const std::vector<float>& p, float binning, atr_thickness* out)
{
    const int n = (int)p.size(), half = n / 2;
    const int bot = (int)(n * 0.1), top = n - bot;
    float mean = 0.0f;
    for (int i = bot; i < top; i++) mean += p[i];
    mean = mean / (top - bot);
    int e0 = bot, e1 = top;
    for (int i = bot; i < half; i++)
        if (p[i] >= mean) {
            e0 = i;
            break;
        }
    float mn = mean;
    int imin = e0;
    for (int i = bot; i < e0; i++)
        if (p[i] < mn) {
            mn = p[i];
            imin = i;
        }
    e0 = (e0 + imin) / 2;
    for (int i = top; i > half; i--)
        if (p[i] >= mean) {
            e1 = i;
            break;
        }
    mn = mean;
    imin = e1;
    for (int i = top; i > e1; i--)
        if (p[i] < mn) {
            mn = p[i];
            imin = i;
        }
    e1 = (e1 + imin) / 2;
    out->low_edge = (int)(e0 * binning);
    out->high_edge = (int)(e1 * binning);
    out->thickness = out->high_edge - out->low_edge;
    out->offset_z = (out->low_edge + out->high_edge) / 2 - (int)(half * binning);
}

}  // namespace

extern "C" {

atr_status atr_thickness_edges(const float* std_profile, int n, float binning, atr_thickness* out,
                               float* smoothed)
{
    if (!std_profile || !out) return fail(ATR_ERR_INVALID_ARG, "atr_thickness_edges: NULL argument");
    if (n < 10) return fail(ATR_ERR_INVALID_ARG, "thickness profile needs at least 10 sections");
    if (!(binning > 0.0f) || !std::isfinite(binning))
        return fail(ATR_ERR_INVALID_ARG, "thickness binning must be positive");
    std::vector<float> p(std_profile, std_profile + n);
    smooth(p);
    if (smoothed) std::memcpy(smoothed, p.data(), sizeof(float) * n);
    edges(p, binning, out);
    out->binning = binning;
    out->n_profile = n;
    return ATR_OK;
}

atr_status atr_calc_vol_thick(const atr_series_desc* desc, float pixel_size, const float* proj_host,
                              atr_thickness* out, float* profile)
{
    if (!desc || !proj_host || !out) return fail(ATR_ERR_INVALID_ARG, "atr_calc_vol_thick: NULL argument");
    if (!(pixel_size > 0.0f) || !std::isfinite(pixel_size))
        return fail(ATR_ERR_INVALID_ARG, "pixel size must be positive");
    std::memset(out, 0, sizeof(*out));
    // CCalcVolThick::DoIt (:47-90): binning max(20 / pixel, 20), a Fourier-
    // binned pre-stage with fill, WBP of (int)(0.8 binned width)/2*2 sections,
    // z-flipped xyz volume
    const float binning = std::fmax(20.0f / pixel_size, 20.0f);
    atr_series_desc d = *desc;
    int ax = d.proj_x, ay = d.proj_y;
    if (d.tilt_axis_deg) aligned_size(d.proj_x, d.proj_y, d.tilt_axis_deg[0], &ax, &ay);
    const int bx = (int)(ax / binning) / 2 * 2, by = (int)(ay / binning) / 2 * 2;
    if (bx < 4 || by < 2) return fail(ATR_ERR_INVALID_ARG, "series too small for binning by 20");
    d.method = ATR_METHOD_WBP;
    d.bin = binning;
    d.vol_z = (int)(bx * 0.8) / 2 * 2;
    d.out_layout = ATR_LAYOUT_XYZ;
    d.storage = ATR_STORAGE_F32;
    d.random_fill = 1;
    if (d.vol_z < 11) return fail(ATR_ERR_INVALID_ARG, "binned volume thinner than 11 sections");
    atr_plan* plan = nullptr;
    atr_status st = atr_plan_create(&d, &plan);
    if (st != ATR_OK) return st;
    int vx = 0, vy = 0, vz = 0;
    atr_plan_get_volume_size(plan, &vx, &vy, &vz);
    const size_t n_in = (d.tilt_stride ? d.tilt_stride : (d.row_stride ? d.row_stride : (size_t)d.proj_x) * d.proj_y) *
                            (d.n_tilts - 1) +
                        (d.row_stride ? d.row_stride : (size_t)d.proj_x) * (d.proj_y - 1) + d.proj_x;
    const size_t n_vol = (size_t)vx * vy * vz;
    const int nz = vz, n = nz - 1;
    const int tx = (int)(vx * 0.4f) / 2 * 2, ty = (int)(vy * 0.4f) / 2 * 2;
    float *d_in = nullptr, *d_vol = nullptr;
    double* d_mom = nullptr;
    std::vector<double> mom(3 * (size_t)n);
    hipStream_t s = nullptr;
    hipError_t e = hipStreamCreate(&s);
    if (e == hipSuccess) e = hipMalloc((void**)&d_in, sizeof(float) * n_in);
    if (e == hipSuccess) e = hipMalloc((void**)&d_vol, sizeof(float) * n_vol);
    if (e == hipSuccess) e = hipMalloc((void**)&d_mom, sizeof(double) * 3 * n);
    if (e == hipSuccess) e = hipMemcpyAsync(d_in, proj_host, sizeof(float) * n_in, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) st = atr_execute(plan, d_in, d_vol, 0, s);
    if (e == hipSuccess && st == ATR_OK && tx > 0 && ty > 0) {
        hipLaunchKernelGGL(k_tile_moments, dim3(n), dim3(256), 0, s, d_vol, vx, vy, tx, ty,
                           (vx - tx) / 2, (vy - ty) / 2, d_mom);
        e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(mom.data(), d_mom, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
    if (d_in) (void)hipFree(d_in);
    if (d_vol) (void)hipFree(d_vol);
    if (d_mom) (void)hipFree(d_mom);
    if (s) (void)hipStreamDestroy(s);
    atr_plan_destroy(plan);
    if (e != hipSuccess) return hip_fail(e, "atr_calc_vol_thick");
    if (st != ATR_OK) return st;
    if (tx <= 0 || ty <= 0) return fail(ATR_ERR_INVALID_ARG, "binned volume too small for the tile");
    // CCalcVolThick::mMeasureStd (:141-170): moments in double, sqrt in float
    std::vector<float> prof(n);
    for (int z = 0; z < n; z++) {
        const double c = mom[3 * z + 2];
        if (c <= 0) {
            prof[z] = 0.0f;
            continue;
        }
        double m1 = mom[3 * z] / c;
        double m2 = mom[3 * z + 1] / c - m1 * m1;
        prof[z] = m2 <= 0 ? 0.0f : std::sqrt((float)m2);
    }
    if (profile) std::memcpy(profile, prof.data(), sizeof(float) * n);
    if ((st = atr_thickness_edges(prof.data(), n, binning, out, nullptr)) != ATR_OK) return st;
    // CCorrTomoStack.cpp:114-115,178 and CCalcVolThick.cpp:75: pixel of the
    // binned stack (aligned height / binned height) over the binning
    out->pixel_size = pixel_size * ((float)ay / (float)by) / binning;
    out->vol_z = vz;
    return ATR_OK;
}

}  // extern "C"
