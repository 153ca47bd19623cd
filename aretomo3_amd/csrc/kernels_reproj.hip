// kernels_reproj.hip -- the ProjAlign re-projector (AreTomo/ProjAlign/GReproj.cu,
// CCalcReproj.cpp) for gfx950: a projection at angle a is predicted from the
// nearby lower-tilt projections by a cos-weighted nearest-sample
// back-projection onto a 2x-wide slice followed by a forward projection at a.
//
// The reference runs one host loop iteration per image row (two launches
// and two copies per row).  Here one launch back-projects a slab of rows and
// one launch forward-projects it:
//   k_reproj_bp  thread per voxel of the slab [rows][vol_z][2 proj_x]
//   k_reproj_fp  thread per output pixel, the ray summed in order
// Coordinates are evaluated in FP64 with contraction off, from host FP64
// tables, so every nearest-sample pick and every in/out decision is the one
// of the FP64 restatement; sums are FP64.
#include <algorithm>

#include "atr_internal.h"

namespace atr {

__global__ __launch_bounds__(256) void k_reproj_bp(const float* __restrict__ projs, ReprojArgs a,
                                                   int y0, float* __restrict__ vol)
{
#pragma clang fp contract(off)
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int z = blockIdx.y, yl = blockIdx.z;
    if (x >= a.vol_x) return;
    const int y = y0 + yl;
    const double fx = x + 0.5 - a.vol_x * 0.5;
    const double fz = z + 0.5 - a.vol_z * 0.5;
    const double cx = a.proj_x * 0.5;
    const int last = a.proj_x - 1;
    double sum = 0.0, cnt = 0.0;
    for (int k = 0; k < a.n_range; k++) {
        const double c = a.cos_t[k], s = a.sin_t[k];
        double u = fx * c + fz * s + cx;
        if (u < 0 || u > last) continue;
        const float v = projs[((size_t)(a.first + k) * a.proj_y + y) * a.proj_x + (int)u];
        if (v < kEmpty) continue;
        sum += v * c * a.weight[k];
        cnt += a.weight[k];
    }
    vol[((size_t)yl * a.vol_z + z) * a.vol_x + x] = cnt < 0.001 ? kSentinel : (float)(sum / cnt);
}

__global__ __launch_bounds__(256) void k_reproj_fp(const float* __restrict__ vol, ReprojArgs a,
                                                   int y0, int rows, float* __restrict__ out)
{
#pragma clang fp contract(off)
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int yl = blockIdx.y;
    if (x >= a.proj_x || yl >= rows) return;
    const double fc = a.ray_cos, fs = a.ray_sin;
    const double xp = x + 0.5 - a.proj_x * 0.5;
    const double tx = a.vol_x * 0.5 + a.ray * 0.5 * fs + xp / fc;
    const double tz = a.vol_z * 0.5 - a.ray * 0.5 * fc;
    const float* slice = vol + (size_t)yl * a.vol_z * a.vol_x;
    double sum = 0.0;
    int cnt = 0;
    for (int i = 0; i < a.ray; i++) {
        double px = tx - i * fs;
        double pz = tz + i * fc;
        if (px < 0 || px >= a.vol_x || pz < 0 || pz >= a.vol_z) continue;
        float v = slice[(size_t)(int)pz * a.vol_x + (int)px];
        if (v <= kEmpty) continue;
        sum += v;
        cnt += 1;
    }
    out[(size_t)(y0 + yl) * a.proj_x + x] = ((float)cnt < a.min_count) ? kSentinel : (float)(sum / cnt);
}

void launch_reproj(const ReprojArgs& a, const float* projs, float* vol_slab, int slab_rows,
                   float* out, hipStream_t s)
{
    for (int y0 = 0; y0 < a.proj_y; y0 += slab_rows) {
        const int rows = std::min(slab_rows, a.proj_y - y0);
        hipLaunchKernelGGL(k_reproj_bp, dim3((a.vol_x + 255) / 256, a.vol_z, rows), dim3(256), 0, s,
                           projs, a, y0, vol_slab);
        hipLaunchKernelGGL(k_reproj_fp, dim3((a.proj_x + 255) / 256, rows), dim3(256), 0, s, vol_slab,
                           a, y0, rows, out);
    }
}

}  // namespace atr
