// kernels_sart.hip -- pre-stage, weighting, SART back-projection and the
// volume fill/output kernels (gfx950); the forward projection + difference
// is in kernels_fp.hip.
//
// Data layout in HBM (one series):
//   d_aligned [N][aln_y][aln_x]      pre-stage output, -1e30 where empty
//   d_sino    [VolY][N][stride]      weighted projections, sinogram-major:
//                                    the sinogram of slice y is contiguous
//   d_diff    [VolY][max_subset][stride]  SART residual of the current subset
//   d_vol     [VolY][VolZ][VolX]     SART state, x fastest (XZY, unflipped)
#include <atomic>
#include <climits>

#include "atr_internal.h"
#include "storage.h"
#include "work_queue.h"

namespace atr {

// ------------------------------------------------------------------------
// Pre-stage: Correct/GCorrPatchShift.cu:58-100 (global + local shift,
// tilt-axis rotation, nearest neighbour, -1e30 outside the raw image),
// geometry in FP64 with contraction off so that the rounding of every
// coordinate is the rounding of the FP64 restatement.  One workgroup per
// (aligned row, tilt); the row's (sum, count) of non-empty pixels is written
// for the fill mean.  The tilt's patch table (centres x, y, shifts x, y,
// good flags; n_patches each) is staged in LDS.
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_prestage(const float* __restrict__ in, size_t tilt_stride,
                                                  size_t row_stride, int in_x, int in_y,
                                                  const double2* __restrict__ axis_cs,
                                                  const float2* __restrict__ shifts,
                                                  const float* __restrict__ local, int n_patches,
                                                  float* __restrict__ out, int out_x, int out_y,
                                                  double2* __restrict__ row_stats)
{
#pragma clang fp contract(off)
    extern __shared__ float s_patch[];
    const int y = blockIdx.x, t = blockIdx.y;
    if (n_patches > 0) {
        const float* lp = local + (size_t)t * 5 * n_patches;
        for (int i = threadIdx.x; i < 5 * n_patches; i += blockDim.x) s_patch[i] = lp[i];
        __syncthreads();
    }
    const float *pcx = s_patch, *pcy = s_patch + n_patches;
    const float *psx = s_patch + 2 * n_patches, *psy = s_patch + 3 * n_patches;
    const float* pgood = s_patch + 4 * n_patches;
    const double ca = axis_cs[t].x, sa = axis_cs[t].y;
    const double sx = (double)shifts[t].x + in_x * 0.5;
    const double sy = (double)shifts[t].y + in_y * 0.5;
    const double yc = y + 0.5 - out_y * 0.5;
    const float* src = in + (size_t)t * tilt_stride;
    float* dst = out + ((size_t)t * out_y + y) * out_x;
    double sum = 0.0;
    int cnt = 0;
    for (int x = threadIdx.x; x < out_x; x += blockDim.x) {
        double xc = x + 0.5 - out_x * 0.5;
        double xr = xc * ca - yc * sa;
        double yr = xc * sa + yc * ca;
        if (n_patches > 0) {
            // Correct/GCorrPatchShift.cu:18-48: Gaussian-weighted mean of the
            // good patches' shifts (good >= 0.9); no weight -> no shift
            double wx = 0.0, wy = 0.0, sw = 0.0;
            int used = 0;
            for (int q = 0; q < n_patches; q++) {
                if (pgood[q] < 0.9f) continue;
                double dx = (xr - (double)pcx[q]) / in_x;
                double dy = (yr - (double)pcy[q]) / in_y;
                double w = exp(-100.0 * (dx * dx + dy * dy));
                wx += (double)psx[q] * w;
                wy += (double)psy[q] * w;
                sw += w;
                used += 1;
            }
            if (used > 0 && sw > 0) {
                xr += wx / sw;
                yr += wy / sw;
            }
        }
        xr += sx;
        yr += sy;
        int ix = (int)(xr + 0.5);
        int iy = (int)(yr + 0.5);
        float v = kSentinel;
        if (ix >= 0 && ix < in_x && iy >= 0 && iy < in_y) {
            v = src[(size_t)iy * row_stride + ix];
            if (!(v < kEmpty)) {
                sum += v;
                cnt += 1;
            }
        }
        dst[x] = v;
    }
    // fixed-order block reduction (deterministic)
    __shared__ double s_sum[256];
    __shared__ int s_cnt[256];
    s_sum[threadIdx.x] = sum;
    s_cnt[threadIdx.x] = cnt;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            s_sum[threadIdx.x] += s_sum[threadIdx.x + o];
            s_cnt[threadIdx.x] += s_cnt[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) row_stats[(size_t)t * out_y + y] = make_double2(s_sum[0], (double)s_cnt[0]);
}

// Per-tilt fill statistics: sum of the row partials in a fixed order.
__global__ __launch_bounds__(256) void k_tilt_stats(const double2* __restrict__ row_stats, int rows,
                                                    double2* __restrict__ tilt_stats)
{
    const int t = blockIdx.x;
    double sum = 0.0, cnt = 0.0;
    for (int y = threadIdx.x; y < rows; y += blockDim.x) {
        double2 r = row_stats[(size_t)t * rows + y];
        sum += r.x;
        cnt += r.y;
    }
    __shared__ double s_sum[256], s_cnt[256];
    s_sum[threadIdx.x] = sum;
    s_cnt[threadIdx.x] = cnt;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            s_sum[threadIdx.x] += s_sum[threadIdx.x + o];
            s_cnt[threadIdx.x] += s_cnt[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) tilt_stats[t] = make_double2(s_sum[0], s_cnt[0]);
}

// Value of aligned pixel i of tilt t after MaUtil/GFillEmpty2D.cu:60-88:
// an empty pixel follows the random chain (<= 100 hops) to the first
// non-empty pixel, else takes the tilt mean.  Reads only the unfilled
// image, so every pixel can be resolved independently.
__device__ __forceinline__ float filled_value(const float* __restrict__ img, int i,
                                              const int* __restrict__ table, float mean)
{
    float v = img[i];
    if (!(v < kEmpty)) return v;
    int next = table[i];
    for (int j = 0; j < 100; j++) {
        float w = img[next];
        if (w < kEmpty) {
            next = table[next];
            continue;
        }
        return w;
    }
    return mean;
}

// ------------------------------------------------------------------------
// Weighting (Recon/GWeightProjs.cu:7-19): s *= cos_t / VolZ, written
// sinogram-major into d_sino, with the pre-stage random fill resolved on the
// fly.  One thread per pixel; grid (x-blocks, tilt, y).
// ------------------------------------------------------------------------
template <typename S>
__global__ __launch_bounds__(256) void k_weight(const float* __restrict__ in, size_t tilt_stride,
                                                size_t row_stride, int fill,
                                                const int* __restrict__ table,
                                                const double2* __restrict__ tilt_stats,
                                                const float2* __restrict__ cs, int n, int proj,
                                                int vol_z, S* __restrict__ sino, int stride,
                                                int* __restrict__ sflag)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int t = blockIdx.y, y = blockIdx.z;
    if (x >= stride) return;
    S* dst = sino + ((size_t)y * n + t) * stride;
    if (x >= proj) {
        st<S>(dst + x, 0.0f);
        return;
    }
    const float w = cs[t].x / (float)vol_z;
    const float* img = in + (size_t)t * tilt_stride;
    float v;
    if (fill) {
        double2 ts = tilt_stats[t];
        float mean = ts.y > 0 ? (float)(ts.x / ts.y) : 0.0f;
        v = filled_value(img, (int)(y * row_stride + x), table, mean);
    } else {
        v = img[(size_t)y * row_stride + x];
    }
    v *= w;
    if (v <= kEmpty) atomicOr(sflag + y, 1);
    st<S>(dst + x, v);
}

// ------------------------------------------------------------------------
// SART back-projection (Recon/GBackProj.cu:12-60).
//
// Workgroup = BP_TX x TZ voxels of one slice, lane = x, each thread walks
// TZ z-rows (TZ = 16 by default, 8 a tuning choice).  For each chunk of CH
// tilts (16 or 8) the needed part of every
// sinogram row is staged in LDS as pairs (p[i] - i d[i], d[i]), d[i] =
// p[i+1]-p[i], so that one ds_read_b64 and one FMA with the window
// coordinate itself give the linear interpolation (slices with empty samples
// stage (p[i], d[i]) and weight d by the fraction).  Inclusion comes
// from the FP64 interval tables: tiles the host classified as fully inside
// skip all per-lane tests and count per tile; partial tiles test x against
// the (tilt, z) interval.  Sums and counts are FP32/int per voxel, in tilt
// order, as the reference.  Everything per tilt (list entry, cos/sin, the
// tile's window start and class) is wave-uniform and read with scalar loads.
// ------------------------------------------------------------------------
// One tile: voxels x0 .. x0+BP_TX-1, z0 .. z0+TZ-1 of slice y.  Ends
// after a barrier, with the LDS free for the next tile.
//
// SENT: the slice's sinogram rows may hold empty samples.  Without them the
// staged pair of window index i is (q_i, d_i) with q_i = q(w0 + i) - i d_i,
// so the interpolation at window coordinate u is the single FMA
// q_i + u d_i (no fraction); with them the pair is (q(w0 + i), d_i) and the
// value q + fract(u) d, whose sentinel test must see the exact -1e30 scale.
template <typename S, int TZ, int CH, bool WBP, bool SENT>
__device__ __forceinline__ void bp_tile_body(cptr<BpArgs> A, float2 (&pairs)[CH][bp_win(TZ)],
                                             int xt, int zt, int y)
{
    const gptr<const S> sino = (gptr<const S>)A->sino;
    const int rows_per_slice = A->rows_per_slice, stride = A->stride, proj = A->proj;
    const cptr<TiltRow> list = as_cptr(A->list);
    const int n_list = A->n_list;
    const cptr<float2> cs = as_cptr(A->cs);
    const cptr<int2> lohi = as_cptr(A->lohi);
    const cptr<int4> tiles = as_cptr(A->tiles);
    const int n_ztiles = A->n_ztiles, n_xtiles = A->n_xtiles;
    const float relax = A->relax;
    const int positivity = A->positivity;
    const float scale = A->scale;
    const gptr<float> vol = as_gptr(A->vol);
    const int vol_x = A->vol_x, vol_z = A->vol_z;
    int* vflag = A->vflag;
    const int tid = threadIdx.x;
    constexpr int WIN = bp_win(TZ);
    const int x0 = xt * BP_TX, z0 = zt * TZ;
    const int x = x0 + tid;
    const int xc = min(x, vol_x - 1);
    const gptr<float> vslice = vol + (size_t)y * vol_z * vol_x;
    // the slice (and below, the slice's sinogram rows) as buffers: a voxel's
    // row offset is a scalar, its column the lane's fixed byte offset
    const __amdgpu_buffer_rsrc_t vres =
        buffer_of((const void*)vslice, (size_t)vol_z * vol_x * sizeof(float));
    const int row_bytes = vol_x * (int)sizeof(float);

    // the voxels this thread updates, loaded now and consumed at the end
    // (the WBP grid is written, not updated)
    float vold[TZ];
#pragma unroll
    for (int k = 0; k < TZ; k++)
        vold[k] = WBP ? 0.0f
                      : ldb<float>(vres, xc * (int)sizeof(float), min(z0 + k, vol_z - 1) * row_bytes);

    float acc[TZ];
    int cnt[TZ];
#pragma unroll
    for (int k = 0; k < TZ; k++) {
        acc[k] = 0.0f;
        cnt[k] = 0;
    }
    int full_cnt = 0;
    const float xf = (float)(xc - x0);   // offset from the tile's first voxel
    const gptr<const S> base = sino + (size_t)y * rows_per_slice * stride;
    const __amdgpu_buffer_rsrc_t sres =
        buffer_of((const void*)base, (size_t)rows_per_slice * stride * sizeof(S));
    const cptr<int4> tile_tab = tiles + (size_t)zt * n_xtiles + xt;
    const size_t tilt_step = (size_t)n_ztiles * n_xtiles;

    for (int c0 = 0; c0 < n_list; c0 += CH) {
        const int nc = min(CH, n_list - c0);
        // ---- stage the row segments.  Pair i of tilt j is (q(g), q(g+1) - q(g))
        // at g = w0 + i, where q is the row extended linearly beyond its ends:
        // q(g) = p[ga] + (g - ga) (p[ga+1] - p[ga]), ga = clamp(g, 0, P-2).
        // Inside [0, P-2] that is the plain pair; at g = -1 it extrapolates, so
        // a coordinate that FP32 rounding puts just below 0 still interpolates
        // the first sample.  Pairs further out are only read by voxels the
        // exact intervals exclude.  Four tilts' loads are issued before any
        // pair is formed. ----
        for (int h = 0; h < nc; h += 4) {
            float pv[4][2], rv[4][2];
            int gv[4][2];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const TiltRow tr = ldc(list + c0 + min(h + q, nc - 1));
                const int w0 = tile_tab[(size_t)tr.tilt * tilt_step].x;
                const int row_off = tr.row * stride * (int)sizeof(S);
#pragma unroll
                for (int m = 0; m < 2; m++) {
                    const int g = w0 + tid + m * BP_TX;
                    const int ga = min(max(g, 0), proj - 2);
                    if (m == 0 || tid < WIN - BP_TX) {
                        pv[q][m] = ldb<S>(sres, ga * (int)sizeof(S), row_off);
                        rv[q][m] = ldb<S>(sres, (ga + 1) * (int)sizeof(S), row_off);
                    }
                    gv[q][m] = g - ga;
                }
            }
#pragma unroll
            for (int q = 0; q < 4; q++) {
                if (h + q >= nc) break;
#pragma unroll
                for (int m = 0; m < 2; m++) {
                    if (m == 1 && tid >= WIN - BP_TX) break;
                    float p = pv[q][m], r = rv[q][m];
                    if constexpr (SENT && !std::is_same<S, float>::value) {
                        // narrow storage cannot hold -1e30: widen its empty
                        // marks (a slice without them has nothing to widen)
                        const float thr = empty_threshold<S>();
                        if (p <= thr) p = kSentinel;
                        if (r <= thr) r = kSentinel;
                    }
                    const float d = r - p;
                    // SENT: the extended value; else that minus i d
                    const int gi = SENT ? gv[q][m] : gv[q][m] - (tid + m * BP_TX);
                    pairs[h + q][tid + m * BP_TX] = make_float2(fmaf((float)gi, d, p), d);
                }
            }
        }
        __syncthreads();
        // ---- accumulate ----
        for (int j = 0; j < nc; j++) {
            const TiltRow tr = ldc(list + c0 + j);
            // {w0, class, u - w0 at the tile's first voxel}
            const int4 tw = ldc(tile_tab + (size_t)tr.tilt * tilt_step);
            const float2 sc = ldc(cs + tr.tilt);
            // both loads issued together, before the class test
            asm volatile("" ::"s"(tw.z), "s"(sc.x), "s"(sc.y));
            if (tw.y == 0) continue;
            // scale is 1 or 1/2: both products are exact
            const float cc = scale * sc.x, ss = scale * sc.y;
            const float u0 = fmaf(xf, cc, __int_as_float(tw.z));
            const float2* pj = pairs[j];
            if (tw.y == 1 && !SENT) {
                full_cnt += 1;
#pragma unroll
                for (int k = 0; k < TZ; k++) {
                    float u = fmaf((float)k, ss, u0);
                    int i = (int)u;
                    float2 pr = pj[i];
                    acc[k] += fmaf(u, pr.y, pr.x);
                }
            } else if (!SENT) {
                const cptr<int2> lh = lohi + (size_t)tr.tilt * vol_z;
#pragma unroll
                for (int k = 0; k < TZ; k++) {
                    const int2 r = ldc(lh + min(z0 + k, vol_z - 1));
                    float u = fmaf((float)k, ss, u0);
                    int i = (int)u;
                    float2 pr = pj[i];
                    float v = fmaf(u, pr.y, pr.x);
                    bool in = (x >= r.x) & (x <= r.y);
                    acc[k] += in ? v : 0.0f;
                    cnt[k] += in ? 1 : 0;
                }
            } else {
                const cptr<int2> lh = lohi + (size_t)tr.tilt * vol_z;
#pragma unroll
                for (int k = 0; k < TZ; k++) {
                    const int2 r = ldc(lh + min(z0 + k, vol_z - 1));
                    float u = fmaf((float)k, ss, u0);
                    int i = (int)u;
                    float f = __builtin_amdgcn_fractf(u);
                    float2 pr = pj[i];
                    float v = fmaf(f, pr.y, pr.x);
                    bool in = (x >= r.x) & (x <= r.y) & !(v <= kEmpty);
                    acc[k] += in ? v : 0.0f;
                    cnt[k] += in ? 1 : 0;
                }
            }
        }
        __syncthreads();
    }
    if (WBP) {
        // ---- mean over the tilts used, 0 when none (GBackProjWbp.cu:48-52) ----
#pragma unroll
        for (int k = 0; k < TZ; k++) {
            const int z = z0 + k;
            const int count = full_cnt + cnt[k];
            const float v = count > 0 ? acc[k] * __builtin_amdgcn_rcpf((float)count) : 0.0f;
            if (x < vol_x && z < vol_z) stb(vres, v, x * (int)sizeof(float), z * row_bytes);
        }
        return;
    }
    // ---- v = relax * sum / count + v, positivity, sentinel (GBackProj.cu:50-59) ----
    bool wrote_sent = false;
#pragma unroll
    for (int k = 0; k < TZ; k++) {
        const int z = z0 + k;
        const int count = full_cnt + cnt[k];
        float v = relax * acc[k] * __builtin_amdgcn_rcpf((float)count) + vold[k];
        const bool empty = !(v > kEmpty);
        v = empty ? kSentinel : (positivity ? fmaxf(v, 0.0f) : v);
        if (x < vol_x && z < vol_z && count > 0) {
            stb(vres, v, x * (int)sizeof(float), z * row_bytes);
            wrote_sent |= empty;
        }
    }
    if (wrote_sent) atomicOr(vflag + y, 1);
}

template <typename S, int TZ, int CH, bool WBP = false>
__device__ __forceinline__ void bp_tile(cptr<BpArgs> A, float2 (&pairs)[CH][bp_win(TZ)],
                                        int xt, int zt, int y)
{
    const cptr<int> sflag = as_cptr(A->sflag);
    if (sflag != nullptr && sflag[y] != 0)
        bp_tile_body<S, TZ, CH, WBP, true>(A, pairs, xt, zt, y);
    else
        bp_tile_body<S, TZ, CH, WBP, false>(A, pairs, xt, zt, y);
}

template <typename S, int TZ, int CH>
__global__ __launch_bounds__(BP_TX) void k_back_proj(const BpArgs* __restrict__ a)
{
    __shared__ float2 pairs[CH][bp_win(TZ)];
    bp_tile<S, TZ, CH>(as_cptr(a), pairs, blockIdx.x, blockIdx.y, blockIdx.z);
}

// WBP back-projection onto the 2x grid (Recon/GBackProjWbp.cu:14-53): the
// same tile machinery, coordinate u = (fx cos + fz sin)/2 + P/2 + 1/2, the
// mean written to the grid buffer.
template <typename S>
__global__ __launch_bounds__(BP_TX) void k_wbp_grid_bp(const BpArgs* __restrict__ a)
{
    __shared__ float2 pairs[BP_CHUNK][bp_win(BP_TZ)];
    bp_tile<S, BP_TZ, BP_CHUNK, true>(as_cptr(a), pairs, blockIdx.x, blockIdx.y, blockIdx.z);
}

// Back-projection of many series in persistent launches: blocks take
// (series, slice, run of BP_ITEM_TILES tiles) items from a queue; entries
// are ordered by their first item.  One launch runs one tile shape (0: 16 z rows x 16 tilts
// per stage, 1: 16 x 8, 2: 8 x 16, 3: 8 x 8), so each shape's kernel holds
// only its own code and registers, as the per-series kernels do.
template <typename S, int SHAPE>
__global__ __launch_bounds__(BP_TX) void k_back_proj_batch(const BpItem* __restrict__ e, int n_e,
                                                           int total, int* __restrict__ counter)
{
    constexpr int TZ = (SHAPE & 2) ? 8 : 16, CH = (SHAPE & 1) ? 8 : 16;
    __shared__ float2 pairs[CH][bp_win(TZ)];
    const cptr<BpItem> E = as_cptr(e);
    int s = 0;
    for (int it = publish_item(claim_item(counter)); it < total;) {
        const int ahead = claim_item(counter);
        s = entry_of(E, n_e, it, s);
        const int local = it - E[s].item0;
        const int per_y = E[s].n_items_y;
        const int y = local / per_y, q = local - y * per_y;
        const int nx = E[s].a.n_xtiles, nt = E[s].a.n_ztiles * nx;
        for (int g = 0; g < BP_ITEM_TILES; g++) {
            const int r = q * BP_ITEM_TILES + g;
            if (r >= nt) break;
            const int zt = r / nx, xt = r - zt * nx;
            bp_tile<S, TZ, CH>(&E[s].a, pairs, xt, zt, y);
        }
        it = publish_item(ahead);
    }
}

// ------------------------------------------------------------------------
// Volume helpers: zero / load the SART state, per-slice fill statistics,
// output with the fill of MaUtil/GFillEmpty2D.cu:60-88 (CTomoSart.cpp:86-89)
// and the z-flip / XYZ layouts of CDoSartRecon.cpp:253-269 and
// DataUtil/CTiltSeries.cpp:286-305.
// ------------------------------------------------------------------------
__global__ void k_fill_zero(float* __restrict__ p, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0.0f;
}

__global__ void k_fill_zero_int(int* __restrict__ p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// SLICE-layout caller volume -> state, flagging slices that hold empties.
__global__ __launch_bounds__(256) void k_load_volume(const float* __restrict__ in,
                                                     float* __restrict__ vol, int vol_x,
                                                     int vol_z, int* __restrict__ vflag)
{
    const int y = blockIdx.y;
    const size_t n = (size_t)vol_x * vol_z;
    bool any = false;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        float v = in[(size_t)y * n + i];
        vol[(size_t)y * n + i] = v;
        any |= v < kEmpty;
    }
    if (any) atomicOr(vflag + y, 1);
}

// (sum, count) of non-empty voxels of each flagged slice, fixed order.
__global__ __launch_bounds__(256) void k_vol_stats(const float* __restrict__ vol, int vol_x,
                                                   int vol_z, const int* __restrict__ vflag,
                                                   double2* __restrict__ stats)
{
    const int y = blockIdx.x;
    if (vflag[y] == 0) return;
    const size_t n = (size_t)vol_x * vol_z;
    const float* v = vol + (size_t)y * n;
    double sum = 0.0, cnt = 0.0;
    for (size_t i = threadIdx.x; i < n; i += blockDim.x) {
        float a = v[i];
        if (!(a < kEmpty)) {
            sum += a;
            cnt += 1.0;
        }
    }
    __shared__ double s_sum[256], s_cnt[256];
    s_sum[threadIdx.x] = sum;
    s_cnt[threadIdx.x] = cnt;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            s_sum[threadIdx.x] += s_sum[threadIdx.x + o];
            s_cnt[threadIdx.x] += s_cnt[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) stats[y] = make_double2(s_sum[0], s_cnt[0]);
}

__device__ __forceinline__ size_t out_index(int layout, int x, int y, int z, int vol_x,
                                            int vol_y, int vol_z)
{
    switch (layout) {
    case ATR_LAYOUT_XYZ: return ((size_t)z * vol_y + y) * vol_x + x;
    case ATR_LAYOUT_XYZ_NOFLIP: return ((size_t)(vol_z - 1 - z) * vol_y + y) * vol_x + x;
    case ATR_LAYOUT_SLICE: return ((size_t)y * vol_z + z) * vol_x + x;
    default: return ((size_t)y * vol_z + (vol_z - 1 - z)) * vol_x + x;
    }
}

// grid (x-blocks, z, y)
__global__ __launch_bounds__(256) void k_output(const float* __restrict__ vol, int vol_x,
                                                int vol_y, int vol_z, const int* __restrict__ vflag,
                                                const double2* __restrict__ stats,
                                                const int* __restrict__ table, int layout,
                                                float* __restrict__ out)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int z = blockIdx.y, y = blockIdx.z;
    if (x >= vol_x) return;
    const float* slice = vol + (size_t)y * vol_z * vol_x;
    float v = slice[(size_t)z * vol_x + x];
    if (vflag[y] != 0 && v < kEmpty) {
        double2 st2 = stats[y];
        float mean = st2.y > 0 ? (float)(st2.x / st2.y) : 0.0f;
        v = filled_value(slice, z * vol_x + x, table, mean);
    }
    out[out_index(layout, x, y, z, vol_x, vol_y, vol_z)] = v;
}

// Plain streaming copy (16 B per lane, grid-stride, four independent loads
// in flight per lane): the calibration kernel for HBM byte counters and the
// attainable-bandwidth reference.
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src,
                                              float4* __restrict__ dst, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        const float4 a = src[i], b = src[i + stride], c = src[i + 2 * stride],
                     d = src[i + 3 * stride];
        dst[i] = a;
        dst[i + stride] = b;
        dst[i + 2 * stride] = c;
        dst[i + 3 * stride] = d;
    }
    for (; i < n; i += stride) dst[i] = src[i];
}

void launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s)
{
    size_t n = bytes / sizeof(float4);
    if (n == 0) return;
    unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, s, (const float4*)src, (float4*)dst, n);
}

// ------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------
void launch_prestage(const Plan& p, const float* in, hipStream_t s)
{
    dim3 grid(p.aln_y, p.n);
    hipLaunchKernelGGL(k_prestage, grid, dim3(256), sizeof(float) * 5 * (size_t)p.n_patches, s, in,
                       p.in_tilt_stride, p.in_row_stride, p.in_x, p.in_y,
                       (const double2*)p.d_axis_cs, (const float2*)p.d_shifts, p.d_local,
                       p.n_patches, p.d_aligned, p.aln_x, p.aln_y, (double2*)p.d_row_stats);
}

void launch_tilt_stats(const Plan& p, hipStream_t s)
{
    hipLaunchKernelGGL(k_tilt_stats, dim3(p.n), dim3(256), 0, s, (const double2*)p.d_row_stats,
                       p.aln_y, (double2*)p.d_tilt_stats);
}

template <typename S>
static void launch_weight_t(const Plan& p, const float* in, hipStream_t s)
{
    const StageSource src0 = stage_source(p, in);
    const float* src = src0.src;
    const size_t ts = src0.tilt_stride, rs = src0.row_stride;
    dim3 grid((p.sino_stride + 255) / 256, p.n, p.vol_y);
    hipLaunchKernelGGL(k_weight<S>, grid, dim3(256), 0, s, src, ts, rs, src0.fill ? 1 : 0,
                       p.d_fill_aln, (const double2*)p.d_tilt_stats, p.d_cs, p.n, p.proj,
                       p.vol_z, (S*)p.d_sino, p.sino_stride, p.d_sflag);
}

void launch_weight_sart(const Plan& p, const float* in, hipStream_t s)
{
    if (p.storage == ATR_STORAGE_F16) launch_weight_t<_Float16>(p, in, s);
    else if (p.storage == ATR_STORAGE_BF16) launch_weight_t<bf16_t>(p, in, s);
    else launch_weight_t<float>(p, in, s);
}

BpArgs bp_args(const Plan& p, const void* sino, int rows, const TiltRow* list, int n_list,
               float relax, bool check_sent)
{
    BpArgs a{};
    a.sino = sino;
    a.rows_per_slice = rows;
    a.stride = p.sino_stride;
    a.proj = p.proj;
    a.list = list;
    a.n_list = n_list;
    a.cs = p.d_cs;
    a.lohi = p.d_bp_lohi;
    a.tiles = p.d_bp_tiles;
    a.n_ztiles = (p.vol_z + p.bp_tz - 1) / p.bp_tz;
    a.n_xtiles = (p.vol_x + BP_TX - 1) / BP_TX;
    a.relax = relax;
    a.positivity = p.positivity;
    a.scale = 1.0f;
    a.offset = (float)(p.proj / 2.0);
    a.sflag = check_sent ? p.d_sflag : nullptr;
    a.vol = p.d_vol;
    a.vol_x = p.vol_x;
    a.vol_z = p.vol_z;
    a.vflag = p.d_vflag;
    return a;
}

template <typename S, int TZ, int CH>
static void launch_bp_v(const Plan& p, int step, hipStream_t s)
{
    const int nxt = (p.vol_x + BP_TX - 1) / BP_TX, nzt = (p.vol_z + TZ - 1) / TZ;
    hipLaunchKernelGGL((k_back_proj<S, TZ, CH>), dim3(nxt, nzt, p.vol_y), dim3(BP_TX), 0, s,
                       p.d_bp_args + step);
}

// Back-projection tile shapes the autotuner chooses from: z rows per thread
// x tilts per LDS stage.
template <typename S>
static void launch_bp_t(const Plan& p, int step, hipStream_t s)
{
    if (p.bp_tz == 8 && p.bp_chunk == 8) launch_bp_v<S, 8, 8>(p, step, s);
    else if (p.bp_tz == 8) launch_bp_v<S, 8, 16>(p, step, s);
    else if (p.bp_chunk == 8) launch_bp_v<S, 16, 8>(p, step, s);
    else launch_bp_v<S, 16, 16>(p, step, s);
}

void launch_bp(const Plan& p, int step, hipStream_t s)
{
    if (p.storage == ATR_STORAGE_F16) launch_bp_t<_Float16>(p, step, s);
    else if (p.storage == ATR_STORAGE_BF16) launch_bp_t<bf16_t>(p, step, s);
    else launch_bp_t<float>(p, step, s);
}

void launch_wbp_bp(const Plan& p, int slab, int ny, hipStream_t s)
{
    const int nxt = (2 * p.vol_x + BP_TX - 1) / BP_TX, nzt = (2 * p.vol_z + BP_TZ - 1) / BP_TZ;
    dim3 grid(nxt, nzt, ny);
    const BpArgs* a = p.d_bp_args + slab;
    if (p.storage == ATR_STORAGE_F16) hipLaunchKernelGGL(k_wbp_grid_bp<_Float16>, grid, dim3(BP_TX), 0, s, a);
    else if (p.storage == ATR_STORAGE_BF16) hipLaunchKernelGGL(k_wbp_grid_bp<bf16_t>, grid, dim3(BP_TX), 0, s, a);
    else hipLaunchKernelGGL(k_wbp_grid_bp<float>, grid, dim3(BP_TX), 0, s, a);
}

std::vector<BpLaunch> lay_out_bp(const std::vector<QueueEntry<BpItem>>& ent, BpItem* out)
{
    std::vector<BpLaunch> launches;
    int first = 0;
    for (int shape = 0; shape < BP_SHAPES; shape++) {
        std::vector<QueueEntry<BpItem>> g;
        for (const auto& e : ent)
            if (e.item.shape == shape) g.push_back(e);
        if (g.empty()) continue;
        const int total = lay_out_entries(g, out + first);
        launches.push_back(BpLaunch{first, (int)g.size(), total, shape});
        first += (int)g.size();
    }
    return launches;
}

// items == nullptr: resident blocks per CU, else the launch
template <typename S, int SHAPE>
static int bp_batch_v(const BpLaunch* l, const BpItem* items, int* counter, int grid,
                      hipStream_t s)
{
    if (!items) {
        int nb = 0;
        const hipError_t e =
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_back_proj_batch<S, SHAPE>, BP_TX, 0);
        return (e == hipSuccess && nb > 0) ? nb : 1;
    }
    hipLaunchKernelGGL((k_back_proj_batch<S, SHAPE>), dim3(grid), dim3(BP_TX), 0, s, items,
                       l->n_entries, l->total, counter);
    return 0;
}

template <typename S>
static int bp_batch_s(const BpLaunch* l, const BpItem* items, int* counter, int grid,
                      hipStream_t s)
{
    switch (l->shape) {
    case 1: return bp_batch_v<S, 1>(l, items, counter, grid, s);
    case 2: return bp_batch_v<S, 2>(l, items, counter, grid, s);
    case 3: return bp_batch_v<S, 3>(l, items, counter, grid, s);
    default: return bp_batch_v<S, 0>(l, items, counter, grid, s);
    }
}

static int bp_batch(int storage, const BpLaunch* l, const BpItem* items, int* counter, int grid,
                    hipStream_t s)
{
    if (storage == ATR_STORAGE_F16) return bp_batch_s<_Float16>(l, items, counter, grid, s);
    if (storage == ATR_STORAGE_BF16) return bp_batch_s<bf16_t>(l, items, counter, grid, s);
    return bp_batch_s<float>(l, items, counter, grid, s);
}

void launch_bp_batch(int storage, const BpLaunch& l, const BpItem* items, int* counter, int grid,
                     hipStream_t s)
{
    (void)bp_batch(storage, &l, items, counter, std::max(1, std::min(grid, l.total)), s);
}

int bp_batch_grid(int storage, const BpLaunch& l)
{
    int dev = 0, cus = 1;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus = 1;
    return cus * bp_batch(storage, &l, nullptr, nullptr, 0, nullptr);
}

void launch_zero(float* ptr, size_t n, hipStream_t s)
{
    size_t blocks = std::min<size_t>((n + 255) / 256, 65536);
    if (blocks == 0) return;
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)blocks), dim3(256), 0, s, ptr, n);
}

void launch_zero_int(int* ptr, int n, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_zero_int, dim3((n + 255) / 256), dim3(256), 0, s, ptr, n);
}

void launch_load_volume(const Plan& p, const float* in, hipStream_t s)
{
    size_t n = (size_t)p.vol_x * p.vol_z;
    unsigned bx = (unsigned)std::min<size_t>((n + 255) / 256, 64);
    hipLaunchKernelGGL(k_load_volume, dim3(bx, p.vol_y), dim3(256), 0, s, in, p.d_vol, p.vol_x,
                       p.vol_z, p.d_vflag);
}

void launch_vol_stats(const Plan& p, hipStream_t s)
{
    hipLaunchKernelGGL(k_vol_stats, dim3(p.vol_y), dim3(256), 0, s, p.d_vol, p.vol_x, p.vol_z,
                       p.d_vflag, (double2*)p.d_vstats);
}

void launch_output(const Plan& p, float* out, hipStream_t s)
{
    dim3 grid((p.vol_x + 255) / 256, p.vol_z, p.vol_y);
    hipLaunchKernelGGL(k_output, grid, dim3(256), 0, s, p.d_vol, p.vol_x, p.vol_y, p.vol_z,
                       p.d_vflag, (const double2*)p.d_vstats, p.d_fill_vol, p.layout, out);
}

}  // namespace atr
