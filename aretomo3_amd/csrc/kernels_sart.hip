// kernels_sart.hip -- pre-stage, weighting, SART back-projection, forward
// projection + difference, and the volume fill/output kernels (gfx950).
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

namespace atr {

// A value every lane holds alike, moved to an SGPR so that the compiler
// treats what is derived from it as wave-uniform.
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

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
            if (tw.y == 0) continue;
            const float2 sc = ldc(cs + tr.tilt);
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

// Block-wide work queue of the persistent launches.  Thread 0 claims item
// indices with an atomic on the queue counter; every thread sees the value
// through LDS.  The item after the current one is claimed when the current
// one starts, so the atomic's round trip overlaps the item's own loads
// instead of stalling the block between items.  Indices are handed out in
// increasing order, each to exactly one block, and a block leaves when it
// holds an index past the end (every block therefore exits).
__device__ __forceinline__ int publish_item(int v)
{
    __shared__ int s_item;
    if (threadIdx.x == 0) s_item = v;
    __syncthreads();
    const int it = uniform(s_item);
    __syncthreads();
    return it;
}

__device__ __forceinline__ int claim_item(int* __restrict__ counter)
{
    return threadIdx.x == 0 ? atomicAdd(counter, 1) : 0;
}

// Entry of queue index `it`.  Entries are ordered by their first item and a
// block's indices only grow, so the search resumes where the last one ended.
template <typename Item>
__device__ __forceinline__ int entry_of(cptr<Item> E, int n_e, int it, int s)
{
    while (s + 1 < n_e && E[s + 1].item0 <= it) s++;
    return s;
}

template <int TZ, int CH>
__device__ __forceinline__ float2 (&bp_pairs(float2* lds))[CH][bp_win(TZ)]
{
    return *reinterpret_cast<float2(*)[CH][bp_win(TZ)]>(lds);
}

// Back-projection of many series in one persistent launch: blocks take
// (series, slice, z-tile, x-tile) items from a queue; entries are ordered by
// their first item.  Each entry runs its own series' tile shape.
template <typename S>
__global__ __launch_bounds__(BP_TX) void k_back_proj_batch(const BpItem* __restrict__ e, int n_e,
                                                           int total, int* __restrict__ counter)
{
    __shared__ float2 lds[BP_CHUNK * bp_win(BP_TZ)];   // the largest shape's stage
    const cptr<BpItem> E = as_cptr(e);
    int s = 0;
    for (int it = publish_item(claim_item(counter)); it < total;) {
        const int ahead = claim_item(counter);
        s = entry_of(E, n_e, it, s);
        const int local = it - E[s].item0;
        const int per_y = E[s].n_items_y;
        const int y = local / per_y, r = local - y * per_y;
        const int zt = r / E[s].a.n_xtiles, xt = r - zt * E[s].a.n_xtiles;
        const cptr<BpArgs> A = &E[s].a;
        switch (E[s].shape) {
        case 1: bp_tile<S, 16, 8>(A, bp_pairs<16, 8>(lds), xt, zt, y); break;
        case 2: bp_tile<S, 8, 16>(A, bp_pairs<8, 16>(lds), xt, zt, y); break;
        case 3: bp_tile<S, 8, 8>(A, bp_pairs<8, 8>(lds), xt, zt, y); break;
        default: bp_tile<S, 16, 16>(A, bp_pairs<16, 16>(lds), xt, zt, y); break;
        }
        it = publish_item(ahead);
    }
}

// ------------------------------------------------------------------------
// Forward projection + difference (Recon/GForProj.cu:10-76 and
// Recon/GDiffProj.cu:7-21 fused).
//
// One workgroup (FP_THREADS) per (slice, group of tilts).  The slice is
// streamed through LDS in z-bands of bz cell rows; band entry (z, e) holds the
// interpolation pair (V[z][x] - e dV, dV), dV = V[z][x+1]-V[z][x], x = e - 1
// (slices with empty voxels: (V[z][x], dV)), so a bilinear sample is two
// ds_read_b64 (rows z and z+1) and three FMAs.  Entry 0 of a row is a guard for
// x in [-1, 0) (linear extrapolation), which lets the entry index be a plain
// truncation of x + 1.  The host assigns the group's (tilt, 64-ray tile)
// pairs to the FP_WAVES waves (FP_MAXP slots each) so that the per-band loads
// of the waves match (fp_pair_schedule); one ray per lane and one FP32
// accumulator per pair.  FP64 tables give, per (tilt, tile, band), the
// samples whose cell row lies in the band and whether every ray of the tile
// uses all of them; only the other blocks test each sample against the ray's
// FP64 interval.  Samples are processed four at a time with their geometry
// and cells loaded before any is used.  The difference raw - fwd (0 where
// either is empty) is written for the back-projection.
// ------------------------------------------------------------------------
// Geometry of four consecutive samples, fetched with one 64-byte scalar load.
struct Geo4 {
    float4 g[4];
};

// The two cells of a bilinear sample: g.x is the byte offset of the
// sample's lower row in the band buffer, xs = x + 1 (the +1 skips the guard
// pair), rowb the byte pitch of a band row.
//
// QF (slices without empty voxels): entry e holds (V_e - e dV_e, dV_e), the
// value at xs is then (V_e - e dV_e) + xs dV_e and the slope's weight is xs
// itself.  Otherwise entry e holds (V_e, dV_e) and the weight is fract(xs),
// so that the sentinel scale of an empty voxel reaches the value unchanged.
struct FpCells {
    float2 e0, e1;  // entries of the lower and upper row
    float f;        // weight of the slopes: xs (QF) or its fraction
};

template <bool QF>
__device__ __forceinline__ FpCells fp_fetch(const char* __restrict__ band, float4 g, float xs,
                                            int rowb)
{
    const int ix = (int)xs;
    const char* p = band + __float_as_int(g.x) + 8 * ix;
    FpCells c;
    c.f = QF ? xs : __builtin_amdgcn_fractf(xs);
    c.e0 = *(const float2*)p;
    c.e1 = *(const float2*)(p + rowb);
    return c;
}

// Bilinear value: top + wz (bot - top), g.y = wz.
__device__ __forceinline__ float fp_value(const FpCells& c, float4 g)
{
    const float top = fmaf(c.f, c.e0.y, c.e0.x);
    const float bot = fmaf(c.f, c.e1.y, c.e1.x);
    return fmaf(g.y, bot - top, top);
}

// acc + (1 - wz) top + wz bot, g.w = 1 - wz: the unmasked form, one VALU
// fewer than forming the value first.
__device__ __forceinline__ float fp_accumulate(const FpCells& c, float4 g, float acc)
{
    const float top = fmaf(c.f, c.e0.y, c.e0.x);
    const float bot = fmaf(c.f, c.e1.y, c.e1.x);
    return fmaf(g.y, bot, fmaf(g.w, top, acc));
}

// Samples [0, count) of one (pair, band): geometry entries G[0..count),
// sample indices k0 + m.  Groups of four (one 64-byte scalar load of
// geometry, the eight LDS reads issued before any is used), then the
// remainder one sample at a time.  MASKED clamps the coordinate and tests
// each sample against the ray's FP64 interval rr = (first, count).
//
// ROWB > 0 is the band row pitch in bytes known at compile time: the upper
// row's read then takes it as the LDS instruction's offset instead of an
// address add.
template <bool MASKED, bool SENT, int ROWB = 0>
__device__ __forceinline__ void fp_run(const char* __restrict__ band, int rowb_rt,
                                       cptr<float4> G, int count, int k0,
                                       float tf1, float xmax1, int2 rr, float& acc, int& cnt)
{
    const int rowb = ROWB > 0 ? ROWB : rowb_rt;
    int m = 0;
    for (; m + 4 <= count; m += 4) {
        const Geo4 g4 = ldc((cptr<Geo4>)(G + m));
        FpCells c[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            float xs = tf1 - g4.g[q].z;  // x + 1
            if (MASKED) xs = __builtin_amdgcn_fmed3f(xs, 1.0f, xmax1);
            c[q] = fp_fetch<!SENT>(band, g4.g[q], xs, rowb);
        }
        // keep the reads of the group ahead of their uses: scheduled next to
        // each use, every sample would wait out a full LDS latency
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (MASKED) {
                const float v = fp_value(c[q], g4.g[q]);
                bool in = (unsigned)(k0 + m + q - rr.x) < (unsigned)rr.y;
                if (SENT) in = in & !(v < kEmpty);
                acc += in ? v : 0.0f;
                if (SENT) cnt += in ? 1 : 0;
            } else {
                acc = fp_accumulate(c[q], g4.g[q], acc);
            }
        }
    }
    // the last 1-3 samples one at a time: no lane computes a sample past
    // the window
    for (; m < count; m++) {
        const float4 g = ldc(G + m);
        float xs = tf1 - g.z;
        if (MASKED) xs = __builtin_amdgcn_fmed3f(xs, 1.0f, xmax1);
        const FpCells c = fp_fetch<!SENT>(band, g, xs, rowb);
        if (MASKED) {
            const float v = fp_value(c, g);
            bool in = (unsigned)(k0 + m - rr.x) < (unsigned)rr.y;
            if (SENT) in = in & !(v < kEmpty);
            acc += in ? v : 0.0f;
            if (SENT) cnt += in ? 1 : 0;
        } else {
            acc = fp_accumulate(c, g, acc);
        }
    }
}

// One (group of tilts, slice) work item.  Returns early, for the whole
// block, when the slice's sentinel flag does not match SENT.  Ends after a
// barrier, with the LDS free for the next item.
// PF: the band-prefetching variant.  Wave FP_WAVES-1 has no pairs; while the
// other waves sample band b, it streams band b+1's volume rows into the raw
// buffer after the pair rows with direct global->LDS loads (one wave
// instruction = 256 floats of a row; no registers, and the sampling waves'
// own loads never wait on them: the memory counter is per wave).  All waves
// then form the pairs from the raw rows, LDS to LDS, with the same
// arithmetic as the direct staging, so both variants give the same pairs.
__device__ __forceinline__ void fp_raw_load(const ATR_GLOBAL_AS float* __restrict__ V, float* raw,
                                            int vol_x, int zb, int rows, int lane)
{
    const int per_row = vol_x >> 8;
    const int chunks = rows * per_row;
    for (int c = 0; c < chunks; c++) {
        const int r = c / per_row, k = c - r * per_row;
        const ATR_GLOBAL_AS float* src = V + (size_t)(zb + r) * vol_x + k * 256 + lane * 4;
        float* dst = raw + r * vol_x + k * 256;
        __builtin_amdgcn_global_load_lds((const ATR_GLOBAL_AS void*)src,
                                         (__attribute__((address_space(3))) void*)dst, 16, 0, 0);
    }
}

// RP > 0: the register-prefetching variant (VolX 1024, 2048 without empty
// voxels; RP = columns per thread, 1 or 2).  Each thread loads the
// next band's rows of its columns into registers right after the current
// band's pairs are in LDS, so the loads are in flight while the band is
// sampled; the pairs are formed from the registers after the band.  Entry
// VolX (x = VolX - 1 with dV = 0) is written by the owner of entry VolX - 1
// from its second value, so no thread stages a column of its own beyond RP.
template <typename S, bool SENT, int ROWB = 0, bool PF = false, int RP = 0>
__device__ __forceinline__ void fp_block(cptr<FpArgs> A, float2* __restrict__ band, int gx, int y)
{
    const gptr<const float> vol = as_gptr(A->vol);
    const int vol_x = A->vol_x, vol_z = A->vol_z;
    const cptr<int> fp_list = as_cptr(A->fp_list);
    const int group = A->group;
    const cptr<int> sched = as_cptr(A->sched);
    const cptr<float2> cs = as_cptr(A->cs);
    const gptr<const int2> fp_range = as_gptr(A->fp_range);
    const cptr<float4> geo = as_cptr(A->geo);
    const cptr<int4> fp_desc = as_cptr(A->fp_desc);
    const int nb = A->nb, bz = A->bz, proj = A->proj;
    const gptr<const S> sino = (gptr<const S>)A->sino;
    const int n = A->n, stride = A->stride;
    const gptr<S> diff = (gptr<S>)A->diff;
    const int diff_rows = A->diff_rows;
    if ((as_cptr(A->vflag)[y] != 0) != SENT) return;
    const int tid = threadIdx.x, lane = tid & 63;
    // wave index made provably uniform so that all per-pair table reads are
    // scalar loads into SGPRs
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int j0 = gx * group;
    const int ntile = (proj + FP_WAVE - 1) / FP_WAVE;
    // this wave's (local tilt << 16 | tile) slots, -1 = unused
    const cptr<int> pair_tab = sched + (size_t)gx * (FP_WAVES * FP_MAXP) + wave * FP_MAXP;
    const int pitch = vol_x + 1;               // pairs per band row (guard + VolX)
    const int rowb = pitch * (int)sizeof(float2);
    const float xmax1 = (float)(vol_x - 1);    // clamp of x + 1 for masked samples
    const gptr<const float> V = vol + (size_t)y * vol_z * vol_x;
    // the slice as a buffer resource (raw, 32-bit records)
    const __amdgpu_buffer_rsrc_t vrsrc = __builtin_amdgcn_make_buffer_rsrc(
        (void*)V, (short)0, vol_z * vol_x * (int)sizeof(float), 0x00020000);
    float acc[FP_MAXP], tf1[FP_MAXP];
    int cnt[FP_MAXP];
#pragma unroll
    for (int i = 0; i < FP_MAXP; i++) {
        acc[i] = 0.0f;
        cnt[i] = 0;
        tf1[i] = 1.0f;
        const int jt = pair_tab[i];
        if (jt >= 0) {
            const int t = fp_list[j0 + (jt >> 16)];
            const int xd = min((jt & 0xffff) * FP_WAVE + lane, proj - 1);
            tf1[i] = (float)(((double)xd - 0.5 * proj) / (double)cs[t].x + vol_x * 0.5 + 1.0);
        }
    }
    const char* bandb = (const char*)band;
    // PF: raw rows of the next band, from the first 256-byte boundary after
    // the pair rows; loaded by the last wave
    float* raw = (float*)(bandb + ((size_t)(bz + 1) * pitch * sizeof(float2) + 255) / 256 * 256);
    const bool loader = PF && wave == FP_WAVES - 1;
    // Work items start at different bands, so that the band reloads of the
    // workgroups on the chip are spread out in time instead of all hitting
    // HBM at once.  (A ray's samples are summed band by band in this order.)
    const int rot = (int)(((unsigned)y * 5u + (unsigned)gx) % (unsigned)nb);
    // RP: rows of the band held in registers, per column c: pa = V[z][xa],
    // pb = V[z][xb] of column tid + c FP_THREADS
    constexpr int RPC = RP > 0 ? RP : 1;
    constexpr int RP_ROWS = RP > 0 ? (FP_BZ_MAX + 1) / RP : 1;
    float pa[RPC][RP_ROWS], pb[RPC][RP_ROWS];
    auto rp_load = [&](int zb_n, int rows_n, auto full) {
        constexpr bool FULL = decltype(full)::value;
        const int zoff = zb_n * vol_x * (int)sizeof(float);
#pragma unroll
        for (int c = 0; c < RPC; c++) {
            const int col = tid + c * FP_THREADS;
            // columns 0 .. VolX - 1 (the threads past them load nothing)
            const int xa = max(col - 1, 0), xb = min(xa + 1, vol_x - 1);
            if (col >= vol_x) continue;
#pragma unroll
            for (int r = 0; r < RP_ROWS; r++) {
                if (FULL || r < rows_n) {
                    const int so = zoff + r * vol_x * (int)sizeof(float);
                    pa[c][r] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(
                        vrsrc, xa * (int)sizeof(float), so, 0));
                    pb[c][r] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(
                        vrsrc, xb * (int)sizeof(float), so, 0));
                }
            }
        }
    };
    auto rp_load_band = [&](int b_n) {
        const int zb_n = __builtin_amdgcn_readfirstlane(b_n * bz);
        const int rows_n = __builtin_amdgcn_readfirstlane(min(bz, vol_z - 1 - zb_n) + 1);
        if (rows_n == RP_ROWS) rp_load(zb_n, rows_n, std::true_type{});
        else rp_load(zb_n, rows_n, std::false_type{});
    };
    auto rp_write = [&](int rows_n, auto full) {
        constexpr bool FULL = decltype(full)::value;
#pragma unroll
        for (int c = 0; c < RPC; c++) {
            const int col = tid + c * FP_THREADS;
            if (col < vol_x) {
                const float ecol = (float)max(col, 1);
                char* wp = (char*)band + col * (int)sizeof(float2);
                const bool last = col == vol_x - 1;
#pragma unroll
                for (int r = 0; r < RP_ROWS; r++) {
                    if (FULL || r < rows_n) {
                        const float dx = pb[c][r] - pa[c][r];
                        *(float2*)(wp + r * ROWB) = make_float2(fmaf(-ecol, dx, pa[c][r]), dx);
                        if (last) *(float2*)(wp + r * ROWB + 8) = make_float2(pb[c][r], 0.0f);
                    }
                }
            }
        }
    };
    if (RP) rp_load_band(rot);
    if (PF) {
        if (loader) {
            const int zb = rot * bz;
            fp_raw_load(V, raw, vol_x, zb, min(bz, vol_z - 1 - zb) + 1, lane);
            __builtin_amdgcn_s_waitcnt(0);
        }
        __syncthreads();
    }
    for (int bi = 0; bi < nb; bi++) {
        const int b = bi + rot < nb ? bi + rot : bi + rot - nb;
        const int zb = b * bz;
        const int rows = min(bz, vol_z - 1 - zb) + 1;  // volume rows zb .. zb+rows-1
        if (PF) {
            // pairs from the raw rows (LDS to LDS), then the next band's loads
            for (int col = tid; col < pitch; col += FP_THREADS) {
                const int xi = col - 1;
                const int xa = max(xi, 0), xb = min(xa + 1, vol_x - 1);
                for (int r = 0; r < rows; r++) {
                    const float va = raw[r * vol_x + xa], vb = raw[r * vol_x + xb];
                    const float dx = vb - va;
                    const float v0 = xi < 0 ? va - dx : va;
                    band[r * pitch + col] = make_float2(SENT ? v0 : fmaf(-(float)col, dx, v0), dx);
                }
            }
            __syncthreads();
            if (loader && bi + 1 < nb) {
                const int b1 = b + 1 < nb ? b + 1 : 0;
                const int zb1 = b1 * bz;
                fp_raw_load(V, raw, vol_x, zb1, min(bz, vol_z - 1 - zb1) + 1, lane);
            }
        }
        // Staging: a thread owns whole columns of interpolation pairs
        // (V[z][x], V[z][x+1]-V[z][x]); column 0 is the guard for x in [-1, 0).
        // Each half of the rows is loaded first (independent, coalesced
        // loads), then the pairs are formed from registers.  The loads are
        // buffer loads over the slice: the lane's byte offset is fixed for
        // the launch and the row offset is a scalar, so they cost no address
        // VALU (and a load past the slice would return 0, not fault).
        // Entry e stores (V - e dV, dV); the guard (e = 0, x = -1) holds the
        // extrapolated V[0] - dV, which is V[0] - 1 dV: the same formula with
        // e taken as 1.
        if (RP) {
            const int rows_u = __builtin_amdgcn_readfirstlane(rows);
            if (rows_u == RP_ROWS) rp_write(rows_u, std::true_type{});
            else rp_write(rows_u, std::false_type{});
            __syncthreads();
            if (bi + 1 < nb) rp_load_band(b + 1 < nb ? b + 1 : 0);
        } else if (!PF) {
            const int rows_u = __builtin_amdgcn_readfirstlane(rows);
            const int zoff = __builtin_amdgcn_readfirstlane(zb) * vol_x * (int)sizeof(float);
            const int rowb_s = ROWB > 0 ? ROWB : rowb;
            for (int col = tid; col < pitch; col += FP_THREADS) {
                const int xi = col - 1;                      // -1 .. VolX-1
                const int xa = max(xi, 0), xb = min(xa + 1, vol_x - 1);
                const int oa = xa * (int)sizeof(float), ob = xb * (int)sizeof(float);
                const float ecol = (float)max(col, 1);
                char* wp = (char*)band + col * (int)sizeof(float2);
                // a batch of FP_STAGE_ROWS rows starting at row h; FULL: all
                // of them exist (no per-row test)
                auto batch = [&](int h, auto full) {
                    constexpr bool FULL = decltype(full)::value;
                    float va[FP_STAGE_ROWS], vb[FP_STAGE_ROWS];
#pragma unroll
                    for (int r = 0; r < FP_STAGE_ROWS; r++) {
                        if (FULL || h + r < rows_u) {
                            const int so = zoff + (h + r) * vol_x * (int)sizeof(float);
                            va[r] = __uint_as_float(
                                __builtin_amdgcn_raw_buffer_load_b32(vrsrc, oa, so, 0));
                            vb[r] = __uint_as_float(
                                __builtin_amdgcn_raw_buffer_load_b32(vrsrc, ob, so, 0));
                        }
                    }
#pragma unroll
                    for (int r = 0; r < FP_STAGE_ROWS; r++) {
                        if (FULL || h + r < rows_u) {
                            const float dx = vb[r] - va[r];
                            // SENT: (V, dV), the guard (V[0] - dV, dV)
                            const float v0 = SENT ? (xi < 0 ? va[r] - dx : va[r])
                                                  : fmaf(-ecol, dx, va[r]);
                            *(float2*)(wp + (h + r) * rowb_s) = make_float2(v0, dx);
                        }
                    }
                };
#pragma unroll
                for (int h = 0; h < FP_BZ_MAX + 1; h += FP_STAGE_ROWS) {
                    if (h + FP_STAGE_ROWS <= rows_u) batch(h, std::true_type{});
                    else if (h < rows_u) batch(h, std::false_type{});
                }
            }
        }
        if (!PF && !RP) __syncthreads();
#pragma unroll
        for (int i = 0; i < FP_MAXP; i++) {
            const int jt = pair_tab[i];
            if (jt >= 0) {
                const int t = fp_list[j0 + (jt >> 16)];
                const int tile = jt & 0xffff;
                const int4 d = ldc(fp_desc + ((size_t)t * ntile + tile) * nb + b);
                if (d.y > 0) {
                    const cptr<float4> G = geo + d.x;
                    if (!SENT && d.w != 0) {
                        // the window's samples summed from zero, then added
                        // to the ray's sum: a fresh loop-carried register
                        // instead of one lane of the 16-register sum array
                        float part = 0.0f;
                        int unused = 0;
                        fp_run<false, false, ROWB>(bandb, rowb, G, d.y, d.z, tf1[i], xmax1,
                                                   make_int2(0, 0), part, unused);
                        acc[i] += part;
                    } else {
                        int xd = tile * FP_WAVE + lane;
                        // keep the interval load inside the band loop: hoisted,
                        // the intervals of all pairs would pin 2*FP_MAXP VGPRs
                        asm volatile("" : "+v"(xd));
                        int2 rr = make_int2(0, 0);
                        if (xd < proj) rr = ldgo(fp_range + (size_t)t * proj + xd);
                        float part = 0.0f;
                        int cpart = 0;
                        fp_run<true, SENT, ROWB>(bandb, rowb, G, d.y, d.z, tf1[i], xmax1, rr, part,
                                                 cpart);
                        acc[i] += part;
                        if (SENT) cnt[i] += cpart;
                    }
                }
            }
        }
        // PF: the next band's rows are in LDS once the loader's loads are done
        if (loader) __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < FP_MAXP; i++) {
        const int jt = pair_tab[i];
        if (jt >= 0) {
            const int j = j0 + (jt >> 16);
            const int t = fp_list[j];
            const int xd = (jt & 0xffff) * FP_WAVE + lane;
            if (xd < proj) {
                const int count = SENT ? cnt[i] : fp_range[(size_t)t * proj + xd].y;
                const float fwd = count <= 0 ? kSentinel : acc[i] / (float)count;
                float raw = ldg(sino + ((size_t)y * n + t) * stride + xd);
                if (raw <= empty_threshold<S>()) raw = kSentinel;
                const float d = (fwd < kEmpty || raw < kEmpty) ? 0.0f : raw - fwd;
                stg(diff + ((size_t)y * diff_rows + j) * stride + xd, d);
            }
        }
    }
}

// ROWB: the band row pitch in bytes, 8 (VolX + 1), when the launch fixes it
// at compile time (VolX 512, 1024, 2048), else 0.
// PF: the band-prefetching variant (VolX a multiple of 256).
// RP: the register-prefetching variant (columns per thread).
template <typename S, bool SENT, int ROWB, bool PF = false, int RP = 0>
__global__ __launch_bounds__(FP_THREADS) void k_for_proj_diff(const FpArgs* __restrict__ a)
{
    extern __shared__ float2 band[];
    fp_block<S, SENT, ROWB, PF, RP>(as_cptr(a), band, blockIdx.x, blockIdx.y);
}

// Forward projection + difference of many series in one persistent launch:
// blocks take (series, slice, tilt group) items from a queue.  A slice with
// empty voxels runs the sentinel variant, the others the variant with the
// band pitch fixed at compile time for VolX 512 / 1024 / 2048 (as the
// per-series launch does), so one launch covers every slice.
template <typename S>
__global__ __launch_bounds__(FP_THREADS) void k_for_proj_diff_batch(const FpItem* __restrict__ e,
                                                                   int n_e, int total,
                                                                   int* __restrict__ counter)
{
    extern __shared__ float2 band[];
    const cptr<FpItem> E = as_cptr(e);
    int s = 0;
    for (int it = publish_item(claim_item(counter)); it < total;) {
        const int ahead = claim_item(counter);
        s = entry_of(E, n_e, it, s);
        const int local = it - E[s].item0;
        const int ng = E[s].n_groups;
        const int y = local / ng, gx = local - y * ng;
        const cptr<FpArgs> A = &E[s].a;
        if (as_cptr(A->vflag)[y] != 0) {
            fp_block<S, true>(A, band, gx, y);
        } else {
            // the staging mode of the series' own plan (its band rows and
            // wave schedule were built for it), so a series computes here
            // exactly what it computes alone
            if (A->pf == FP_PF_DMA) {
                switch (A->vol_x) {
                case 512: fp_block<S, false, 513 * 8, true>(A, band, gx, y); break;
                case 1024: fp_block<S, false, 1025 * 8, true>(A, band, gx, y); break;
                case 2048: fp_block<S, false, 2049 * 8, true>(A, band, gx, y); break;
                default: fp_block<S, false, 0, true>(A, band, gx, y); break;
                }
            } else {
                const bool rp = A->pf == FP_PF_REG;
                switch (A->vol_x) {
                case 512: fp_block<S, false, 513 * 8>(A, band, gx, y); break;
                case 1024:
                    if (rp) fp_block<S, false, 1025 * 8, false, 1>(A, band, gx, y);
                    else fp_block<S, false, 1025 * 8>(A, band, gx, y);
                    break;
                case 2048:
                    if (rp) fp_block<S, false, 2049 * 8, false, 2>(A, band, gx, y);
                    else fp_block<S, false, 2049 * 8>(A, band, gx, y);
                    break;
                default: fp_block<S, false, 0>(A, band, gx, y); break;
                }
            }
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

void launch_bp_batch(int storage, const BpItem* items, int n_items, int total, int* counter,
                     int grid, hipStream_t s)
{
    if (storage == ATR_STORAGE_F16)
        hipLaunchKernelGGL(k_back_proj_batch<_Float16>, dim3(grid), dim3(BP_TX), 0, s, items, n_items, total, counter);
    else if (storage == ATR_STORAGE_BF16)
        hipLaunchKernelGGL(k_back_proj_batch<bf16_t>, dim3(grid), dim3(BP_TX), 0, s, items, n_items, total, counter);
    else
        hipLaunchKernelGGL(k_back_proj_batch<float>, dim3(grid), dim3(BP_TX), 0, s, items, n_items, total, counter);
}

FpArgs fp_args(const Plan& p, const int* fp_list, int n_list, const int* sched)
{
    FpArgs a{};
    a.vol = p.d_vol;
    a.vol_x = p.vol_x;
    a.vol_z = p.vol_z;
    a.vflag = p.d_vflag;
    a.fp_list = fp_list;
    a.n_list = n_list;
    a.group = fp_group(p.proj, n_list, fp_waves(p.fp_pf));
    a.pf = p.fp_pf;
    a.sched = sched;
    a.cs = p.d_cs;
    a.fp_range = p.d_fp_range;
    a.geo = p.d_fp_geo;
    a.fp_desc = p.d_fp_desc;
    a.nb = p.geo.fp_nbands;
    a.bz = p.geo.fp_bz;
    a.proj = p.proj;
    a.sino = p.d_sino;
    a.n = p.n;
    a.stride = p.sino_stride;
    a.diff = p.d_diff;
    a.diff_rows = p.max_subset;
    return a;
}

// Pair rows of a band; with the prefetch, then (from a 256-byte boundary) the
// raw rows of the next band.
static size_t fp_pair_bytes(int bz, int vol_x) { return (size_t)(bz + 1) * (vol_x + 1) * sizeof(float2); }

size_t fp_lds_bytes(int bz, int vol_x, int pf)
{
    const size_t pairs = fp_pair_bytes(bz, vol_x);
    if (pf != FP_PF_DMA) return pairs;
    return (pairs + 255) / 256 * 256 + (size_t)(bz + 1) * vol_x * sizeof(float);
}

// Kernels with more than 64 KiB of dynamic LDS must say so once per device;
// `done` holds one bit per device the attribute was set on (devices 64 and
// up set it at every launch).
template <typename K>
static void allow_big_lds(K kernel, std::atomic<unsigned long long>& done, int bytes = FP_LDS_BUDGET)
{
    int dev = -1;
    const bool known = hipGetDevice(&dev) == hipSuccess && dev >= 0 && dev < 64;
    const unsigned long long bit = known ? 1ull << dev : 0ull;
    if (known && (done.load(std::memory_order_relaxed) & bit)) return;
    const hipError_t e = hipFuncSetAttribute((const void*)kernel,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e == hipSuccess && known) done.fetch_or(bit, std::memory_order_relaxed);
    // a refusal surfaces as the launch's own error, not as a stale one
    if (e != hipSuccess) (void)hipGetLastError();
}

template <typename S, bool SENT, int ROWB, bool PF = false, int RP = 0>
static void launch_fp_k(dim3 grid, size_t shm, const FpArgs* a, hipStream_t s)
{
    static std::atomic<unsigned long long> done{0};
    allow_big_lds(k_for_proj_diff<S, SENT, ROWB, PF, RP>, done, PF ? FP_LDS_PF : FP_LDS_BUDGET);
    hipLaunchKernelGGL((k_for_proj_diff<S, SENT, ROWB, PF, RP>), grid, dim3(FP_THREADS), shm, s, a);
}

template <typename S>
static void launch_fp_t(const Plan& p, int subset, hipStream_t s)
{
    const int n_list = p.subset_fp_off[subset + 1] - p.subset_fp_off[subset];
    const int group = fp_group(p.proj, n_list, fp_waves(p.fp_pf));
    dim3 grid((n_list + group - 1) / group, p.vol_y);
    const size_t shm = fp_lds_bytes(p.geo.fp_bz, p.vol_x, p.fp_pf);
    const FpArgs* a = p.d_fp_args + subset;
    // the sentinel-free slices (nearly all) take a variant with the band
    // pitch fixed at compile time for the common widths
    if (p.fp_pf == FP_PF_REG) {   // VolX 1024 or 2048
        if (p.vol_x == 1024) launch_fp_k<S, false, 1025 * 8, false, 1>(grid, shm, a, s);
        else launch_fp_k<S, false, 2049 * 8, false, 2>(grid, shm, a, s);
        launch_fp_k<S, true, 0>(grid, shm, a, s);
        return;
    }
    if (p.fp_pf == FP_PF_DMA) {
        switch (p.vol_x) {
        case 512: launch_fp_k<S, false, 513 * 8, true>(grid, shm, a, s); break;
        case 1024: launch_fp_k<S, false, 1025 * 8, true>(grid, shm, a, s); break;
        case 2048: launch_fp_k<S, false, 2049 * 8, true>(grid, shm, a, s); break;
        default: launch_fp_k<S, false, 0, true>(grid, shm, a, s); break;
        }
        launch_fp_k<S, true, 0, true>(grid, shm, a, s);
        return;
    }
    switch (p.vol_x) {
    case 512: launch_fp_k<S, false, 513 * 8>(grid, shm, a, s); break;
    case 1024: launch_fp_k<S, false, 1025 * 8>(grid, shm, a, s); break;
    case 2048: launch_fp_k<S, false, 2049 * 8>(grid, shm, a, s); break;
    default: launch_fp_k<S, false, 0>(grid, shm, a, s); break;
    }
    launch_fp_k<S, true, 0>(grid, shm, a, s);
}

void launch_fp_diff(const Plan& p, int subset, hipStream_t s)
{
    if (p.storage == ATR_STORAGE_F16) launch_fp_t<_Float16>(p, subset, s);
    else if (p.storage == ATR_STORAGE_BF16) launch_fp_t<bf16_t>(p, subset, s);
    else launch_fp_t<float>(p, subset, s);
}

template <typename S>
static void fp_batch_attrs()
{
    static std::atomic<unsigned long long> done{0};
    allow_big_lds(k_for_proj_diff_batch<S>, done, FP_LDS_PF);   // the largest mode's LDS
}

template <typename S>
static void launch_fp_batch_t(const FpItem* items, int n_items, int total, int* counters,
                              size_t shm, int grid, hipStream_t s)
{
    fp_batch_attrs<S>();
    hipLaunchKernelGGL((k_for_proj_diff_batch<S>), dim3(grid), dim3(FP_THREADS), shm, s,
                       items, n_items, total, counters);
}

void launch_fp_batch(int storage, const FpItem* items, int n_items, int total, int* counters,
                     size_t shm, int grid, hipStream_t s)
{
    if (storage == ATR_STORAGE_F16) launch_fp_batch_t<_Float16>(items, n_items, total, counters, shm, grid, s);
    else if (storage == ATR_STORAGE_BF16) launch_fp_batch_t<bf16_t>(items, n_items, total, counters, shm, grid, s);
    else launch_fp_batch_t<float>(items, n_items, total, counters, shm, grid, s);
}

// Resident blocks per CU of the batched kernels (the persistent grid size).
template <typename S>
static int batch_blocks_t(size_t fp_shm, bool fp)
{
    fp_batch_attrs<S>();
    int nb = 0;
    hipError_t e = fp ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                            &nb, k_for_proj_diff_batch<S>, FP_THREADS, fp_shm)
                      : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_back_proj_batch<S>,
                                                                     BP_TX, 0);
    return (e == hipSuccess && nb > 0) ? nb : 1;
}

int batch_blocks_per_cu(int storage, bool fp, size_t fp_shm)
{
    if (storage == ATR_STORAGE_F16) return batch_blocks_t<_Float16>(fp_shm, fp);
    if (storage == ATR_STORAGE_BF16) return batch_blocks_t<bf16_t>(fp_shm, fp);
    return batch_blocks_t<float>(fp_shm, fp);
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
