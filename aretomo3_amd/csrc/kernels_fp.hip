// kernels_fp.hip -- SART forward projection + difference (gfx950): the
// per-series kernel, the batched persistent kernel and their launchers.
//
// Built as its own translation unit: the per-pair ray sums of a wave are
// registers that each (pair, band) window adds to, and with private arrays
// promoted to vector registers the compiler copied the whole set in and out
// of every window.  This file is compiled with that promotion off.
#include <atomic>
#include <climits>

#include "atr_internal.h"
#include "storage.h"
#include "work_queue.h"

namespace atr {

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
// next band's rows of its columns into registers while the current band is
// sampled (after the wave's first FP_RP_ISSUE pair slots), and the pairs are
// formed from the registers after the band.  Entry
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
    // this wave's slots: local tilt << 24 | pair index (tilt * ntile + tile,
    // the row of the pair's band descriptors), -1 = unused
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
            const int t = fp_list[j0 + (jt >> 24)];
            const int tile = (jt & 0xffffff) - t * ntile;
            const int xd = min(tile * FP_WAVE + lane, proj - 1);
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
    // Every one of the RP_ROWS rows is loaded (a short band repeats its last
    // row), so the loads' registers are the same on every path: a register
    // that one path loads and another does not would make the compiler wait
    // for the loads in flight wherever the paths join.
    auto rp_load_band = [&](int b_n) {
        const int zb_n = __builtin_amdgcn_readfirstlane(b_n * bz);
        const int last = __builtin_amdgcn_readfirstlane(min(bz, vol_z - 1 - zb_n));
        const int zoff = zb_n * vol_x * (int)sizeof(float);
#pragma unroll
        for (int c = 0; c < RPC; c++) {
            const int col = tid + c * FP_THREADS;
            // columns 0 .. VolX - 1 (a thread past them loads column VolX - 1
            // and writes nothing)
            const int cl = min(col, vol_x - 1);
            const int xa = max(cl - 1, 0), xb = min(xa + 1, vol_x - 1);
#pragma unroll
            for (int r = 0; r < RP_ROWS; r++) {
                const int so = zoff + min(r, last) * vol_x * (int)sizeof(float);
                pa[c][r] = __uint_as_float(
                    __builtin_amdgcn_raw_buffer_load_b32(vrsrc, xa * (int)sizeof(float), so, 0));
                pb[c][r] = __uint_as_float(
                    __builtin_amdgcn_raw_buffer_load_b32(vrsrc, xb * (int)sizeof(float), so, 0));
            }
        }
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
            // RP: the next band's loads are issued after the wave's first
            // FP_RP_ISSUE pair slots: a masked window's interval load waits
            // for every load issued before it, and the first slots' masked
            // windows then do not wait for the prefetch
            if (RP && i == FP_RP_ISSUE && bi + 1 < nb) rp_load_band(b + 1 < nb ? b + 1 : 0);
            const int jt = pair_tab[i];
            if (jt >= 0) {
                const int4 d = ldc(fp_desc + (size_t)(jt & 0xffffff) * nb + b);
                // all four fields in one scalar load, before the count test
                asm volatile("" ::"s"(d.x), "s"(d.w));
                if (d.y > 0) {
                    const cptr<float4> G = geo + d.x;
                    if (!SENT && d.w != 0) {
                        // the window's samples summed from zero, then added
                        // to the ray's sum
                        float part = 0.0f;
                        int unused = 0;
                        fp_run<false, false, ROWB>(bandb, rowb, G, d.y, d.z, tf1[i], xmax1,
                                                   make_int2(0, 0), part, unused);
                        acc[i] += part;
                    } else {
                        const int t = fp_list[j0 + (jt >> 24)];
                        const int tile = (jt & 0xffffff) - t * ntile;
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
            const int j = j0 + (jt >> 24);
            const int t = fp_list[j];
            const int xd = ((jt & 0xffffff) - t * ntile) * FP_WAVE + lane;
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

// Forward projection + difference of many series in persistent launches:
// blocks take (series, slice, tilt group) items from a queue.  One launch
// runs one staging variant (the staging mode and the band pitch fixed at
// compile time for VolX 512 / 1024 / 2048, as the per-series launch does),
// so each variant's kernel holds only its own code and registers; variant
// FP_VARIANT_EMPTY is the empty-voxel projector.  Every block of a work item
// whose slice does not match the variant's empty-voxel flag leaves at once.
template <typename S, int V>
__device__ __forceinline__ void fp_variant_block(cptr<FpArgs> A, float2* band, int gx, int y)
{
    // VolX class: 0: 512, 1: 1024, 2: 2048, 3: any (pitch at run time)
    constexpr int W = V < 8 ? V % 4 : V - 7;
    constexpr int ROWB = W == 0 ? 513 * 8 : W == 1 ? 1025 * 8 : W == 2 ? 2049 * 8 : 0;
    if constexpr (V == FP_VARIANT_EMPTY) fp_block<S, true>(A, band, gx, y);
    else if constexpr (V < 4) fp_block<S, false, ROWB, true>(A, band, gx, y);
    else if constexpr (V < 8) fp_block<S, false, ROWB>(A, band, gx, y);
    else fp_block<S, false, ROWB, false, W>(A, band, gx, y);   // RP: 1 or 2 columns
}

template <typename S, int V>
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
        fp_variant_block<S, V>(&E[s].a, band, gx, y);
        it = publish_item(ahead);
    }
}

// ------------------------------------------------------------------------
// Host side: arguments, LDS sizes, launches.
// ------------------------------------------------------------------------
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

// Staging variant of a series' plan (slices without empty voxels).
int fp_variant(const FpArgs& a)
{
    const int w = a.vol_x == 512 ? 0 : a.vol_x == 1024 ? 1 : a.vol_x == 2048 ? 2 : 3;
    if (a.pf == FP_PF_DMA) return w;
    if (a.pf == FP_PF_REG && (w == 1 || w == 2)) return 7 + w;
    return 4 + w;
}

std::vector<FpLaunch> lay_out_fp(const std::vector<QueueEntry<FpItem>>& ent, FpItem* out)
{
    std::vector<FpLaunch> launches;
    int first = 0;
    for (int v = 0; v < FP_VARIANTS; v++) {
        std::vector<QueueEntry<FpItem>> g;
        size_t shm = 0;
        for (const auto& e : ent) {
            const FpArgs& a = e.item.a;
            if (v != FP_VARIANT_EMPTY && fp_variant(a) != v) continue;
            g.push_back(e);
            shm = std::max(shm, fp_lds_bytes(a.bz, a.vol_x, v == FP_VARIANT_EMPTY ? 0 : a.pf));
        }
        if (g.empty()) continue;
        const int total = lay_out_entries(g, out + first);
        launches.push_back(FpLaunch{first, (int)g.size(), total, v, shm});
        first += (int)g.size();
    }
    return launches;
}

template <typename S, int V>
static int fp_batch_v(const FpLaunch* l, const FpItem* items, int* counter, int grid,
                      hipStream_t s)
{
    static std::atomic<unsigned long long> done{0};
    // the DMA variants' LDS holds the raw rows too (and the queue's word)
    allow_big_lds(k_for_proj_diff_batch<S, V>, done, V < 4 ? FP_LDS_PF : FP_LDS_BUDGET);
    if (!items) {
        // resident blocks per CU at this LDS size
        int nb = 0;
        const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, k_for_proj_diff_batch<S, V>, FP_THREADS, l->shm);
        return (e == hipSuccess && nb > 0) ? nb : 1;
    }
    hipLaunchKernelGGL((k_for_proj_diff_batch<S, V>), dim3(grid), dim3(FP_THREADS), l->shm, s,
                       items, l->n_entries, l->total, counter);
    return 0;
}

template <typename S, int V = 0>
static int fp_batch_s(const FpLaunch* l, const FpItem* items, int* counter, int grid,
                      hipStream_t s)
{
    if constexpr (V < FP_VARIANTS) {
        if (l->variant == V) return fp_batch_v<S, V>(l, items, counter, grid, s);
        return fp_batch_s<S, V + 1>(l, items, counter, grid, s);
    }
    return 1;
}

static int fp_batch(int storage, const FpLaunch* l, const FpItem* items, int* counter, int grid,
                    hipStream_t s)
{
    if (storage == ATR_STORAGE_F16) return fp_batch_s<_Float16>(l, items, counter, grid, s);
    if (storage == ATR_STORAGE_BF16) return fp_batch_s<bf16_t>(l, items, counter, grid, s);
    return fp_batch_s<float>(l, items, counter, grid, s);
}

void launch_fp_batch(int storage, const FpLaunch& l, const FpItem* items, int* counter, int grid,
                     hipStream_t s)
{
    (void)fp_batch(storage, &l, items, counter, std::max(1, std::min(grid, l.total)), s);
}

int fp_batch_grid(int storage, const FpLaunch& l)
{
    int dev = 0, cus = 1;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus = 1;
    return cus * fp_batch(storage, &l, nullptr, nullptr, 0, nullptr);
}

}  // namespace atr
