// kernels_wbp.hip -- weighted back-projection path (gfx950):
//   k_filter        weighting + ramp/Hamming filter, LDS row FFT
//                   (Recon/GWeightProjs.cu:7-19, Recon/GRWeight.cu:7-69)
//   k_wbp_rowfft    row FFT of two rows of the 2x grid (the back-projection
//                   onto that grid is k_wbp_grid_bp in kernels_sart.hip)
//                   (Recon/GBackProjWbp.cu:126)
//   k_wbp_colfft    column FFT, normalisation, crop, inverse column FFT
//                   (MaUtil/CCufft2D.cpp:65-80, MaUtil/GFourierResize2D.cu:7-39)
//   k_wbp_rowc2r    real inverse row FFT + output layout
//                   (MaUtil/CCufft2D.cpp:108-123, MaUtil/CPad2D.cpp:29-39,
//                    Recon/CDoWbpRecon.cpp:107-123)
// Scratch: d_wbp_grid [slab][2VolZ][2VolX], d_wbp_t [slab][2VolZ][VolX/2+1],
// d_wbp_u [slab][VolZ][VolX/2+1].
#include <algorithm>

#include "atr_internal.h"
#include "storage.h"

namespace atr {

__device__ __forceinline__ float filled_value_w(const float* __restrict__ img, int i,
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
// Fourier binning of the aligned stack (Correct/CBinStack.cpp:54-106,
// -AtBin > 1): per image, R2C over ax x ay normalised by 1/(ax ay), crop to
// bx x by (MaUtil/GFourierResize2D.cu:7-39), unnormalised C2R.
//   k_bin_rowfft  two image rows packed as one complex row, FFT of length
//                 ax, the columns kx < bx/2+1 of both half spectra kept
//   k_bin_colfft  per block of columns: FFT of length ay, normalisation,
//                 crop to by rows, inverse FFT of length by
//   k_bin_rowc2r  two binned rows: packed C2R of length bx
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bin_rowfft(const float* __restrict__ in, size_t tilt_stride,
                                                    size_t row_stride, int fill,
                                                    const int* __restrict__ table,
                                                    const double2* __restrict__ tilt_stats,
                                                    int ax, int ay, int kxb, FftDesc fd,
                                                    float2* __restrict__ tbuf)
{
    extern __shared__ float2 lds[];
    float2* a = lds;
    float2* b = lds + ax;
    const int t = blockIdx.y;
    const int y0 = 2 * blockIdx.x, y1 = y0 + 1;
    const bool has1 = y1 < ay;
    const float* img = in + (size_t)t * tilt_stride;
    float mean = 0.0f;
    if (fill) {
        double2 ts = tilt_stats[t];
        mean = ts.y > 0 ? (float)(ts.x / ts.y) : 0.0f;
    }
    for (int x = threadIdx.x; x < ax; x += blockDim.x) {
        float va, vb = 0.0f;
        if (fill) {
            va = filled_value_w(img, (int)(y0 * row_stride + x), table, mean);
            if (has1) vb = filled_value_w(img, (int)(y1 * row_stride + x), table, mean);
        } else {
            va = img[(size_t)y0 * row_stride + x];
            if (has1) vb = img[(size_t)y1 * row_stride + x];
        }
        a[x] = make_float2(va, vb);
    }
    __syncthreads();
    float2* r = lds_fft(a, b, fd, 1, -1.0f, threadIdx.x, blockDim.x);
    float2* ta = tbuf + ((size_t)t * ay + y0) * kxb;
    float2* tb = tbuf + ((size_t)t * ay + y1) * kxb;
    for (int k = threadIdx.x; k < kxb; k += blockDim.x) {
        float2 z1 = r[k];
        float2 z2 = r[(ax - k) % ax];
        ta[k] = make_float2(0.5f * (z1.x + z2.x), 0.5f * (z1.y - z2.y));
        if (has1) tb[k] = make_float2(0.5f * (z1.y + z2.y), -0.5f * (z1.x - z2.x));
    }
}

__global__ __launch_bounds__(256) void k_bin_colfft(const float2* __restrict__ tbuf,
                                                    float2* __restrict__ ubuf, int ax, int ay,
                                                    int by, int kxb, int kb, FftDesc fy,
                                                    FftDesc fby)
{
    extern __shared__ float2 lds[];
    float2* a = lds;
    float2* b = lds + kb * ay;
    const int k0 = blockIdx.x * kb, t = blockIdx.y;
    const int nk = min(kb, kxb - k0);
    const float2* src = tbuf + (size_t)t * ay * kxb;
    for (int i = threadIdx.x; i < kb * ay; i += blockDim.x) {
        int col = i % kb, z = i / kb;
        a[col * ay + z] = col < nk ? src[(size_t)z * kxb + k0 + col] : make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    float2* r = lds_fft(a, b, fy, kb, -1.0f, threadIdx.x, blockDim.x);
    float2* o = (r == a) ? b : a;
    const float norm = (float)(1.0 / ((double)ax * (double)ay));
    for (int i = threadIdx.x; i < kb * by; i += blockDim.x) {
        int col = i / by, j = i % by;
        int srcz = j <= by / 2 ? j : j - by + ay;
        float2 v = r[col * ay + srcz];
        o[col * by + j] = make_float2(v.x * norm, v.y * norm);
    }
    __syncthreads();
    float2* q = lds_fft(o, r, fby, kb, 1.0f, threadIdx.x, blockDim.x);
    float2* dst = ubuf + (size_t)t * by * kxb;
    for (int i = threadIdx.x; i < kb * by; i += blockDim.x) {
        int col = i % kb, z = i / kb;
        if (col < nk) dst[(size_t)z * kxb + k0 + col] = q[col * by + z];
    }
}

__global__ __launch_bounds__(256) void k_bin_rowc2r(const float2* __restrict__ ubuf, int bx, int by,
                                                    FftDesc fx, float* __restrict__ out)
{
    extern __shared__ float2 lds[];
    float2* a = lds;
    float2* b = lds + bx;
    const int kxb = bx / 2 + 1, half = bx / 2;
    const int t = blockIdx.y, ya = 2 * blockIdx.x, yb = ya + 1;
    const bool has_b = yb < by;
    const float2* ha = ubuf + ((size_t)t * by + ya) * kxb;
    const float2* hb = ubuf + ((size_t)t * by + (has_b ? yb : ya)) * kxb;
    for (int k = threadIdx.x; k < bx; k += blockDim.x) {
        float2 va, vb;
        if (k == 0 || k == half) {
            va = make_float2(ha[k].x, 0.0f);
            vb = make_float2(hb[k].x, 0.0f);
        } else if (k < half) {
            va = ha[k];
            vb = hb[k];
        } else {
            float2 u = ha[bx - k], w = hb[bx - k];
            va = make_float2(u.x, -u.y);
            vb = make_float2(w.x, -w.y);
        }
        if (!has_b) vb = make_float2(0.0f, 0.0f);
        a[k] = make_float2(va.x - vb.y, va.y + vb.x);
    }
    __syncthreads();
    float2* r = lds_fft(a, b, fx, 1, 1.0f, threadIdx.x, blockDim.x);
    float* oa = out + ((size_t)t * by + ya) * bx;
    for (int x = threadIdx.x; x < bx; x += blockDim.x) {
        oa[x] = r[x].x;
        if (has_b) oa[bx + x] = r[x].y;
    }
}

// One workgroup per (tilt, pair of rows): the two weighted rows are packed as
// real and imaginary parts of one complex row, transformed, scaled by the
// real even filter (which keeps the two parts separate), transformed back.
template <typename S>
__global__ __launch_bounds__(256) void k_filter(const float* __restrict__ in, size_t tilt_stride,
                                                size_t row_stride, int fill,
                                                const int* __restrict__ table,
                                                const double2* __restrict__ tilt_stats,
                                                const float2* __restrict__ cs, int n, int rows,
                                                int proj, int vol_z, FftDesc fd,
                                                const float* __restrict__ filt,
                                                S* __restrict__ sino, int stride,
                                                int* __restrict__ sflag)
{
    extern __shared__ float2 lds[];
    float2* a = lds;
    float2* b = lds + proj;
    const int t = blockIdx.y;
    const int y0 = 2 * blockIdx.x, y1 = y0 + 1;
    const bool has1 = y1 < rows;
    const float w = cs[t].x / (float)vol_z;
    const float* img = in + (size_t)t * tilt_stride;
    float mean = 0.0f;
    if (fill) {
        double2 ts = tilt_stats[t];
        mean = ts.y > 0 ? (float)(ts.x / ts.y) : 0.0f;
    }
    for (int x = threadIdx.x; x < proj; x += blockDim.x) {
        float va, vb = 0.0f;
        if (fill) {
            va = filled_value_w(img, (int)(y0 * row_stride + x), table, mean);
            if (has1) vb = filled_value_w(img, (int)(y1 * row_stride + x), table, mean);
        } else {
            va = img[(size_t)y0 * row_stride + x];
            if (has1) vb = img[(size_t)y1 * row_stride + x];
        }
        a[x] = make_float2(va * w, vb * w);
    }
    __syncthreads();
    float2* r = lds_fft(a, b, fd, 1, -1.0f, threadIdx.x, blockDim.x);
    for (int k = threadIdx.x; k < proj; k += blockDim.x) {
        int kk = k <= proj / 2 ? k : proj - k;
        float f = filt[kk];
        r[k] = make_float2(r[k].x * f, r[k].y * f);
    }
    __syncthreads();
    float2* o = (r == a) ? b : a;
    r = lds_fft(r, o, fd, 1, 1.0f, threadIdx.x, blockDim.x);
    S* d0 = sino + ((size_t)y0 * n + t) * stride;
    S* d1 = sino + ((size_t)y1 * n + t) * stride;
    bool sent = false;
    for (int x = threadIdx.x; x < stride; x += blockDim.x) {
        float va = x < proj ? r[x].x : 0.0f;
        float vb = x < proj ? r[x].y : 0.0f;
        sent |= va <= kEmpty;
        st<S>(d0 + x, va);
        if (has1) {
            sent |= vb <= kEmpty;
            st<S>(d1 + x, vb);
        }
    }
    if (sent) {
        atomicOr(sflag + y0, 1);
        if (has1) atomicOr(sflag + y1, 1);
    }
}

// One workgroup per (slice, pair of 2x-grid rows).  Every grid point is the
// mean of the linear interpolations of the tilts whose FP64 inclusion
// interval holds it (0 when none).  The two rows are transformed together as
// one complex row of length 2VolX; bins 0..VolX/2 of each are written.
// Row FFT of the 2x grid (Recon/GBackProjWbp.cu:126-130 forward R2C along x):
// two grid rows packed as one complex row, one LDS FFT of length 2 VolX, the
// two half spectra split apart.
__global__ __launch_bounds__(256) void k_wbp_rowfft(const float* __restrict__ grid, int vol_x,
                                                    int vol_z, FftDesc fd,
                                                    float2* __restrict__ tbuf)
{
    extern __shared__ float2 lds[];
    const int nx2 = 2 * vol_x, nz2 = 2 * vol_z, kxn = vol_x / 2 + 1;
    float2* a = lds;
    float2* b = lds + nx2;
    const int zp = blockIdx.x, ys = blockIdx.y;
    const int za = 2 * zp, zb = za + 1;
    const float* ra = grid + ((size_t)ys * nz2 + za) * nx2;
    const float* rb = ra + nx2;
    for (int x = threadIdx.x; x < nx2; x += blockDim.x) a[x] = make_float2(ra[x], rb[x]);
    __syncthreads();
    float2* r = lds_fft(a, b, fd, 1, -1.0f, threadIdx.x, blockDim.x);
    float2* ta = tbuf + ((size_t)ys * nz2 + za) * kxn;
    float2* tb = tbuf + ((size_t)ys * nz2 + zb) * kxn;
    for (int k = threadIdx.x; k < kxn; k += blockDim.x) {
        float2 z1 = r[k];
        float2 z2 = r[(nx2 - k) % nx2];
        // A = (Z[k] + conj(Z[N-k]))/2, B = (Z[k] - conj(Z[N-k]))/(2i)
        ta[k] = make_float2(0.5f * (z1.x + z2.x), 0.5f * (z1.y - z2.y));
        tb[k] = make_float2(0.5f * (z1.y + z2.y), -0.5f * (z1.x - z2.x));
    }
}

// One workgroup per (slice, block of kb spectrum columns): forward FFT along
// z (length 2VolZ), scale by 1/(2VolX*2VolZ) (MaUtil/CCufft2D.cpp:75-78),
// crop to VolZ rows (row j <= VolZ/2 from j, else from j+VolZ), inverse FFT
// along z (length VolZ).
__global__ __launch_bounds__(256) void k_wbp_colfft(const float2* __restrict__ tbuf,
                                                    float2* __restrict__ ubuf, int vol_x,
                                                    int vol_z, int kb, FftDesc fz2, FftDesc fz1)
{
    extern __shared__ float2 lds[];
    const int nz2 = 2 * vol_z, kxn = vol_x / 2 + 1;
    float2* a = lds;
    float2* b = lds + kb * nz2;
    const int k0 = blockIdx.x * kb, ys = blockIdx.y;
    const int nk = min(kb, kxn - k0);
    const float2* src = tbuf + (size_t)ys * nz2 * kxn;
    for (int i = threadIdx.x; i < kb * nz2; i += blockDim.x) {
        int col = i % kb, z = i / kb;
        a[col * nz2 + z] = col < nk ? src[(size_t)z * kxn + k0 + col] : make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    float2* r = lds_fft(a, b, fz2, kb, -1.0f, threadIdx.x, blockDim.x);
    float2* o = (r == a) ? b : a;
    const float norm = (float)(1.0 / ((double)(2 * vol_x) * (double)nz2));
    for (int i = threadIdx.x; i < kb * vol_z; i += blockDim.x) {
        int col = i / vol_z, j = i % vol_z;
        int srcz = j <= vol_z / 2 ? j : j + vol_z;
        float2 v = r[col * nz2 + srcz];
        o[col * vol_z + j] = make_float2(v.x * norm, v.y * norm);
    }
    __syncthreads();
    float2* q = lds_fft(o, r, fz1, kb, 1.0f, threadIdx.x, blockDim.x);
    float2* dst = ubuf + (size_t)ys * vol_z * kxn;
    for (int i = threadIdx.x; i < kb * vol_z; i += blockDim.x) {
        int col = i % kb, z = i / kb;
        if (col < nk) dst[(size_t)z * kxn + k0 + col] = q[col * vol_z + z];
    }
}

__device__ __forceinline__ size_t out_index_w(int layout, int x, int y, int z, int vol_x,
                                              int vol_y, int vol_z)
{
    switch (layout) {
    case ATR_LAYOUT_XYZ: return ((size_t)z * vol_y + y) * vol_x + x;
    case ATR_LAYOUT_XYZ_NOFLIP: return ((size_t)(vol_z - 1 - z) * vol_y + y) * vol_x + x;
    case ATR_LAYOUT_SLICE: return ((size_t)y * vol_z + z) * vol_x + x;
    default: return ((size_t)y * vol_z + (vol_z - 1 - z)) * vol_x + x;
    }
}

// One workgroup per (slice, pair of output rows): the unnormalised C2R of
// length VolX of two half spectra, packed as one complex inverse FFT.  The
// imaginary parts of the DC and Nyquist bins are ignored (numpy irfft).
__global__ __launch_bounds__(256) void k_wbp_rowc2r(const float2* __restrict__ ubuf, int vol_x,
                                                    int vol_y, int vol_z, int y0, FftDesc fx,
                                                    int layout, float* __restrict__ out)
{
    extern __shared__ float2 lds[];
    float2* a = lds;
    float2* b = lds + vol_x;
    const int kxn = vol_x / 2 + 1, half = vol_x / 2;
    const int zp = blockIdx.x, ys = blockIdx.y, y = y0 + ys;
    const int za = 2 * zp, zb = za + 1;
    const bool has_b = zb < vol_z;
    const float2* ha = ubuf + ((size_t)ys * vol_z + za) * kxn;
    const float2* hb = ubuf + ((size_t)ys * vol_z + (has_b ? zb : za)) * kxn;
    for (int k = threadIdx.x; k < vol_x; k += blockDim.x) {
        float2 va, vb;
        if (k == 0 || k == half) {
            va = make_float2(ha[k].x, 0.0f);
            vb = make_float2(hb[k].x, 0.0f);
        } else if (k < half) {
            va = ha[k];
            vb = hb[k];
        } else {
            float2 u = ha[vol_x - k], w = hb[vol_x - k];
            va = make_float2(u.x, -u.y);
            vb = make_float2(w.x, -w.y);
        }
        if (!has_b) vb = make_float2(0.0f, 0.0f);
        a[k] = make_float2(va.x - vb.y, va.y + vb.x);
    }
    __syncthreads();
    float2* r = lds_fft(a, b, fx, 1, 1.0f, threadIdx.x, blockDim.x);
    for (int x = threadIdx.x; x < vol_x; x += blockDim.x) {
        out[out_index_w(layout, x, y, za, vol_x, vol_y, vol_z)] = r[x].x;
        if (has_b) out[out_index_w(layout, x, y, zb, vol_x, vol_y, vol_z)] = r[x].y;
    }
}

// ------------------------------------------------------------------------
template <typename S>
static void launch_filter_t(const Plan& p, const float* in, hipStream_t s)
{
    const StageSource src0 = stage_source(p, in);
    const float* src = src0.src;
    const size_t ts = src0.tilt_stride, rs = src0.row_stride;
    dim3 grid((p.vol_y + 1) / 2, p.n);
    size_t shm = 2 * (size_t)p.proj * sizeof(float2);
    hipLaunchKernelGGL(k_filter<S>, grid, dim3(256), shm, s, src, ts, rs, src0.fill ? 1 : 0,
                       p.d_fill_aln, (const double2*)p.d_tilt_stats, p.d_cs, p.n, p.vol_y, p.proj,
                       p.vol_z, p.fft_proj, p.d_filter, (S*)p.d_sino, p.sino_stride, p.d_sflag);
}

void launch_filter_wbp(const Plan& p, const float* in, hipStream_t s)
{
    if (p.storage == ATR_STORAGE_F16) launch_filter_t<_Float16>(p, in, s);
    else if (p.storage == ATR_STORAGE_BF16) launch_filter_t<bf16_t>(p, in, s);
    else launch_filter_t<float>(p, in, s);
}

void launch_wbp_rowfft(const Plan& p, int ny, hipStream_t s)
{
    dim3 grid(p.vol_z, ny);
    size_t shm = 2 * (size_t)(2 * p.vol_x) * sizeof(float2);
    hipLaunchKernelGGL(k_wbp_rowfft, grid, dim3(256), shm, s, p.d_wbp_grid, p.vol_x, p.vol_z,
                       p.fft_x2, p.d_wbp_t);
}

int wbp_col_block(int vol_z)
{
    int kb = 2048 / (2 * vol_z);
    if (kb < 1) kb = 1;
    if (kb > 16) kb = 16;
    return kb;
}

void launch_wbp_colfft(const Plan& p, int y0, int ny, hipStream_t s)
{
    (void)y0;
    int kb = wbp_col_block(p.vol_z);
    int kxn = p.vol_x / 2 + 1;
    dim3 grid((kxn + kb - 1) / kb, ny);
    size_t shm = 2 * (size_t)kb * (2 * p.vol_z) * sizeof(float2);
    hipLaunchKernelGGL(k_wbp_colfft, grid, dim3(256), shm, s, p.d_wbp_t, p.d_wbp_u, p.vol_x,
                       p.vol_z, kb, p.fft_z2, p.fft_z);
}

void launch_wbp_rowc2r(const Plan& p, int y0, int ny, float* out, hipStream_t s)
{
    dim3 grid((p.vol_z + 1) / 2, ny);
    size_t shm = 2 * (size_t)p.vol_x * sizeof(float2);
    hipLaunchKernelGGL(k_wbp_rowc2r, grid, dim3(256), shm, s, p.d_wbp_u, p.vol_x, p.vol_y,
                       p.vol_z, y0, p.fft_x, p.layout, out);
}

// Fourier binning of the stage input into p.d_binned.
void launch_bin(const Plan& p, const float* in, hipStream_t s)
{
    const float* src = p.prestage ? p.d_aligned : in;
    size_t ts = p.prestage ? (size_t)p.aln_x * p.aln_y : p.in_tilt_stride;
    size_t rs = p.prestage ? (size_t)p.aln_x : p.in_row_stride;
    const int ax = p.aln_x, ay = p.aln_y, bx = p.bin_x, by = p.bin_y;
    const int kxb = bx / 2 + 1;
    hipLaunchKernelGGL(k_bin_rowfft, dim3((ay + 1) / 2, p.n), dim3(256), 2 * (size_t)ax * sizeof(float2),
                       s, src, ts, rs, p.fill ? 1 : 0, p.d_fill_aln, (const double2*)p.d_tilt_stats,
                       ax, ay, kxb, p.fft_bin_ax, p.d_bin_t);
    const int kb = bin_col_block(ay);
    hipLaunchKernelGGL(k_bin_colfft, dim3((kxb + kb - 1) / kb, p.n), dim3(256),
                       2 * (size_t)kb * ay * sizeof(float2), s, p.d_bin_t, p.d_bin_u, ax, ay, by,
                       kxb, kb, p.fft_bin_ay, p.fft_bin_by);
    hipLaunchKernelGGL(k_bin_rowc2r, dim3((by + 1) / 2, p.n), dim3(256), 2 * (size_t)bx * sizeof(float2),
                       s, p.d_bin_u, bx, by, p.fft_bin_bx, p.d_binned);
}

int bin_col_block(int ay)
{
    int kb = 4096 / ay;
    return std::max(1, std::min(kb, 16));
}

}  // namespace atr
