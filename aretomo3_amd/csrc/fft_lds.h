// fft_lds.h -- workgroup FFT on complex data held in LDS (gfx950).
//
// Replaces the cuFFT plans of the reference path (MaUtil/GFFT1D.cu:45-88 for
// the row filter, MaUtil/CCufft2D.cpp:23-123 for the WBP Fourier crop).
// Stockham auto-sort formulation: each stage reads `src`, writes `dst` in
// natural-order-preserving positions, so no bit reversal is needed and the
// result of an s-stage transform sits in (s even ? a : b).  Radices 4, 2, 3
// and 5 have register butterflies; any other prime factor p is a direct
// length-p DFT per output (p complex products each), so every size up to
// 8192 is accepted and sizes with large prime factors cost O(n p).
// Twiddles come from a table W[e] = (cos 2 pi e/n, sin 2 pi e/n) computed in
// double on the host and rounded to float.  All arithmetic is FP32.
#pragma once
#include <hip/hip_runtime.h>

#define ATR_FFT_MAX_STAGES 16

struct FftDesc {
    int n;
    int n_stages;
    int radix[ATR_FFT_MAX_STAGES];
    const float2* tw;  // n entries
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by sign * i
__device__ __forceinline__ float2 cmul_si(float2 a, float sign) { return make_float2(-sign * a.y, sign * a.x); }

template <int R>
__device__ __forceinline__ void dft_small(float2* v, float sign)
{
    if constexpr (R == 2) {
        float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 3) {
        const float h = 0.86602540378443864676f;
        float2 t1 = cadd(v[1], v[2]);
        float2 t2 = make_float2(v[0].x - 0.5f * t1.x, v[0].y - 0.5f * t1.y);
        float2 d = cscale(csub(v[1], v[2]), h);
        float2 id = cmul_si(d, sign);
        v[0] = cadd(v[0], t1);
        v[1] = cadd(t2, id);
        v[2] = csub(t2, id);
    } else if constexpr (R == 4) {
        float2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        float2 s13 = cadd(v[1], v[3]), d13 = cmul_si(csub(v[1], v[3]), sign);
        v[0] = cadd(s02, s13);
        v[2] = csub(s02, s13);
        v[1] = cadd(d02, d13);
        v[3] = csub(d02, d13);
    } else if constexpr (R == 5) {
        const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
        const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
        float2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
        float2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
        float2 b1 = make_float2(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
        float2 b2 = make_float2(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
        float2 d1 = make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y);
        float2 d2 = make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y);
        float2 i1 = cmul_si(d1, sign), i2 = cmul_si(d2, sign);
        v[0] = make_float2(v[0].x + t1.x + t2.x, v[0].y + t1.y + t2.y);
        v[1] = cadd(b1, i1);
        v[4] = csub(b1, i1);
        v[2] = cadd(b2, i2);
        v[3] = csub(b2, i2);
    }
}

template <int R>
__device__ __forceinline__ void fft_stage(const float2* __restrict__ src, float2* __restrict__ dst,
                                          int n, int batch, int ns, const float2* __restrict__ tw,
                                          float sign, int tid, int nthreads)
{
    const int m = n / R;          // butterflies per sequence
    const int step = n / (ns * R);
    const int total = batch * m;
    for (int idx = tid; idx < total; idx += nthreads) {
        int bi = idx / m;
        int j = idx - bi * m;
        const float2* s = src + bi * n;
        float2* d = dst + bi * n;
        int k = j % ns;
        float2 v[R];
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = s[j + r * m];
        if (ns > 1) {
#pragma unroll
            for (int r = 1; r < R; r++) {
                float2 w = tw[r * k * step];
                w.y *= sign;
                v[r] = cmul(v[r], w);
            }
        }
        dft_small<R>(v, sign);
        int o = (j / ns) * ns * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) d[o + r * ns] = v[r];
    }
}

// Stage of an arbitrary radix R (a prime above 5), one output per thread:
// output q of butterfly (j, k = j mod ns) is sum_r src[j + r m] W^(r c),
// c = k n/(ns R) + q m, which folds the inter-stage twiddle W^(r k step)
// and the length-R DFT kernel W_R^(q r) = W^(q r m) into one table index.
__device__ __forceinline__ void fft_stage_any(const float2* __restrict__ src, float2* __restrict__ dst,
                                              int n, int batch, int ns, int R,
                                              const float2* __restrict__ tw, float sign, int tid,
                                              int nthreads)
{
    const int m = n / R;
    const int step = n / (ns * R);
    const int total = batch * n;
    for (int idx = tid; idx < total; idx += nthreads) {
        int bi = idx / n;
        int pos = idx - bi * n;
        int k = pos % ns;
        int t = pos / ns;
        int q = t % R;
        int j = (t / R) * ns + k;
        const float2* s = src + bi * n;
        int c = k * step + q * m;
        int e = 0;
        float2 acc = make_float2(0.0f, 0.0f);
        for (int r = 0; r < R; r++) {
            float2 w = tw[e];
            w.y *= sign;
            float2 v = s[j + r * m];
            acc.x += v.x * w.x - v.y * w.y;
            acc.y += v.x * w.y + v.y * w.x;
            e += c;
            if (e >= n) e -= n;
        }
        dst[bi * n + pos] = acc;
    }
}

// Transforms `batch` sequences of length fd.n stored contiguously in `a`
// (b is scratch of the same size).  sign = -1 forward, +1 inverse
// (unnormalised).  Returns the buffer holding the result.  Must be called by
// every thread of the workgroup; ends with a barrier.
__device__ __forceinline__ float2* lds_fft(float2* a, float2* b, const FftDesc& fd, int batch,
                                           float sign, int tid, int nthreads)
{
    int ns = 1;
    float2* src = a;
    float2* dst = b;
    for (int s = 0; s < fd.n_stages; s++) {
        int r = fd.radix[s];
        if (r == 4) fft_stage<4>(src, dst, fd.n, batch, ns, fd.tw, sign, tid, nthreads);
        else if (r == 2) fft_stage<2>(src, dst, fd.n, batch, ns, fd.tw, sign, tid, nthreads);
        else if (r == 3) fft_stage<3>(src, dst, fd.n, batch, ns, fd.tw, sign, tid, nthreads);
        else if (r == 5) fft_stage<5>(src, dst, fd.n, batch, ns, fd.tw, sign, tid, nthreads);
        else fft_stage_any(src, dst, fd.n, batch, ns, r, fd.tw, sign, tid, nthreads);
        __syncthreads();
        ns *= r;
        float2* t = src;
        src = dst;
        dst = t;
    }
    return src;
}
