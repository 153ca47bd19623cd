/*
 * oracle.c -- FP64 CPU restatement of AreTomo3's reconstruction hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the reported CPU baseline).
 * The product never links or calls this file.
 *
 * Each function cites the reference file:line it restates (paths relative to
 * the AreTomo3 source tree).  Build: see oracle/Makefile (gcc -O2 -fopenmp
 * -ffp-contract=off).
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_EMPTY (-1e10)   /* sentinel threshold used throughout the reference */
#define ORC_SENT  (-1e30)   /* sentinel value the reference writes */

int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---------------------------------------------------------------------- */
/* Random index table for the empty-pixel fill.                            */
/* The reference draws it with cuRAND seeded by time(0)                    */
/* (MaUtil/GGenRandoms.cu:7-21,34-49), so its values are not reproducible. */
/* The rebuild makes the table an explicit seeded input:                   */
/*   r_i   = splitmix64(seed + i * 0xD1B54A32D192ED03)                      */
/*   tbl_i = ((r_i >> 32) * n) >> 32            (uniform in [0, n))         */
/* ---------------------------------------------------------------------- */
uint64_t orc_splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

void orc_fill_table(uint64_t seed, int n, int* table)
{
    for (int i = 0; i < n; i++) {
        uint64_t r = orc_splitmix64(seed + (uint64_t)i * 0xD1B54A32D192ED03ull);
        table[i] = (int)(((r >> 32) * (uint64_t)n) >> 32);
    }
}

/* Recon/CTomoBase.cpp:55-65: angle = 0.017453f * tilt (float product),
 * cos/sin evaluated in double and stored as float. */
void orc_cos_sin(const float* tilts_deg, int n, float* cos_sin)
{
    const float rad = 0.017453f;
    for (int i = 0; i < n; i++) {
        float ang = rad * tilts_deg[i];
        cos_sin[2 * i] = (float)cos((double)ang);
        cos_sin[2 * i + 1] = (float)sin((double)ang);
    }
}

/* AreTomo/Util/CSplitItems.cpp:25-48: equal blocks, the first (n % s)
 * blocks one larger, contiguous starts. */
void orc_split_items(int n_items, int n_splits, int* starts, int* sizes)
{
    int size = n_items / n_splits;
    for (int i = 0; i < n_splits; i++) sizes[i] = size;
    int left = n_items % n_splits;
    for (int i = 0; i < left; i++) sizes[i] += 1;
    starts[0] = 0;
    for (int i = 1; i < n_splits; i++) starts[i] = starts[i - 1] + sizes[i - 1];
}

/* Tilt indices of one SART subset.  order 0 = contiguous blocks exactly as
 * Recon/CTomoSart.cpp:60-61,74-80 (CSplitItems over [start, start+n)).
 * order 1 = interleaved: subset s = {start + s + k*n_subsets}. */
int orc_subset_tilts(int start_tilt, int n_tilts, int n_subsets, int order,
                     int subset, int* tilts)
{
    int n = 0;
    if (order == 0) {
        int* starts = (int*)malloc(sizeof(int) * n_subsets);
        int* sizes = (int*)malloc(sizeof(int) * n_subsets);
        orc_split_items(n_tilts, n_subsets, starts, sizes);
        for (int k = 0; k < sizes[subset]; k++)
            tilts[n++] = start_tilt + starts[subset] + k;
        free(starts);
        free(sizes);
    } else {
        for (int k = subset; k < n_tilts; k += n_subsets) tilts[n++] = start_tilt + k;
    }
    return n;
}

/* Correct/CCorrectUtil.cpp:11-21 (note the 3.14 there). */
void orc_aligned_size(int raw_x, int raw_y, float tilt_axis_deg, int* aln_x, int* aln_y)
{
    double rot = fabs(sin(tilt_axis_deg * 3.14 / 180.0));
    *aln_x = raw_x;
    *aln_y = raw_y;
    if (rot <= 0.707) return;
    *aln_x = raw_y;
    *aln_y = raw_x;
}

/* ---------------------------------------------------------------------- */
/* FP64 mixed-radix FFT (unnormalised).  sign = -1 forward, +1 inverse.    */
/* Used where the reference calls cuFFT (MaUtil/GFFT1D.cu:45-88,          */
/* MaUtil/CCufft2D.cpp:23-123).  Recursive decimation in time over the     */
/* smallest prime factor; twiddles from one double cos/sin table.          */
/* ---------------------------------------------------------------------- */
static int smallest_factor(int n)
{
    if (n % 2 == 0) return 2;
    for (int p = 3; p * p <= n; p += 2)
        if (n % p == 0) return p;
    return n;
}

static void fft_rec(const double* xr, const double* xi, int s, int n,
                    double* yr, double* yi, const double* wr, const double* wi,
                    int big_n, double* tr, double* ti)
{
    if (n == 1) {
        yr[0] = xr[0];
        yi[0] = xi[0];
        return;
    }
    int p = smallest_factor(n), m = n / p;
    for (int r = 0; r < p; r++)
        fft_rec(xr + (size_t)r * s, xi + (size_t)r * s, s * p, m, yr + r * m,
                yi + r * m, wr, wi, big_n, tr, ti);
    int step = big_n / n;
    for (int k = 0; k < n; k++) {
        double ar = 0.0, ai = 0.0;
        int km = k % m;
        for (int r = 0; r < p; r++) {
            long e = ((long)r * k) % n;
            double cr = wr[e * step], ci = wi[e * step];
            double vr = yr[r * m + km], vi = yi[r * m + km];
            ar += vr * cr - vi * ci;
            ai += vr * ci + vi * cr;
        }
        tr[k] = ar;
        ti[k] = ai;
    }
    memcpy(yr, tr, sizeof(double) * n);
    memcpy(yi, ti, sizeof(double) * n);
}

void orc_fft(double* re, double* im, int n, int sign)
{
    double* buf = (double*)malloc(sizeof(double) * 6 * (size_t)n);
    double *wr = buf, *wi = buf + n, *yr = buf + 2 * n, *yi = buf + 3 * n;
    double *tr = buf + 4 * n, *ti = buf + 5 * n;
    const double two_pi = 6.283185307179586476925286766559;
    for (int e = 0; e < n; e++) {
        double a = two_pi * (double)e / (double)n;
        wr[e] = cos(a);
        wi[e] = sign * sin(a);
    }
    fft_rec(re, im, 1, n, yr, yi, wr, wi, n, tr, ti);
    memcpy(re, yr, sizeof(double) * n);
    memcpy(im, yi, sizeof(double) * n);
    free(buf);
}

/* ---------------------------------------------------------------------- */
/* Weighting                                                               */
/* ---------------------------------------------------------------------- */

/* Recon/GWeightProjs.cu:7-19 (launch :41-47): s[t][x] *= cos_t / VolZ for
 * x < proj_x. */
void orc_weight_projs(double* sino, int n_tilts, int pad_x, int proj_x,
                      const float* cos_sin, int vol_z)
{
    for (int t = 0; t < n_tilts; t++) {
        double w = (double)cos_sin[2 * t] / (double)vol_z;
        for (int x = 0; x < proj_x; x++) sino[(size_t)t * pad_x + x] *= w;
    }
}

/* Recon/GRWeight.cu:7-23,57-69 with MaUtil/GFFT1D.cu:45-88: row R2C of
 * length fft_size, bin x in [0, N/2] scaled by
 *   (1/N) * 2 r (0.55 + 0.45 cos(6.2831852 r)),  r = x / N,
 * then unnormalised C2R.  The filter is real and even, so the complex FFT of
 * the real row followed by the symmetric scaling and a complex inverse gives
 * the same Hermitian C2R result. */
void orc_rweight(double* sino, int n_tilts, int pad_x, int fft_size)
{
    int n = fft_size;
    const double c2pi = (double)6.2831852f;
    double* w = (double*)malloc(sizeof(double) * (n / 2 + 1));
    for (int x = 0; x <= n / 2; x++) {
        double r = (double)x / (double)n;
        w[x] = 2.0 * r * (0.55 + 0.45 * cos(c2pi * r)) / (double)n;
    }
    double* re = (double*)malloc(sizeof(double) * n);
    double* im = (double*)malloc(sizeof(double) * n);
    for (int t = 0; t < n_tilts; t++) {
        double* row = sino + (size_t)t * pad_x;
        for (int x = 0; x < n; x++) {
            re[x] = row[x];
            im[x] = 0.0;
        }
        orc_fft(re, im, n, -1);
        for (int k = 0; k < n; k++) {
            int kk = k <= n / 2 ? k : n - k;
            re[k] *= w[kk];
            im[k] *= w[kk];
        }
        orc_fft(re, im, n, +1);
        for (int x = 0; x < n; x++) row[x] = re[x];
    }
    free(w);
    free(re);
    free(im);
}

/* ---------------------------------------------------------------------- */
/* Back-projection, forward projection, difference                         */
/* ---------------------------------------------------------------------- */

/* Recon/GBackProj.cu:12-60.  For each voxel: linear interpolation of every
 * listed, not-excluded tilt whose coordinate u = fx cos + fz sin + ProjX/2
 * lies in [0, ProjX-2]; samples <= -1e10 skipped; v += relax * sum / count;
 * positivity clamp (bSart); <= -1e10 -> -1e30; count 0 leaves v unchanged. */
void orc_back_proj(const double* sino, int pad_x, int proj_x,
                   const float* cos_sin, const unsigned char* no_proj,
                   const int* tilts, int n_list, int positivity, double relax,
                   double* vol, int vol_x, int vol_z)
{
    double cx = proj_x / 2.0;
    int end_x = proj_x - 2;
    for (int z = 0; z < vol_z; z++) {
        double fz = z - vol_z * 0.5;
        for (int x = 0; x < vol_x; x++) {
            double fx = x - vol_x * 0.5;
            double sum = 0.0;
            int count = 0;
            for (int k = 0; k < n_list; k++) {
                int t = tilts[k];
                if (no_proj != 0 && no_proj[t]) continue;
                double c = cos_sin[2 * t], s = cos_sin[2 * t + 1];
                double u = fx * c + fz * s + cx;
                if (u < 0 || u > end_x) continue;
                int iu = (int)u;
                double f = u - iu;
                const double* row = sino + (size_t)t * pad_x;
                double v = row[iu] * (1 - f) + row[iu + 1] * f;
                if (v <= ORC_EMPTY) continue;
                sum += v;
                count += 1;
            }
            if (count <= 0) continue;
            size_t i = (size_t)z * vol_x + x;
            double val = relax * sum / count + vol[i];
            if (val > ORC_EMPTY) vol[i] = positivity ? (val > 0 ? val : 0.0) : val;
            else vol[i] = ORC_SENT;
        }
    }
}

/* Recon/GForProj.cu:10-76: ray (tilt t, detector x) of L = (int)(VolZ/cos +
 * 1.5) unit steps centred on the slice; sample k at
 *   z = VolZ/2 + (k - L/2) cos,  x = (xd - ProjX/2)/cos + VolX/2 - (k - L/2) sin;
 * bilinear inside [0,VolX-2]x[0,VolZ-2]; samples < -1e10 skipped; mean of the
 * rest, -1e30 when none.  Output row t of fwd (pad_x stride). */
void orc_for_proj(const double* vol, int vol_x, int vol_z,
                  const float* cos_sin, const int* tilts, int n_list,
                  int proj_x, int pad_x, double* fwd)
{
    int end_x = vol_x - 2, end_z = vol_z - 2;
    for (int k = 0; k < n_list; k++) {
        int t = tilts[k];
        double c = cos_sin[2 * t], s = cos_sin[2 * t + 1];
        int len = (int)(vol_z / c + 1.5);
        for (int xd = 0; xd < proj_x; xd++) {
            double xp = xd - 0.5 * proj_x;
            double tx = xp / c + vol_x * 0.5;
            double tz = vol_z * 0.5;
            double sum = 0.0, count = 0.0;
            for (int i = 0; i < len; i++) {
                double fzk = i - len * 0.5;
                double x = tx - fzk * s;
                double z = tz + fzk * c;
                if (x < 0 || z < 0 || x > end_x || z > end_z) continue;
                int ix = (int)x, iz = (int)z;
                double ax = x - ix, az = z - iz;
                const double* r0 = vol + (size_t)iz * vol_x + ix;
                const double* r1 = r0 + vol_x;
                double v = r0[0] * (1 - ax) * (1 - az) + r0[1] * ax * (1 - az)
                         + r1[0] * (1 - ax) * az + r1[1] * ax * az;
                if (v < ORC_EMPTY) continue;
                sum += v;
                count += 1.0;
            }
            fwd[(size_t)t * pad_x + xd] = count <= 0 ? ORC_SENT : sum / count;
        }
    }
}

/* Recon/GDiffProj.cu:7-21 (used in place, CTomoSart.cpp:103-113):
 * d = raw - fwd, or 0 if either is < -1e10. */
void orc_diff_proj(const double* raw, double* fwd, const int* tilts,
                   int n_list, int proj_x, int pad_x)
{
    for (int k = 0; k < n_list; k++) {
        size_t base = (size_t)tilts[k] * pad_x;
        for (int x = 0; x < proj_x; x++) {
            size_t i = base + x;
            if (fwd[i] < ORC_EMPTY) fwd[i] = 0.0;
            else if (raw[i] < ORC_EMPTY) fwd[i] = 0.0;
            else fwd[i] = raw[i] - fwd[i];
        }
    }
}

/* MaUtil/GFillEmpty2D.cu:7-143: pixels < -1e10 are empty; each empty pixel
 * follows the random index chain (<= 100 hops) to the first non-empty pixel
 * and takes its value, else the mean of all non-empty pixels (0 if none). */
void orc_fill_empty(double* img, int w, int h, int pad_x, const int* table)
{
    int n = w * h;
    unsigned char* empty = (unsigned char*)malloc(n);
    double sum = 0.0, count = 0.0;
    for (int i = 0; i < n; i++) {
        double v = img[(size_t)(i / w) * pad_x + i % w];
        empty[i] = v < ORC_EMPTY;
        if (!empty[i]) {
            sum += v;
            count += 1.0;
        }
    }
    double mean = count <= 0 ? 0.0 : sum / count;
    for (int i = 0; i < n; i++) {
        if (!empty[i]) continue;
        double val = mean;
        int next = table[i];
        for (int j = 0; j < 100; j++) {
            if (empty[next]) {
                next = table[next];
                continue;
            }
            val = img[(size_t)(next / w) * pad_x + next % w];
            break;
        }
        img[(size_t)(i / w) * pad_x + i % w] = val;
    }
    free(empty);
}

/* ---------------------------------------------------------------------- */
/* SART (Recon/CTomoSart.cpp:54-125)                                        */
/* ---------------------------------------------------------------------- */
void orc_tomo_sart(double* sino, int n_tilts, int pad_x, int proj_x,
                   const float* cos_sin, const unsigned char* no_proj,
                   double* vol, int vol_x, int vol_z, int start_tilt,
                   int n_sart_tilts, int iters, int n_subsets, double relax,
                   int order, int positivity, const int* vol_fill_table)
{
    /* weighting of all projections (:65-66) */
    orc_weight_projs(sino, n_tilts, pad_x, proj_x, cos_sin, vol_z);
    /* initial back-projection of every projection, relax 1 (:68-69) */
    int* all = (int*)malloc(sizeof(int) * n_tilts);
    for (int t = 0; t < n_tilts; t++) all[t] = t;
    orc_back_proj(sino, pad_x, proj_x, cos_sin, no_proj, all, n_tilts,
                  positivity, 1.0, vol, vol_x, vol_z);
    free(all);
    /* relax = 1/nSubsets in the reference (:71) */
    double r = relax > 0 ? relax : 1.0 / n_subsets;
    double* fwd = (double*)calloc((size_t)n_tilts * pad_x, sizeof(double));
    int* list = (int*)malloc(sizeof(int) * (n_sart_tilts + 1));
    for (int it = 0; it < iters; it++) {
        for (int s = 0; s < n_subsets; s++) {
            int n = orc_subset_tilts(start_tilt, n_sart_tilts, n_subsets, order, s, list);
            orc_for_proj(vol, vol_x, vol_z, cos_sin, list, n, proj_x, pad_x, fwd);
            orc_diff_proj(sino, fwd, list, n, proj_x, pad_x);
            orc_back_proj(fwd, pad_x, proj_x, cos_sin, no_proj, list, n,
                          positivity, r, vol, vol_x, vol_z);
        }
    }
    free(list);
    free(fwd);
    /* GFillEmpty2D on the slice (:86-89) */
    orc_fill_empty(vol, vol_x, vol_z, vol_x, vol_fill_table);
}

/* ---------------------------------------------------------------------- */
/* WBP (Recon/CTomoWbp.cpp:38-60, Recon/GBackProjWbp.cu:14-143)            */
/* ---------------------------------------------------------------------- */

/* Recon/GBackProjWbp.cu:14-53: back-projection onto the 2x grid
 * (vol_x2 = 2 VolX, vol_z2 = 2 VolZ):
 *   u = (fx cos + fz sin)/2 + (ProjX/2) + 0.5,  fx = x - vol_x2/2, fz = z - vol_z2/2;
 * mean over valid tilts, 0 when none. */
void orc_back_proj_wbp(const double* sino, int pad_x, int proj_x,
                       const float* cos_sin, const unsigned char* no_proj,
                       int n_tilts, double* img2, int vol_x2, int vol_z2)
{
    int cx = proj_x / 2;
    int end_x = proj_x - 2;
    for (int z = 0; z < vol_z2; z++) {
        double fz = z - vol_z2 / 2.0;
        for (int x = 0; x < vol_x2; x++) {
            double fx = x - vol_x2 / 2.0;
            double sum = 0.0;
            int count = 0;
            for (int t = 0; t < n_tilts; t++) {
                if (no_proj != 0 && no_proj[t]) continue;
                double c = cos_sin[2 * t], s = cos_sin[2 * t + 1];
                double u = (fx * c + fz * s) / 2 + cx + 0.5;
                if (u < 0 || u > end_x) continue;
                int iu = (int)u;
                double f = u - iu;
                const double* row = sino + (size_t)t * pad_x;
                double v = row[iu] * (1 - f) + row[iu + 1] * f;
                if (v <= ORC_EMPTY) continue;
                sum += v;
                count += 1;
            }
            img2[(size_t)z * vol_x2 + x] = count <= 0 ? 0.0 : sum / count;
        }
    }
}

/* Fourier crop of a real image ax x ay to bx x by (bx <= ax, by <= ay, all
 * even): forward R2C normalised by 1/(ax*ay) (MaUtil/CCufft2D.cpp:65-80),
 * crop (MaUtil/GFourierResize2D.cu:7-39: output row j <= by/2 from input row
 * j, j > by/2 from input row j - by + ay; columns 0..bx/2), unnormalised
 * C2R (MaUtil/CCufft2D.cpp:108-123).  The crop leaves the Nyquist row and
 * column non-Hermitian; the C2R is defined here as a complex inverse along y
 * followed by a real inverse along x that ignores the imaginary parts of the
 * DC and Nyquist bins (numpy irfft2). */
void orc_fourier_crop(const double* img, int ax, int ay, int bx, int by, double* out)
{
    int kx_n = bx / 2 + 1;
    double* ar = (double*)malloc(sizeof(double) * (size_t)ay * kx_n);
    double* ai = (double*)malloc(sizeof(double) * (size_t)ay * kx_n);
    int big = ax > ay ? ax : ay;
    double* re = (double*)malloc(sizeof(double) * big);
    double* im = (double*)malloc(sizeof(double) * big);
    /* forward along x */
    for (int z = 0; z < ay; z++) {
        for (int x = 0; x < ax; x++) {
            re[x] = img[(size_t)z * ax + x];
            im[x] = 0.0;
        }
        orc_fft(re, im, ax, -1);
        for (int k = 0; k < kx_n; k++) {
            ar[(size_t)z * kx_n + k] = re[k];
            ai[(size_t)z * kx_n + k] = im[k];
        }
    }
    double norm = 1.0 / ((double)ax * (double)ay);
    double* hr = (double*)malloc(sizeof(double) * (size_t)by * kx_n);
    double* hi = (double*)malloc(sizeof(double) * (size_t)by * kx_n);
    double* gr = (double*)malloc(sizeof(double) * by);
    double* gi = (double*)malloc(sizeof(double) * by);
    /* forward along y, normalise, crop, inverse along y */
    for (int k = 0; k < kx_n; k++) {
        for (int z = 0; z < ay; z++) {
            re[z] = ar[(size_t)z * kx_n + k];
            im[z] = ai[(size_t)z * kx_n + k];
        }
        orc_fft(re, im, ay, -1);
        for (int j = 0; j < by; j++) {
            int src = j <= by / 2 ? j : j - by + ay;
            gr[j] = re[src] * norm;
            gi[j] = im[src] * norm;
        }
        orc_fft(gr, gi, by, +1);
        for (int z = 0; z < by; z++) {
            hr[(size_t)z * kx_n + k] = gr[z];
            hi[(size_t)z * kx_n + k] = gi[z];
        }
    }
    /* real inverse along x */
    for (int z = 0; z < by; z++) {
        const double* r = hr + (size_t)z * kx_n;
        const double* i = hi + (size_t)z * kx_n;
        re[0] = r[0];
        im[0] = 0.0;
        re[bx / 2] = r[bx / 2];
        im[bx / 2] = 0.0;
        for (int k = 1; k < bx / 2; k++) {
            re[k] = r[k];
            im[k] = i[k];
            re[bx - k] = r[k];
            im[bx - k] = -i[k];
        }
        orc_fft(re, im, bx, +1);
        for (int x = 0; x < bx; x++) out[(size_t)z * bx + x] = re[x];
    }
    free(ar);
    free(ai);
    free(hr);
    free(hi);
    free(gr);
    free(gi);
    free(re);
    free(im);
}

/* Recon/GBackProjWbp.cu:126-141: the 2x grid (2 VolX x 2 VolZ) cropped to
 * VolX x VolZ (orc_fourier_crop). */
void orc_wbp_crop(const double* img2, int vol_x, int vol_z, double* vol)
{
    orc_fourier_crop(img2, 2 * vol_x, 2 * vol_z, vol_x, vol_z, vol);
}

/* Correct/CBinStack.cpp:54-106 (-AtBin > 1): every image of the aligned stack
 * Fourier-cropped to the binned size of MaUtil/GFourierResize2D.cu:24-38,
 * ((int)(size / bin)) / 2 * 2 in each direction.  in [n][ay][ax] ->
 * out [n][by][bx]. */
void orc_bin_size(int ax, int ay, float bin, int* bx, int* by)
{
    *bx = ax;
    *by = ay;
    if (bin == 1.0f) return;
    *bx = (int)(ax / bin) / 2 * 2;
    *by = (int)(ay / bin) / 2 * 2;
}

void orc_bin_stack(const double* in, int n, int ax, int ay, int bx, int by, double* out)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int t = 0; t < n; t++)
        orc_fourier_crop(in + (size_t)t * ax * ay, ax, ay, bx, by, out + (size_t)t * bx * by);
}

void orc_tomo_wbp(double* sino, int n_tilts, int pad_x, int proj_x,
                  const float* cos_sin, const unsigned char* no_proj,
                  double* vol, int vol_x, int vol_z)
{
    orc_weight_projs(sino, n_tilts, pad_x, proj_x, cos_sin, vol_z);
    orc_rweight(sino, n_tilts, pad_x, proj_x);
    int nx2 = 2 * vol_x, nz2 = 2 * vol_z;
    double* img2 = (double*)malloc(sizeof(double) * (size_t)nx2 * nz2);
    orc_back_proj_wbp(sino, pad_x, proj_x, cos_sin, no_proj, n_tilts, img2, nx2, nz2);
    orc_wbp_crop(img2, vol_x, vol_z, vol);
    free(img2);
}

/* ---------------------------------------------------------------------- */
/* Pre-stage: Correct/GCorrPatchShift.cu:58-100 (global + local shift,      */
/* tilt-axis rotation, nearest neighbour) and :134-154 (random fill)       */
/* ---------------------------------------------------------------------- */
/* Correct/GCorrPatchShift.cu:18-48: local shift at rotated, centred (x, y)
 * from the patch table of one tilt (cx[n], cy[n], sx[n], sy[n], good[n]):
 * average of the good patches' shifts (good >= 0.9) weighted by
 * exp(-100 ((x - cx)/raw_x)^2 + ((y - cy)/raw_y)^2), 0 if no good patch. */
static void orc_local_shift(const float* lp, int np, double x, double y,
                            int raw_x, int raw_y, double* lsx, double* lsy)
{
    const float *cx = lp, *cy = lp + np, *sx = lp + 2 * np, *sy = lp + 3 * np;
    const float* good = lp + 4 * np;
    double wx = 0.0, wy = 0.0, sw = 0.0;
    int count = 0;
    for (int p = 0; p < np; p++) {
        if (good[p] < 0.9f) continue;
        double dx = (x - cx[p]) / raw_x, dy = (y - cy[p]) / raw_y;
        double w = exp(-100.0 * (dx * dx + dy * dy));
        wx += sx[p] * w;
        wy += sy[p] * w;
        sw += w;
        count += 1;
    }
    *lsx = (count > 0 && sw > 0) ? wx / sw : 0.0;
    *lsy = (count > 0 && sw > 0) ? wy / sw : 0.0;
}

void orc_corr_proj(const float* raw, int raw_x, int raw_y, float tilt_axis_deg,
                   float shift_x, float shift_y, int n_patches,
                   const float* local_params, double* out, int out_x,
                   int out_y, int out_pad_x, int random_fill,
                   const int* fill_table)
{
    const float d2r = 3.141592654f / 180.0f;
    float ang = tilt_axis_deg * d2r;
    double ca = cos((double)ang), sa = sin((double)ang);
    for (int y = 0; y < out_y; y++) {
        for (int x = 0; x < out_x; x++) {
            double xc = x + 0.5 - out_x * 0.5;
            double yc = y + 0.5 - out_y * 0.5;
            double xr = xc * ca - yc * sa;
            double yr = xc * sa + yc * ca;
            if (n_patches > 0 && local_params) {
                double lsx, lsy;
                orc_local_shift(local_params, n_patches, xr, yr, raw_x, raw_y, &lsx, &lsy);
                xr += lsx;
                yr += lsy;
            }
            xr += (shift_x + raw_x * 0.5);
            yr += (shift_y + raw_y * 0.5);
            int ix = (int)(xr + 0.5);
            int iy = (int)(yr + 0.5);
            double v = ORC_SENT;
            if (ix >= 0 && ix < raw_x && iy >= 0 && iy < raw_y)
                v = raw[(size_t)iy * raw_x + ix];
            out[(size_t)y * out_pad_x + x] = v;
        }
    }
    if (random_fill) orc_fill_empty(out, out_x, out_y, out_pad_x, fill_table);
}

/* Correct/CCorrTomoStack.cpp:185-222 per tilt, sizes from
 * CCorrectUtil::CalcAlignedSize with the axis of frame 0
 * (AreTomo/CAreTomoMain.cpp:612-615).  Output [t][aln_y][aln_x]. */
void orc_prestage(const float* raw, int n_tilts, int raw_x, int raw_y,
                  const float* tilt_axis_deg, const float* shifts_xy,
                  int n_patches, const float* local_params,
                  int random_fill, uint64_t fill_seed, double* aligned,
                  int aln_x, int aln_y)
{
    int pad = (aln_x / 2 + 1) * 2;
    int* table = (int*)malloc(sizeof(int) * (size_t)aln_x * aln_y);
    orc_fill_table(fill_seed, aln_x * aln_y, table);
#pragma omp parallel for schedule(dynamic, 1)
    for (int t = 0; t < n_tilts; t++) {
        double* buf = (double*)malloc(sizeof(double) * (size_t)pad * aln_y);
        float sx = shifts_xy ? shifts_xy[2 * t] : 0.0f;
        float sy = shifts_xy ? shifts_xy[2 * t + 1] : 0.0f;
        const float* lp = (n_patches > 0 && local_params)
                              ? local_params + (size_t)t * 5 * n_patches : 0;
        orc_corr_proj(raw + (size_t)t * raw_x * raw_y, raw_x, raw_y,
                      tilt_axis_deg[t], sx, sy, lp ? n_patches : 0, lp, buf,
                      aln_x, aln_y, pad, random_fill, table);
        for (int y = 0; y < aln_y; y++)
            memcpy(aligned + ((size_t)t * aln_y + y) * aln_x, buf + (size_t)y * pad,
                   sizeof(double) * aln_x);
        free(buf);
    }
    free(table);
}

/* Per-series slice loop: Recon/CDoWbpRecon.cpp:49-133 and
 * Recon/CDoSartRecon.cpp:193-279.  VolX = ProjX/2*2, pad = (ProjX/2+1)*2;
 * the sinogram of slice y is row y of every tilt; each slice starts from a
 * zeroed volume; the result is stored z-flipped (frame y, row VolZ-1-z).
 * Computes only the listed slices: out[k] = frame slices[k]. */
void orc_recon_slices(const double* aligned, int n_tilts, int proj_x,
                      int proj_y, const float* tilts_deg,
                      const unsigned char* no_proj, int method, int vol_z,
                      int start_tilt, int n_sart_tilts, int iters,
                      int n_subsets, double relax, int order, int positivity,
                      uint64_t fill_seed, const int* slices, int n_slices,
                      double* out)
{
    int vol_x = proj_x / 2 * 2;
    int pad = (proj_x / 2 + 1) * 2;
    int proj_eff = (pad / 2 - 1) * 2;  /* GBackProj.cu:74, GRWeight.cu:66 */
    float* cs = (float*)malloc(sizeof(float) * 2 * n_tilts);
    orc_cos_sin(tilts_deg, n_tilts, cs);
    int* table = (int*)malloc(sizeof(int) * (size_t)vol_x * vol_z);
    orc_fill_table(fill_seed ^ 0x5A5A5A5A5A5A5A5Aull, vol_x * vol_z, table);
    (void)proj_y;
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < n_slices; k++) {
        int y = slices[k];
        double* sino = (double*)calloc((size_t)pad * n_tilts, sizeof(double));
        double* vol = (double*)calloc((size_t)vol_x * vol_z, sizeof(double));
        for (int t = 0; t < n_tilts; t++)
            memcpy(sino + (size_t)t * pad,
                   aligned + ((size_t)t * proj_y + y) * proj_x,
                   sizeof(double) * proj_x);
        if (method == 0)
            orc_tomo_wbp(sino, n_tilts, pad, proj_eff, cs, no_proj, vol, vol_x, vol_z);
        else
            orc_tomo_sart(sino, n_tilts, pad, proj_eff, cs, no_proj, vol, vol_x,
                          vol_z, start_tilt, n_sart_tilts, iters, n_subsets,
                          relax, order, positivity, table);
        double* dst = out + (size_t)k * vol_x * vol_z;
        for (int z = 0; z < vol_z; z++)
            memcpy(dst + (size_t)(vol_z - 1 - z) * vol_x, vol + (size_t)z * vol_x,
                   sizeof(double) * vol_x);
        free(sino);
        free(vol);
    }
    free(cs);
    free(table);
}

/* DataUtil/CTiltSeries.cpp:286-305: XZY frames (frame y holds vol_z rows)
 * -> XYZ frames; new frame index = flip ? (vol_z-1-row) : row. */
/* ---------------------------------------------------------------------- */
/* ProjAlign re-projector: AreTomo/ProjAlign/CCalcReproj.cpp:45-153 and    */
/* GReproj.cu:22-203.  Angles use the reference's float constant 0.01745f  */
/* (GReproj.cu:6, CCalcReproj.cpp:10); geometry in double.                 */
/* ---------------------------------------------------------------------- */

/* CCalcReproj::mFindProjRange (CCalcReproj.cpp:73-124): the first and last
 * non-skipped tilt within 20.5 deg of the projection angle whose cos stretch
 * cos(t)/cos(a) is at most 1.2; at most 10 tilts; when none qualifies, the
 * neighbour one step toward 0 deg.  range[0..1] inclusive. */
void orc_reproj_range(const float* tilts, const unsigned char* skip, int n,
                      int proj_idx, int* range)
{
    const float d2r = 0.01745f;
    float a = tilts[proj_idx];
    double cos_a = cos((double)(a * d2r));
    int start = -1, end = -1;
    for (int i = 0; i < n && start < 0; i++) {
        if (skip && skip[i]) continue;
        if (fabs((double)(a - tilts[i])) > 20.5) continue;
        if ((float)(cos((double)(tilts[i] * d2r)) / cos_a) > 1.20f) continue;
        start = i;
    }
    for (int i = start < 0 ? n : start; i < n; i++) {
        if (skip && skip[i]) continue;
        if (fabs((double)(a - tilts[i])) > 20.5) continue;
        if ((float)(cos((double)(tilts[i] * d2r)) / cos_a) > 1.20f) continue;
        end = i;
    }
    if (start < 0 || end < 0) {
        int sign = a > 0 ? 1 : -1;
        start = end = proj_idx - sign;
    }
    if (end - start > 9) end = start + 9;
    range[0] = start;
    range[1] = end;
}

/* GReproj::DoIt per row y: mGBackProj (GReproj.cu:22-59) of tilts
 * [range0, range1] onto a 2 proj_x x vol_z slice -- nearest sample
 * (truncation) at u = x cos + z sin + proj_x/2, weighted by cos t *
 * cos(a - t), normalised by the weight sum (-1e30 below 0.001) -- then
 * mGForProj (:63-110) at angle a along rays of length
 * L = (int)(vol_z / cos a + 0.5f): mean of the non-empty samples, -1e30 when
 * fewer than 0.8 L.  projs [n][proj_y][proj_x], out [proj_y][proj_x]. */
void orc_reproj(const float* projs, const float* tilts, int n_tilts, int proj_x,
                int proj_y, int vol_z, int proj_idx, const int* range, double* out)
{
    const float d2r = 0.01745f;
    const int vol_x = 2 * proj_x;
    const float a = tilts[proj_idx];
    const int r0 = range[0], r1 = range[1];
    const int nr = r1 - r0 + 1;
    double* cs = (double*)malloc(sizeof(double) * 3 * (size_t)(nr > 0 ? nr : 1));
    for (int k = 0; k < nr; k++) {
        double v = (double)tilts[r0 + k] * (double)d2r;
        cs[3 * k] = cos(v);
        cs[3 * k + 1] = sin(v);
        cs[3 * k + 2] = cos(((double)a - (double)tilts[r0 + k]) * (double)d2r);
    }
    const double rad = 4.0 * atan(1.0) / 180.0;
    const float fc = (float)cos(rad * a), fs = (float)sin(rad * a);
    const int ray = (int)((float)vol_z / fc + 0.5f);
    const float thr = 0.8f * (float)ray;
    (void)n_tilts;
#pragma omp parallel for schedule(dynamic, 1)
    for (int y = 0; y < proj_y; y++) {
        float* vol = (float*)malloc(sizeof(float) * (size_t)vol_x * vol_z);
        for (int z = 0; z < vol_z; z++) {
            double fz = z + 0.5 - vol_z * 0.5;
            for (int x = 0; x < vol_x; x++) {
                double fx = x + 0.5 - vol_x * 0.5;
                double sum = 0.0, cnt = 0.0;
                for (int k = 0; k < nr; k++) {
                    double u = fx * cs[3 * k] + fz * cs[3 * k + 1] + proj_x * 0.5;
                    if (u < 0 || u > proj_x - 1) continue;
                    float v = projs[((size_t)(r0 + k) * proj_y + y) * proj_x + (int)u];
                    if (v < (float)ORC_EMPTY) continue;
                    sum += v * cs[3 * k] * cs[3 * k + 2];
                    cnt += cs[3 * k + 2];
                }
                vol[(size_t)z * vol_x + x] = cnt < 0.001 ? (float)ORC_SENT : (float)(sum / cnt);
            }
        }
        const double tz = vol_z * 0.5 - ray * 0.5 * (double)fc;
        for (int x = 0; x < proj_x; x++) {
            double xp = x + 0.5 - proj_x * 0.5;
            double tx = vol_x * 0.5 + ray * 0.5 * (double)fs + xp / (double)fc;
            double sum = 0.0;
            int cnt = 0;
            for (int i = 0; i < ray; i++) {
                double px = tx - i * (double)fs;
                double pz = tz + i * (double)fc;
                if (px < 0 || px >= vol_x || pz < 0 || pz >= vol_z) continue;
                float v = vol[(size_t)(int)pz * vol_x + (int)px];
                if (v <= (float)ORC_EMPTY) continue;
                sum += v;
                cnt += 1;
            }
            out[(size_t)y * proj_x + x] = ((float)cnt < thr) ? ORC_SENT : sum / cnt;
        }
        free(vol);
    }
    free(cs);
}

/* ---------------------------------------------------------------------- */
/* Sample thickness: Recon/CCalcVolThick.cpp:47-285.                      */
/* ---------------------------------------------------------------------- */

/* CCalcVolThick::mSetup + mMeasureStd (CCalcVolThick.cpp:116-170): the
 * standard deviation of the central tile ((int)(0.4 n)/2*2 per side) of
 * sections z = 0 .. nz-2 of an xyz volume, samples below -1e20 left out;
 * moments in double, sqrt taken in float as the reference does. */
void orc_thick_profile(const float* vol, int nx, int ny, int nz, float* out)
{
    int tx = (int)(nx * 0.4f) / 2 * 2, ty = (int)(ny * 0.4f) / 2 * 2;
    int x0 = (nx - tx) / 2, y0 = (ny - ty) / 2;
    for (int z = 0; z + 1 < nz; z++) {
        double s1 = 0.0, s2 = 0.0;
        int cnt = 0;
        for (int j = 0; j < ty; j++) {
            const float* row = vol + ((size_t)z * ny + y0 + j) * nx + x0;
            for (int i = 0; i < tx; i++) {
                float v = row[i];
                if (v < (float)-1e20) continue;
                s1 += v;
                s2 += (double)(v * v);
                cnt += 1;
            }
        }
        if (cnt == 0) {
            out[z] = 0.0f;
            continue;
        }
        s1 = s1 / cnt;
        s2 = s2 / cnt - s1 * s1;
        out[z] = s2 <= 0 ? 0.0f : sqrtf((float)s2);
    }
}

/* CCalcVolThick::mSmooth (:93-110): 11-point moving average with the ends
 * clamped, double sum then (float) sum / 11. */
void orc_thick_smooth(float* ccs, int n)
{
    const int win = 11;
    float* buf = (float*)malloc(sizeof(float) * (size_t)n);
    for (int i = 0; i < n; i++) {
        double sum = 0.0;
        for (int j = 0; j < win; j++) {
            int k = i - win / 2 + j;
            if (k < 0) k = 0;
            else if (k >= n) k = n - 1;
            sum += ccs[k];
        }
        buf[i] = (float)sum / win;
    }
    memcpy(ccs, buf, sizeof(float) * (size_t)n);
    free(buf);
}

/* CCalcVolThick::mDetectEdges (:185-255) on the smoothed profile: mean of
 * [n/10, n - n/10) in float, the first crossing of the mean from each end
 * averaged with the lowest point outside it, scaled by the binning (int *=
 * float).  out: low edge, high edge, thickness, z offset of the sample
 * centre from the volume centre.  A side with no crossing keeps its search
 * bound (the reference leaves the edge uninitialised). */
void orc_thick_edges(const float* ccs, int n, float binning, int* out)
{
    int half = n / 2;
    int bot = (int)(n * 0.1), top = n - bot;
    float mean = 0.0f;  /* the reference's std of the profile is unused */
    for (int i = bot; i < top; i++) mean += ccs[i];
    mean = mean / (top - bot);
    int e0 = bot, e1 = top;
    for (int i = bot; i < half; i++) {
        if (ccs[i] < mean) continue;
        e0 = i;
        break;
    }
    float mn = mean;
    int imin = e0;
    for (int i = bot; i < e0; i++) {
        if (ccs[i] >= mn) continue;
        mn = ccs[i];
        imin = i;
    }
    e0 = (e0 + imin) / 2;
    for (int i = top; i > half; i--) {
        if (ccs[i] < mean) continue;
        e1 = i;
        break;
    }
    mn = mean;
    imin = e1;
    for (int i = top; i > e1; i--) {
        if (ccs[i] >= mn) continue;
        mn = ccs[i];
        imin = i;
    }
    e1 = (e1 + imin) / 2;
    e0 = (int)(e0 * binning);
    e1 = (int)(e1 * binning);
    out[0] = e0;
    out[1] = e1;
    out[2] = e1 - e0;
    out[3] = (e0 + e1) / 2 - (int)(half * binning);
}

void orc_flip_vol(const double* vol_xzy, int vol_x, int vol_z, int vol_y,
                  int flip, double* vol_xyz)
{
    for (int y = 0; y < vol_z; y++) {
        int nf = flip ? (vol_z - 1 - y) : y;
        for (int z = 0; z < vol_y; z++)
            memcpy(vol_xyz + ((size_t)nf * vol_y + z) * vol_x,
                   vol_xzy + ((size_t)z * vol_z + y) * vol_x,
                   sizeof(double) * vol_x);
    }
}
