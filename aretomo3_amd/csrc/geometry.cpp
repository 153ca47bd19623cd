// geometry.cpp -- per-plan geometry evaluated once on the host in FP64.
//
// The reference decides, per voxel and tilt, whether a projection sample is
// used with FP32 tests such as `fXp < 0 || fXp > iProjEndX`
// (Recon/GBackProj.cu:35-37, Recon/GBackProjWbp.cu:34-36,
// Recon/GForProj.cu:145-149).  Those decisions are the only discontinuous
// part of the method, so this build evaluates them in double, once per plan,
// and hands the kernels integer intervals: per (tilt, z) the x-range a
// back-projection row uses, per (tilt, detector x) the sample range a
// forward-projection ray uses.  Inside the intervals the kernels interpolate
// in FP32.  The expressions are written exactly as the FP64 restatement of
// the reference evaluates them, and this file is compiled with
// -ffp-contract=off, so every inclusion decision is reproducible bit for bit.
#include <algorithm>
#include <cmath>
#include <functional>
#include <thread>

#include "atr_internal.h"

namespace atr {

uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Seeded stand-in for the cuRAND table of MaUtil/GGenRandoms.cu:7-21 (which
// the reference seeds with time(0)): tbl_i = ((r_i >> 32) * n) >> 32 with
// r_i = splitmix64(seed + i * 0xD1B54A32D192ED03).
void fill_table(uint64_t seed, int n, std::vector<int>& table)
{
    table.resize(n);
    for (int i = 0; i < n; i++) {
        uint64_t r = splitmix64(seed + (uint64_t)i * 0xD1B54A32D192ED03ull);
        table[i] = (int)(((r >> 32) * (uint64_t)n) >> 32);
    }
}

// Recon/CTomoBase.cpp:55-65.
void cos_sin_table(const float* tilts, int n, std::vector<float>& cs)
{
    const float rad = 0.017453f;
    cs.resize(2 * (size_t)n);
    for (int i = 0; i < n; i++) {
        float ang = rad * tilts[i];
        cs[2 * i] = (float)std::cos((double)ang);
        cs[2 * i + 1] = (float)std::sin((double)ang);
    }
}

// AreTomo/Util/CSplitItems.cpp:25-48.
void split_items(int n_items, int n_splits, std::vector<int>& starts, std::vector<int>& sizes)
{
    starts.assign(n_splits, 0);
    sizes.assign(n_splits, n_items / n_splits);
    for (int i = 0; i < n_items % n_splits; i++) sizes[i] += 1;
    for (int i = 1; i < n_splits; i++) starts[i] = starts[i - 1] + sizes[i - 1];
}

// Correct/CCorrectUtil.cpp:11-21.
void aligned_size(int raw_x, int raw_y, float axis_deg, int* ax, int* ay)
{
    double rot = std::fabs(std::sin(axis_deg * 3.14 / 180.0));
    *ax = raw_x;
    *ay = raw_y;
    if (rot <= 0.707) return;
    *ax = raw_y;
    *ay = raw_x;
}

// Interval {lo, hi} of integer i in [0, n) where pred(i) holds, for a
// predicate that holds on one contiguous run (monotone geometry).  Empty
// interval = {n, -1}.
template <class P>
static int2 run_of(int n, P pred)
{
    int lo = -1;
    for (int i = 0; i < n; i++)
        if (pred(i)) {
            lo = i;
            break;
        }
    if (lo < 0) return make_int2(n, -1);
    int hi = lo;
    while (hi + 1 < n && pred(hi + 1)) hi++;
    return make_int2(lo, hi);
}

// Interval search on a monotone run: binary search for the two edges around a
// known inside point would be faster, but the plain scan is already cheap
// (VolX * VolZ * N evaluations) and is obviously exact.  It is split over
// threads by tilt.
template <class F>
static void parallel_tilts(int n, F fn)
{
    unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned w = 0; w < hw; w++)
        th.emplace_back([=]() {
            for (int t = (int)w; t < n; t += (int)hw) fn(t);
        });
    for (auto& x : th) x.join();
}

void parallel_for(int n, const std::function<void(int)>& fn) { parallel_tilts(n, fn); }

// Recon/GBackProj.cu:25-37: u = fx cos + fz sin + ProjX/2, used iff
// 0 <= u <= ProjX-2, fx = x - VolX*0.5, fz = z - VolZ*0.5.
void build_bp_intervals(const std::vector<float>& cs, int n, int vol_x, int vol_z, int proj,
                        std::vector<int2>& lohi)
{
    lohi.assign((size_t)n * vol_z, make_int2(vol_x, -1));
    const double cx = proj / 2.0;
    const int end_x = proj - 2;
    parallel_tilts(n, [&](int t) {
        double c = cs[2 * t], s = cs[2 * t + 1];
        for (int z = 0; z < vol_z; z++) {
            double fz = z - vol_z * 0.5;
            lohi[(size_t)t * vol_z + z] = run_of(vol_x, [&](int x) {
                double fx = x - vol_x * 0.5;
                double u = fx * c + fz * s + cx;
                return !(u < 0 || u > end_x);
            });
        }
    });
}

// Per (tilt, z-tile, x-tile) of a back-projection grid: the start w0 of the
// window of BP_TX + tz + 16 row samples the tile stages, and the tile class
// (1 = every voxel of the tile uses the tilt, 0 = none does, 2 = some do).
// The grid coordinate is u = scale (fx cos + fz sin) + offset (SART: scale 1,
// offset P/2; the WBP 2x grid: scale 1/2, offset P/2 + 1/2).
// w0 = floor(min u over the tile corners) - 3 in FP64; the kernel's FP32
// coordinates relative to w0 then stay in [2, window - 2] (the tile spans at
// most scale ((BP_TX-1)|cos| + (tz-1)|sin|) in u).  The third field holds
// the FP32 bits of u - w0 at the tile's first voxel (x0, z0), formed in FP64:
// the kernel adds (x - x0) scale cos and k scale sin to it in FP32.
void build_bp_tiles(const std::vector<float>& cs, const std::vector<int2>& lohi, int n, int vol_x,
                    int vol_z, double scale, double offset, int tz, std::vector<int4>& tiles)
{
    int nzt = (vol_z + tz - 1) / tz, nxt = (vol_x + BP_TX - 1) / BP_TX;
    tiles.assign((size_t)n * nzt * nxt, int4{0, 0, 0, 0});
    for (int t = 0; t < n; t++) {
        const double c = cs[2 * t], s = cs[2 * t + 1];
        for (int zt = 0; zt < nzt; zt++)
            for (int xt = 0; xt < nxt; xt++) {
                int x0 = xt * BP_TX, x1 = std::min(x0 + BP_TX, vol_x) - 1;
                bool full = true, any = false;
                for (int z = zt * tz; z < std::min((zt + 1) * tz, vol_z); z++) {
                    int2 r = lohi[(size_t)t * vol_z + z];
                    if (!(r.x <= x0 && r.y >= x1)) full = false;
                    if (!(r.x > x1 || r.y < x0)) any = true;
                }
                const double fx0 = x0 - vol_x * 0.5, fx1 = x1 - vol_x * 0.5;
                const double fz0 = zt * tz - vol_z * 0.5, fz1 = fz0 + (tz - 1);
                const double umin =
                    offset + scale * (std::min(c * fx0, c * fx1) + std::min(s * fz0, s * fz1));
                const int w0 = (int)std::floor(umin) - 3;
                const float u00 = (float)(scale * (c * fx0 + s * fz0) + offset - w0);
                tiles[((size_t)t * nzt + zt) * nxt + xt] =
                    int4{w0, full ? 1 : (any ? 2 : 0), __builtin_bit_cast(int, u00), 0};
            }
    }
}

// Recon/GBackProjWbp.cu:23-36 on the 2x grid (vol_x2 = 2 VolX, vol_z2 =
// 2 VolZ): u = (fx cos + fz sin)/2 + (ProjX/2) + 0.5, used iff 0 <= u <= ProjX-2.
void build_wbp_intervals(const std::vector<float>& cs, int n, int vol_x, int vol_z, int proj,
                         std::vector<int2>& lohi)
{
    const int vx2 = 2 * vol_x, vz2 = 2 * vol_z;
    lohi.assign((size_t)n * vz2, make_int2(vx2, -1));
    const int cx = proj / 2;
    const int end_x = proj - 2;
    parallel_tilts(n, [&](int t) {
        double c = cs[2 * t], s = cs[2 * t + 1];
        for (int z = 0; z < vz2; z++) {
            double fz = z - vz2 / 2.0;
            lohi[(size_t)t * vz2 + z] = run_of(vx2, [&](int x) {
                double fx = x - vx2 / 2.0;
                double u = (fx * c + fz * s) / 2 + cx + 0.5;
                return !(u < 0 || u > end_x);
            });
        }
    });
}

// Recon/GForProj.cu:131-149: ray (t, xd) has L = (int)(VolZ/cos + 1.5)
// samples; sample k at fzk = k - L*0.5, x = ((xd - 0.5 ProjX)/cos + VolX*0.5)
// - fzk sin, z = VolZ*0.5 + fzk cos, used iff inside [0,VolX-2]x[0,VolZ-2].
// Tables: per ray its sample interval (as first sample and count); per
// (tilt, k) the cell row, the z weight and fzk*sin; per (tilt, band) the
// samples whose cell row lies in the band; per (tilt, 64-ray tile, band)
// whether every ray of the tile uses the whole band window.
bool build_fp_tables(const std::vector<float>& cs, int n, int vol_x, int vol_z, int proj,
                     int bz, Geometry& g)
{
    const int end_x = vol_x - 2, end_z = vol_z - 2;
    const int ntile = (proj + FP_WAVE - 1) / FP_WAVE;
    g.fp_len.assign(n, 0);
    g.fp_geo_off.assign(n, 0);
    size_t total = 0;
    for (int t = 0; t < n; t++) {
        double c = cs[2 * t];
        if (!(c > 0.0)) return false;
        double l = vol_z / c + 1.5;
        if (!(l < 1e6)) return false;
        int len = (int)l;
        g.fp_len[t] = len;
        g.fp_geo_off[t] = (int)total;
        total += (size_t)len;
    }
    // four spare entries: the kernels read geometry in groups of four
    g.fp_geo.assign(total + 4, make_float4(0, 0, 0, 0));
    g.fp_range.assign((size_t)n * proj, make_int2(0, 0));
    const int nb = (vol_z - 1 + bz - 1) / bz;
    g.fp_bz = bz;
    g.fp_nbands = nb;
    g.fp_band.assign((size_t)n * nb, make_int2(0, -1));
    parallel_tilts(n, [&](int t) {
        double c = cs[2 * t], s = cs[2 * t + 1];
        int len = g.fp_len[t];
        float4* geo = g.fp_geo.data() + g.fp_geo_off[t];
        double tz = vol_z * 0.5;
        int kz_lo = len, kz_hi = -1;
        for (int k = 0; k < len; k++) {
            double fzk = k - len * 0.5;
            double z = tz + fzk * c;
            bool zin = !(z < 0 || z > end_z);
            if (zin) {
                kz_lo = std::min(kz_lo, k);
                kz_hi = std::max(kz_hi, k);
            }
            double zc = std::min(std::max(z, 0.0), (double)end_z);
            int iz = (int)zc;
            float wz = (float)(zc - iz);
            float q = (float)(fzk * s);
            // byte offset of the sample's lower row inside its band buffer
            // (row iz % bz, (VolX + 1) interpolation pairs per row)
            const int row_bytes = (iz % bz) * (vol_x + 1) * 8;
            geo[k] = make_float4(__builtin_bit_cast(float, row_bytes), wz, q,
                                 (float)(1.0 - (zc - iz)));
            if (zin) {
                int2& bb = g.fp_band[(size_t)t * nb + iz / bz];
                if (bb.y < bb.x) bb = make_int2(k, k);
                else bb.y = k;
            }
        }
        for (int xd = 0; xd < proj; xd++) {
            double xp = xd - 0.5 * proj;
            double tx = xp / c + vol_x * 0.5;
            int lo = -1, hi = -2;
            for (int k = std::max(kz_lo, 0); k <= kz_hi; k++) {
                double fzk = k - len * 0.5;
                double x = tx - fzk * s;
                double z = tz + fzk * c;
                bool in = !(x < 0 || z < 0 || x > end_x || z > end_z);
                if (in) {
                    if (lo < 0) lo = k;
                    hi = k;
                } else if (lo >= 0) {
                    break;
                }
            }
            g.fp_range[(size_t)t * proj + xd] = lo < 0 ? make_int2(0, 0) : make_int2(lo, hi - lo + 1);
        }
    });
    // One descriptor per (tilt, 64-ray tile, band): the band's sample window
    // clipped to the union of the tile's ray intervals, as {first geometry
    // entry, sample count, first sample index, 1 if every ray of the tile
    // uses every sample of the window}.
    g.fp_desc.assign((size_t)n * ntile * nb, make_int4(0, 0, 0, 0));
    parallel_tilts(n, [&](int t) {
        for (int tile = 0; tile < ntile; tile++) {
            const int a = tile * FP_WAVE, e = std::min(proj, a + FP_WAVE) - 1;
            int ulo = 1 << 30, uhi = -1;
            for (int xd = a; xd <= e; xd++) {
                int2 r = g.fp_range[(size_t)t * proj + xd];
                if (r.y > 0) {
                    ulo = std::min(ulo, r.x);
                    uhi = std::max(uhi, r.x + r.y - 1);
                }
            }
            for (int b = 0; b < nb; b++) {
                const int2 kb = g.fp_band[(size_t)t * nb + b];
                const int lo = std::max(kb.x, ulo), hi = std::min(kb.y, uhi);
                int4 d = make_int4(0, 0, 0, 0);
                if (kb.x <= kb.y && lo <= hi) {
                    bool full = true;
                    for (int xd = a; xd <= e && full; xd++) {
                        int2 r = g.fp_range[(size_t)t * proj + xd];
                        full = r.y > 0 && r.x <= lo && r.x + r.y - 1 >= hi;
                    }
                    d = make_int4(g.fp_geo_off[t] + lo, hi - lo + 1, lo, full ? 1 : 0);
                }
                g.fp_desc[((size_t)t * ntile + tile) * nb + b] = d;
            }
        }
    });
    return true;
}

// Factorisation into FFT stages: radix 4, 2, 3, 5 butterflies first, then
// one direct-DFT stage per remaining prime factor.
bool make_fft(int n, FftDesc& fd)
{
    fd = FftDesc{};
    fd.n = n;
    int m = n, s = 0;
    if (n < 2) return false;
    while (m % 4 == 0 && s < ATR_FFT_MAX_STAGES) { fd.radix[s++] = 4; m /= 4; }
    while (m % 2 == 0 && s < ATR_FFT_MAX_STAGES) { fd.radix[s++] = 2; m /= 2; }
    while (m % 3 == 0 && s < ATR_FFT_MAX_STAGES) { fd.radix[s++] = 3; m /= 3; }
    while (m % 5 == 0 && s < ATR_FFT_MAX_STAGES) { fd.radix[s++] = 5; m /= 5; }
    // remaining prime factors: direct-DFT stages (fft_stage_any)
    for (int q = 7; q * q <= m && s < ATR_FFT_MAX_STAGES; q += 2)
        while (m % q == 0 && s < ATR_FFT_MAX_STAGES) { fd.radix[s++] = q; m /= q; }
    if (m > 1 && s < ATR_FFT_MAX_STAGES) { fd.radix[s++] = m; m = 1; }
    fd.n_stages = s;
    return m == 1;
}

std::vector<float2> twiddles(int n)
{
    std::vector<float2> w(n);
    const double two_pi = 6.283185307179586476925286766559;
    for (int e = 0; e < n; e++) {
        double a = two_pi * (double)e / (double)n;
        w[e] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    return w;
}

// Recon/GRWeight.cu:7-23: bin x of the length-n row spectrum is scaled by
// 2 r (0.55 + 0.45 cos(6.2831852 r)) / n, r = x/n (evaluated in double).
std::vector<float> ramp_filter(int n)
{
    std::vector<float> w(n / 2 + 1);
    const double c2pi = (double)6.2831852f;
    for (int x = 0; x <= n / 2; x++) {
        double r = (double)x / (double)n;
        w[x] = (float)(2.0 * r * (0.55 + 0.45 * std::cos(c2pi * r)) / (double)n);
    }
    return w;
}

}  // namespace atr

namespace atr {
// Cell rows per forward-projection band: bz + 1 rows of (VolX + 1)
// interpolation pairs must fit the LDS budget of one workgroup.
int fp_band_rows(int vol_x, int pf)
{
    if (pf != FP_PF_DMA) {
        int rows = FP_LDS_BUDGET / ((vol_x + 1) * (int)sizeof(float2));
        return std::min(FP_BZ_MAX, std::max(1, rows - 1));
    }
    // pair rows and raw rows together (fp_lds_bytes) within FP_LDS_PF
    int bz = FP_BZ_MAX;
    while (bz > 1 && fp_lds_bytes(bz, vol_x, 1) > (size_t)FP_LDS_PF) bz--;
    return bz;
}

// The VolZ - 1 cell rows split into equal bands of at most bz_max rows: a
// short last band costs a full band's reload and per-pair overhead.
int fp_balanced_rows(int bz_max, int vol_z)
{
    const int cells = std::max(1, vol_z - 1);
    bz_max = std::max(1, bz_max);
    const int nb = (cells + bz_max - 1) / bz_max;
    return (cells + nb - 1) / nb;
}

// Band heights the autotuner times: the largest that fits the LDS, then
// about 2/3 and 1/2 of it, each balanced over the slice's cell rows, none
// above the LDS limit, no height twice.  At VolX >= 6400 only one row fits
// (rows of VolX + 1 pairs), and the list is {1}.
void fp_band_candidates(int vol_x, int vol_z, std::vector<int>& rows, int pf)
{
    const int bzmax = fp_band_rows(vol_x, pf);
    rows.clear();
    for (int want : {bzmax, bzmax * 2 / 3, bzmax / 2}) {
        const int bz = std::min(bzmax, fp_balanced_rows(std::max(1, want), vol_z));
        if (std::find(rows.begin(), rows.end(), bz) == rows.end()) rows.push_back(bz);
    }
}

// Tilts per forward-projection workgroup for a list of n_list tilts: at most
// the FP_WAVES * FP_MAXP pair slots of 64-ray tiles, and the list split into
// groups of equal size (13 tilts with room for 8 -> 7 + 6, not 8 + 5).
int fp_group(int proj, int n_list, int n_waves)
{
    const int ntile = (proj + FP_WAVE - 1) / FP_WAVE;
    const int cap = std::max(1, n_waves * FP_MAXP / ntile);
    const int groups = (std::max(n_list, 1) + cap - 1) / cap;
    return (std::max(n_list, 1) + groups - 1) / groups;
}

// Assignment of the (tilt, 64-ray tile) pairs of one forward-projection
// launch to the FP_WAVES waves of a workgroup (FP_MAXP slots each).  Waves
// meet at a barrier after every z-band, so what a band costs is the largest
// per-wave load in it; the schedule minimises the sum over bands of that
// maximum: greedy placement by decreasing cost, then pairwise exchanges.
// Cost of a pair in a band = the samples its descriptor covers, weighted by
// 1 + |sin theta| (the 64 rays of a tile read LDS 1 / cos theta columns
// apart, so bank conflicts grow with the tilt) and by 1.4 for masked
// windows, plus a fixed 8 per non-empty descriptor.
// tab gets one block of FP_WAVES * FP_MAXP entries per group of `group`
// tilts: entry [w * FP_MAXP + i] = (local tilt << 24) | (tilt * ntile +
// tile), the second part being the pair's row of band descriptors, or -1.
// Only waves 0 .. n_waves - 1 get pairs (the band-prefetching projector
// keeps its last wave for the loads).
void fp_pair_schedule(const Geometry& g, const int* list, int n_list, int proj, int group,
                      std::vector<int>& tab, int n_waves)
{
    const int ntile = (proj + FP_WAVE - 1) / FP_WAVE;
    const int nb = g.fp_nbands;
    for (int j0 = 0; j0 < n_list; j0 += group) {
        const int ng = std::min(group, n_list - j0);
        const int n_items = ng * ntile;
        std::vector<double> cost((size_t)n_items * nb);
        std::vector<double> total(n_items, 0.0);
        for (int it = 0; it < n_items; it++) {
            const int t = list[j0 + it / ntile], tile = it % ntile;
            for (int b = 0; b < nb; b++) {
                const int4 d = g.fp_desc[((size_t)t * ntile + tile) * nb + b];
                const double wt = 1.0 + std::fabs((double)g.cos_sin[2 * t + 1]);
                double c = d.y > 0 ? d.y * wt * (d.w ? 1.0 : 1.4) + 8.0 : 0.0;
                cost[(size_t)it * nb + b] = c;
                total[it] += c;
            }
        }
        std::vector<int> order(n_items);
        for (int i = 0; i < n_items; i++) order[i] = i;
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return total[a] > total[b]; });
        std::vector<double> load((size_t)n_waves * nb, 0.0);
        std::vector<int> owner(n_items, -1), fill(n_waves, 0);
        auto band_max = [&](int b) {
            double m = 0.0;
            for (int w = 0; w < n_waves; w++) m = std::max(m, load[(size_t)w * nb + b]);
            return m;
        };
        std::vector<double> bmax(nb, 0.0);
        for (int it : order) {
            int best = -1;
            double best_obj = 0.0, best_tot = 0.0;
            for (int w = 0; w < n_waves; w++) {
                if (fill[w] >= FP_MAXP) continue;
                double obj = 0.0, tot = 0.0;
                for (int b = 0; b < nb; b++) {
                    const double l = load[(size_t)w * nb + b] + cost[(size_t)it * nb + b];
                    obj += std::max(bmax[b], l);
                    tot += l;
                }
                if (best < 0 || obj < best_obj - 1e-9 || (obj < best_obj + 1e-9 && tot < best_tot)) {
                    best = w;
                    best_obj = obj;
                    best_tot = tot;
                }
            }
            owner[it] = best;
            fill[best]++;
            for (int b = 0; b < nb; b++) {
                load[(size_t)best * nb + b] += cost[(size_t)it * nb + b];
                bmax[b] = std::max(bmax[b], load[(size_t)best * nb + b]);
            }
        }
        // pairwise exchanges between waves while they lower the objective.
        // An exchange between waves u and v changes only their loads, so a
        // candidate is scored from the three largest loads of each band
        // (the largest one not on u or v stands for every other wave): O(nb)
        // per candidate instead of O(nb * n_waves).
        double cur = 0.0;
        for (int b = 0; b < nb; b++) cur += band_max(b);
        std::vector<double> top_v((size_t)nb * 3);
        std::vector<int> top_w((size_t)nb * 3);
        auto rank_band = [&](int b) {
            double* tv = &top_v[(size_t)b * 3];
            int* tw = &top_w[(size_t)b * 3];
            for (int k = 0; k < 3; k++) {
                tv[k] = 0.0;
                tw[k] = -1;
            }
            for (int w = 0; w < n_waves; w++) {
                const double l = load[(size_t)w * nb + b];
                for (int k = 0; k < 3; k++) {
                    if (tw[k] < 0 || l > tv[k]) {
                        for (int m = 2; m > k; m--) {
                            tv[m] = tv[m - 1];
                            tw[m] = tw[m - 1];
                        }
                        tv[k] = l;
                        tw[k] = w;
                        break;
                    }
                }
            }
        };
        for (int b = 0; b < nb; b++) rank_band(b);
        for (int pass = 0; pass < 4; pass++) {
            bool improved = false;
            for (int a = 0; a < n_items; a++)
                for (int c = a + 1; c < n_items; c++) {
                    const int u = owner[a], v = owner[c];
                    if (u == v) continue;
                    double o = 0.0;
                    for (int b = 0; b < nb; b++) {
                        const double d = cost[(size_t)c * nb + b] - cost[(size_t)a * nb + b];
                        const double lu = load[(size_t)u * nb + b] + d;
                        const double lv = load[(size_t)v * nb + b] - d;
                        double m = std::max(lu, lv);
                        for (int k = 0; k < 3; k++) {
                            const int w = top_w[(size_t)b * 3 + k];
                            if (w >= 0 && w != u && w != v) {
                                m = std::max(m, top_v[(size_t)b * 3 + k]);
                                break;
                            }
                        }
                        o += m;
                    }
                    if (o < cur - 1e-9) {
                        cur = o;
                        std::swap(owner[a], owner[c]);
                        improved = true;
                        for (int b = 0; b < nb; b++) {
                            const double d = cost[(size_t)c * nb + b] - cost[(size_t)a * nb + b];
                            load[(size_t)u * nb + b] += d;
                            load[(size_t)v * nb + b] -= d;
                            rank_band(b);
                        }
                    }
                }
            if (!improved) break;
        }
        const size_t base = tab.size();
        tab.resize(base + FP_WAVES * FP_MAXP, -1);
        std::vector<int> slot(FP_WAVES, 0);
        for (int it = 0; it < n_items; it++) {
            const int w = owner[it];
            const int t = list[j0 + it / ntile];
            tab[base + w * FP_MAXP + slot[w]++] = ((it / ntile) << 24) | (t * ntile + it % ntile);
        }
    }
}
}  // namespace atr
