// api.cpp -- plan construction, stream-ordered execution and the C ABI of
// libatr_recon.so.
//
// Host control flow mirrors the reference operators:
//   SART  Recon/CTomoSart.cpp:54-90   weight -> BP(all, relax 1) ->
//         iters x subsets {FP, diff, BP(relax)} -> fill empty
//   WBP   Recon/CTomoWbp.cpp:38-60    weight -> ramp filter -> 2x-grid BP ->
//         Fourier crop
// but every kernel processes every slice of the series at once (the
// reference loops over y-slices on the host, Recon/CDoWbpRecon.cpp:69-80),
// and all state stays on the device between kernels.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>

#include "atr_internal.h"

namespace atr {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

atr_status fail(atr_status st, const std::string& msg)
{
    g_err = msg;
    return st;
}

atr_status hip_fail(hipError_t e, const char* what)
{
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? ATR_ERR_OUT_OF_MEMORY : ATR_ERR_HIP;
}

// Device memory of a plan: its allocator when it has one (the server's
// arena), else hipMalloc / hipFree.
static hipError_t plan_malloc(const Plan& p, void** dst, size_t b)
{
    return p.alloc ? p.alloc->alloc(dst, b) : hipMalloc(dst, b);
}

static void plan_mfree(const Plan& p, void* q)
{
    if (q == nullptr) return;
    if (p.alloc) p.alloc->release(q);
    else (void)hipFree(q);
}

template <class T>
static atr_status upload(const Plan& p, T** dst, const std::vector<T>& src, size_t& bytes)
{
    *dst = nullptr;
    if (src.empty()) return ATR_OK;
    size_t b = sizeof(T) * src.size();
    ATR_HIP(plan_malloc(p, (void**)dst, b));
    if (p.upload_stream_set)
        ATR_HIP(hipMemcpyAsync(*dst, src.data(), b, hipMemcpyHostToDevice, p.upload_stream));
    else
        ATR_HIP(hipMemcpy(*dst, src.data(), b, hipMemcpyHostToDevice));
    bytes += b;
    return ATR_OK;
}

static atr_status dalloc(const Plan& p, void** dst, size_t b, size_t& bytes)
{
    *dst = nullptr;
    if (b == 0) return ATR_OK;
    ATR_HIP(plan_malloc(p, dst, b));
    bytes += b;
    return ATR_OK;
}

static size_t storage_bytes(int storage) { return storage == ATR_STORAGE_F32 ? 4 : 2; }

// Sizes, validation and the SART subset lists; no device work.
static atr_status plan_sizes(const atr_series_desc* d, Plan& p)
{
    if (d == nullptr) return fail(ATR_ERR_INVALID_ARG, "descriptor is NULL");
    p.desc = *d;
    if (d->proj_x < 4 || d->proj_y < 1 || d->n_tilts < 1)
        return fail(ATR_ERR_INVALID_ARG, "proj_x >= 4, proj_y >= 1 and n_tilts >= 1 required");
    if (d->tilt_deg == nullptr) return fail(ATR_ERR_INVALID_ARG, "tilt_deg is NULL");
    if (d->vol_z < 2) return fail(ATR_ERR_INVALID_ARG, "vol_z >= 2 required");
    if (d->method != ATR_METHOD_WBP && d->method != ATR_METHOD_SART)
        return fail(ATR_ERR_INVALID_ARG, "unknown method");
    if (d->storage < ATR_STORAGE_F32 || d->storage > ATR_STORAGE_BF16)
        return fail(ATR_ERR_INVALID_ARG, "unknown storage");
    if (d->out_layout < ATR_LAYOUT_XZY || d->out_layout > ATR_LAYOUT_SLICE)
        return fail(ATR_ERR_INVALID_ARG, "unknown output layout");
    p.n = d->n_tilts;
    p.in_x = d->proj_x;
    p.in_y = d->proj_y;
    p.in_row_stride = d->row_stride ? d->row_stride : (size_t)d->proj_x;
    p.in_tilt_stride = d->tilt_stride ? d->tilt_stride : p.in_row_stride * d->proj_y;
    if (p.in_row_stride < (size_t)d->proj_x ||
        p.in_tilt_stride < p.in_row_stride * (d->proj_y - 1) + d->proj_x)
        return fail(ATR_ERR_INVALID_ARG, "input strides smaller than the image");
    for (int t = 0; t < p.n; t++)
        if (!std::isfinite(d->tilt_deg[t])) return fail(ATR_ERR_INVALID_ARG, "non-finite tilt");
    p.prestage = d->tilt_axis_deg != nullptr;
    if (p.prestage) {
        aligned_size(p.in_x, p.in_y, d->tilt_axis_deg[0], &p.aln_x, &p.aln_y);
        for (int t = 0; t < p.n; t++)
            if (!std::isfinite(d->tilt_axis_deg[t]) ||
                (d->shift_xy && (!std::isfinite(d->shift_xy[2 * t]) ||
                                 !std::isfinite(d->shift_xy[2 * t + 1]))))
                return fail(ATR_ERR_INVALID_ARG, "non-finite tilt axis or shift");
    } else {
        p.aln_x = p.in_x;
        p.aln_y = p.in_y;
    }
    p.n_patches = 0;
    if (d->n_patches < 0 || d->n_patches > 2048)
        return fail(ATR_ERR_INVALID_ARG, "n_patches must be in [0, 2048]");
    if (d->n_patches > 0) {
        if (!p.prestage || d->local_params == nullptr)
            return fail(ATR_ERR_INVALID_ARG, "local patch shifts need tilt_axis_deg and local_params");
        for (size_t i = 0; i < (size_t)p.n * 5 * d->n_patches; i++)
            if (!std::isfinite(d->local_params[i]))
                return fail(ATR_ERR_INVALID_ARG, "non-finite local alignment parameter");
        p.n_patches = d->n_patches;
    }
    // -AtBin: CBinStack Fourier binning to ((int)(size / bin)) / 2 * 2
    // (MaUtil/GFourierResize2D.cu:24-38), fill forced on (CCorrTomoStack.cpp:107)
    const float bin = d->bin;
    if (!(bin >= 1.0f) || !std::isfinite(bin)) return fail(ATR_ERR_INVALID_ARG, "bin must be >= 1");
    p.binned = bin != 1.0f;
    p.bin_x = p.aln_x;
    p.bin_y = p.aln_y;
    if (p.binned) {
        p.bin_x = (int)(p.aln_x / bin) / 2 * 2;
        p.bin_y = (int)(p.aln_y / bin) / 2 * 2;
        if (p.bin_x < 4 || p.bin_y < 2)
            return fail(ATR_ERR_INVALID_ARG, "binned image smaller than 4 x 2");
        FftDesc f;
        if (!make_fft(p.aln_x, f) || !make_fft(p.aln_y, f) || !make_fft(p.bin_x, f) ||
            !make_fft(p.bin_y, f))
            return fail(ATR_ERR_UNSUPPORTED_SIZE,
                        "Fourier binning: image size has no FFT plan");
        if (p.aln_x > 8192 || p.aln_y > 8192)
            return fail(ATR_ERR_UNSUPPORTED_SIZE, "Fourier binning supports images up to 8192");
    }
    p.fill = p.prestage && (d->random_fill != 0 || p.binned);
    // CDoWbpRecon.cpp:37-38, GBackProj.cu:74: VolX = ProjX/2*2 = effective ProjX
    p.proj = p.bin_x / 2 * 2;
    p.vol_x = p.proj;
    p.vol_y = p.bin_y;
    p.vol_z = d->vol_z;
    p.sino_stride = (p.proj + 63) / 64 * 64;
    p.method = d->method;
    p.storage = d->storage;
    p.layout = d->out_layout;

    if (p.method == ATR_METHOD_SART) {
        p.iters = d->sart_iters;
        p.n_subsets = d->sart_subsets;
        p.sart_start = d->sart_start_tilt;
        p.sart_n = d->sart_n_tilts > 0 ? d->sart_n_tilts : p.n - p.sart_start;
        p.order = d->subset_order;
        p.positivity = d->positivity ? 1 : 0;
        if (p.iters < 0) return fail(ATR_ERR_INVALID_ARG, "sart_iters < 0");
        if (p.order != ATR_SUBSETS_CONTIGUOUS && p.order != ATR_SUBSETS_INTERLEAVED)
            return fail(ATR_ERR_INVALID_ARG, "unknown subset order");
        if (p.sart_start < 0 || p.sart_n < 1 || p.sart_start + p.sart_n > p.n)
            return fail(ATR_ERR_INVALID_ARG, "SART tilt range outside the series");
        if (p.n_subsets < 1 || p.n_subsets > p.sart_n)
            return fail(ATR_ERR_INVALID_ARG, "sart_subsets must be in [1, number of SART tilts]");
        p.relax = d->sart_relax > 0 ? d->sart_relax : 1.0f / (float)p.n_subsets;
        if (2 * (p.vol_x + 1) * (int)sizeof(float2) > FP_LDS_BUDGET)
            return fail(ATR_ERR_UNSUPPORTED_SIZE,
                        "SART supports VolX <= " +
                            std::to_string(FP_LDS_BUDGET / (2 * (int)sizeof(float2)) - 1));
        // subset lists (Recon/CTomoSart.cpp:60-61,74-80; AreTomo/Util/CSplitItems.cpp)
        std::vector<int> starts, sizes;
        split_items(p.sart_n, p.n_subsets, starts, sizes);
        p.subset_off.assign(1, 0);
        p.subset_fp_off.assign(1, 0);
        p.subset_tilts.clear();
        p.subset_fp.clear();
        p.max_subset = 0;
        for (int s = 0; s < p.n_subsets; s++) {
            std::vector<int> tl;
            if (p.order == ATR_SUBSETS_CONTIGUOUS) {
                for (int k = 0; k < sizes[s]; k++) tl.push_back(p.sart_start + starts[s] + k);
            } else {
                for (int k = s; k < p.sart_n; k += p.n_subsets) tl.push_back(p.sart_start + k);
            }
            for (int j = 0; j < (int)tl.size(); j++) {
                p.subset_fp.push_back(tl[j]);
                if (d->exclude && d->exclude[tl[j]]) continue;
                p.subset_tilts.push_back(TiltRow{tl[j], j});
            }
            p.subset_off.push_back((int)p.subset_tilts.size());
            p.subset_fp_off.push_back((int)p.subset_fp.size());
            p.max_subset = std::max(p.max_subset, (int)tl.size());
        }
    }
    p.all_tilts.clear();
    for (int t = 0; t < p.n; t++)
        if (!(d->exclude && d->exclude[t])) p.all_tilts.push_back(TiltRow{t, t});

    if (p.method == ATR_METHOD_WBP) {
        FftDesc f;
        if (!make_fft(p.proj, f) || !make_fft(2 * p.vol_x, f) || !make_fft(2 * p.vol_z, f) ||
            !make_fft(p.vol_z, f) || !make_fft(p.vol_x, f))
            return fail(ATR_ERR_UNSUPPORTED_SIZE,
                        "WBP: ProjX, VolX or VolZ has no FFT plan");
        if (p.vol_x > 2048 || p.vol_z > 2048)
            return fail(ATR_ERR_UNSUPPORTED_SIZE,
                        "WBP supports ProjX (= VolX) <= 2048 and VolZ <= 2048");
        // slab of slices whose spectra are buffered between the WBP kernels
        size_t per_slice = (size_t)3 * p.vol_z * (p.vol_x / 2 + 1) * sizeof(float2) +
                           (size_t)4 * p.vol_z * p.vol_x * sizeof(float);
        size_t budget = (size_t)1 << 30;
        p.wbp_slab = (int)std::max<size_t>(1, std::min<size_t>(p.vol_y, budget / per_slice));
    }
    return ATR_OK;
}

static size_t plan_bytes_estimate(const Plan& p)
{
    size_t b = 0;
    size_t n = p.n, sb = storage_bytes(p.storage);
    b += n * 64 + (size_t)p.vol_y * 64;
    if (p.prestage) {
        b += n * p.aln_x * (size_t)p.aln_y * 4 + n * p.aln_y * 16 + (size_t)p.aln_x * p.aln_y * 4;
        b += n * 5 * (size_t)p.n_patches * 4;
    }
    if (p.binned) {
        size_t kx = p.bin_x / 2 + 1;
        b += n * (size_t)p.bin_x * p.bin_y * 4 + n * kx * (p.aln_y + p.bin_y) * 8;
        b += (size_t)(p.aln_x + p.aln_y + p.bin_x + p.bin_y) * 8;
    }
    b += (size_t)p.vol_y * n * p.sino_stride * sb;
    if (p.method == ATR_METHOD_SART) {
        size_t vox = (size_t)p.vol_x * p.vol_y * p.vol_z;
        b += vox * 4 + (size_t)p.vol_y * p.max_subset * p.sino_stride * sb;
        b += n * p.vol_z * 8 + n * p.proj * 8 + (size_t)p.vol_x * p.vol_z * 4;
        b += n * (size_t)(p.vol_z * 2.5 + 16) * 16;  // FP sample tables (bounded by 60 deg)
    } else {
        b += (size_t)p.wbp_slab * 3 * p.vol_z * (p.vol_x / 2 + 1) * sizeof(float2);
        b += (size_t)p.wbp_slab * 4 * p.vol_z * p.vol_x * sizeof(float);
        b += n * 2 * p.vol_z * 8 + n * 16;
        b += n * (size_t)((2 * p.vol_z + BP_TZ - 1) / BP_TZ) * ((2 * p.vol_x + BP_TX - 1) / BP_TX) * 16;
    }
    return b;
}

template <class T>
static atr_status reupload(const Plan& p, T** dst, const std::vector<T>& src, size_t& bytes,
                           size_t& old_bytes)
{
    if (*dst) {
        plan_mfree(p, *dst);
        *dst = nullptr;
        bytes -= old_bytes;
    }
    old_bytes = sizeof(T) * src.size();
    return upload(p, dst, src, bytes);
}

// Host half of the SART tables that depend on the tile shapes: back-
// projection tile windows/classes for z-tiles of tz rows, forward-projection
// band tables and wave schedules for bands of bz cell rows, over all waves
// or, with the band prefetch (pf), all but the loading wave.  No device work.
static atr_status sart_tables_host(Plan& p, int tz, int chunk, int bz, int pf)
{
    Geometry& g = p.geo;
    p.bp_tz = tz;
    p.bp_chunk = chunk;
    p.fp_pf = fp_mode_supported(pf, p.vol_x) ? pf : 0;
    build_bp_tiles(g.cos_sin, g.bp_lohi, p.n, p.vol_x, p.vol_z, 1.0, p.proj / 2.0, tz, g.bp_tiles);
    if (!build_fp_tables(g.cos_sin, p.n, p.vol_x, p.vol_z, p.proj, bz, g))
        return fail(ATR_ERR_INVALID_ARG, "SART needs every tilt in (-90, 90) degrees");
    std::vector<int>& sched = p.host.sched;
    sched.clear();
    p.subset_pair_off.clear();
    const int waves = fp_waves(p.fp_pf);
    // the subsets' schedules are independent: built on host threads, then
    // laid out in subset order
    std::vector<std::vector<int>> part(p.n_subsets);
    parallel_for(p.n_subsets, [&](int sb) {
        const int f0 = p.subset_fp_off[sb], f1 = p.subset_fp_off[sb + 1];
        fp_pair_schedule(g, p.subset_fp.data() + f0, f1 - f0, p.proj,
                         fp_group(p.proj, f1 - f0, waves), part[sb], waves);
    });
    for (int sb = 0; sb < p.n_subsets; sb++) {
        p.subset_pair_off.push_back((int)sched.size());
        sched.insert(sched.end(), part[sb].begin(), part[sb].end());
    }
    return ATR_OK;
}

// Device half: uploads the shape-dependent tables and the per-step kernel
// arguments (which hold device pointers).
static atr_status sart_tables_device(Plan& p)
{
    Geometry& g = p.geo;
    size_t& B = p.device_bytes;
    atr_status st;
    if ((st = reupload(p, &p.d_bp_tiles, g.bp_tiles, B, p.tab_bytes[0])) != ATR_OK) return st;
    if ((st = reupload(p, &p.d_fp_range, g.fp_range, B, p.tab_bytes[1])) != ATR_OK) return st;
    if ((st = reupload(p, &p.d_fp_geo_off, g.fp_geo_off, B, p.tab_bytes[2])) != ATR_OK) return st;
    if ((st = reupload(p, &p.d_fp_len, g.fp_len, B, p.tab_bytes[3])) != ATR_OK) return st;
    if ((st = reupload(p, &p.d_fp_geo, g.fp_geo, B, p.tab_bytes[4])) != ATR_OK) return st;
    if ((st = reupload(p, &p.d_fp_desc, g.fp_desc, B, p.tab_bytes[5])) != ATR_OK) return st;
    if ((st = reupload(p, &p.d_fp_pairs, p.host.sched, B, p.tab_bytes[6])) != ATR_OK) return st;
    // kernel arguments of every SART step, read by the kernels from device
    // memory: [0] initial back-projection, [1 + s] subset s
    const int n_all = (int)p.all_tilts.size();
    // kept in the plan's host tables: a stream-ordered upload reads them later
    std::vector<BpArgs>& bpa = p.host.bp_args;
    std::vector<FpArgs>& fpa = p.host.fp_args;
    bpa.clear();
    fpa.clear();
    bpa.push_back(bp_args(p, p.d_sino, p.n, p.d_lists, n_all, 1.0f, true));
    for (int s = 0; s < p.n_subsets; s++) {
        const int b0 = p.subset_off[s], b1 = p.subset_off[s + 1];
        const int f0 = p.subset_fp_off[s], f1 = p.subset_fp_off[s + 1];
        bpa.push_back(bp_args(p, p.d_diff, p.max_subset, p.d_lists + n_all + b0, b1 - b0, p.relax,
                              false));
        fpa.push_back(fp_args(p, p.d_fp_lists + f0, f1 - f0, p.d_fp_pairs + p.subset_pair_off[s]));
    }
    if ((st = reupload(p, &p.d_bp_args, bpa, B, p.tab_bytes[7])) != ATR_OK) return st;
    if ((st = reupload(p, &p.d_fp_args, fpa, B, p.tab_bytes[8])) != ATR_OK) return st;
    return ATR_OK;
}

// Rebuilds the shape-dependent tables (autotuner, atr_plan_set_tile_shape).
static atr_status sart_tables(Plan& p, int tz, int chunk, int bz, int pf)
{
    atr_status st = sart_tables_host(p, tz, chunk, bz, pf);
    if (st != ATR_OK) return st;
    return sart_tables_device(p);
}

// Host half of a plan: every table the kernels read, computed on the host in
// FP64 from the descriptor (and the arrays it points to).  No device work, so
// it can run on any thread (the server builds plans while earlier series
// are on the GPU).
static atr_status plan_host(Plan& p)
{
    const atr_series_desc& d = p.desc;
    Geometry& g = p.geo;
    HostTables& h = p.host;
    cos_sin_table(d.tilt_deg, p.n, g.cos_sin);
    h.cs.resize(p.n);
    for (int t = 0; t < p.n; t++) h.cs[t] = make_float2(g.cos_sin[2 * t], g.cos_sin[2 * t + 1]);
    h.excl.assign(p.n, 0);
    if (d.exclude)
        for (int t = 0; t < p.n; t++) h.excl[t] = d.exclude[t] ? 1 : 0;
    if (p.prestage) {
        // Correct/GCorrPatchShift.cu:104,142: angle = axis * (3.141592654f/180)
        // in float, cos/sin in double.
        const float d2r = 3.141592654f / 180.0f;
        h.axis_cs.assign(2 * (size_t)p.n, 0.0);
        h.shifts.assign(2 * (size_t)p.n, 0.0f);
        for (int t = 0; t < p.n; t++) {
            float ang = d.tilt_axis_deg[t] * d2r;
            h.axis_cs[2 * t] = std::cos((double)ang);
            h.axis_cs[2 * t + 1] = std::sin((double)ang);
            if (d.shift_xy) {
                h.shifts[2 * t] = d.shift_xy[2 * t];
                h.shifts[2 * t + 1] = d.shift_xy[2 * t + 1];
            }
        }
        if (p.fill) fill_table(d.fill_seed, p.aln_x * p.aln_y, h.fill_aln);
        if (p.n_patches > 0)
            h.local.assign(d.local_params, d.local_params + (size_t)p.n * 5 * p.n_patches);
    }
    if (p.binned) {
        FftDesc* fds[4] = {&p.fft_bin_ax, &p.fft_bin_ay, &p.fft_bin_by, &p.fft_bin_bx};
        const int sizes[4] = {p.aln_x, p.aln_y, p.bin_y, p.bin_x};
        for (int i = 0; i < 4; i++) {
            make_fft(sizes[i], *fds[i]);
            h.tw_bin[i] = twiddles(sizes[i]);
        }
    }
    if (p.method == ATR_METHOD_SART) {
        build_bp_intervals(g.cos_sin, p.n, p.vol_x, p.vol_z, p.proj, g.bp_lohi);
        h.lists = p.all_tilts;
        h.lists.insert(h.lists.end(), p.subset_tilts.begin(), p.subset_tilts.end());
        fill_table(d.fill_seed ^ 0x5A5A5A5A5A5A5A5Aull, p.vol_x * p.vol_z, h.fill_vol);
        // default forward-projector staging: register prefetch where it is
        // built (VolX 1024, 2048: measured faster at the full band height),
        // else direct
        return sart_tables_host(p, p.bp_tz, p.bp_chunk,
                                fp_balanced_rows(fp_band_rows(p.vol_x), p.vol_z),
                                fp_rp_supported(p.vol_x) ? FP_PF_REG : 0);
    }
    build_wbp_intervals(g.cos_sin, p.n, p.vol_x, p.vol_z, p.proj, g.wbp_lohi);
    // 2x-grid back-projection: tile windows/classes (excluded tilts are left
    // out of the list, GBackProjWbp.cu:31)
    build_bp_tiles(g.cos_sin, g.wbp_lohi, p.n, 2 * p.vol_x, 2 * p.vol_z, 0.5,
                   (double)(p.proj / 2) + 0.5, BP_TZ, h.wbp_tiles);
    h.filter = ramp_filter(p.proj);
    make_fft(p.proj, p.fft_proj);
    make_fft(2 * p.vol_x, p.fft_x2);
    make_fft(2 * p.vol_z, p.fft_z2);
    make_fft(p.vol_z, p.fft_z);
    make_fft(p.vol_x, p.fft_x);
    h.tw_proj = twiddles(p.proj);
    h.tw_x2 = twiddles(2 * p.vol_x);
    h.tw_z2 = twiddles(2 * p.vol_z);
    h.tw_z = twiddles(p.vol_z);
    h.tw_x = twiddles(p.vol_x);
    return ATR_OK;
}

// Device half: allocations (through the plan's allocator) and uploads of the
// host tables, then the kernel arguments that hold device pointers.  The
// host copies of the uploaded tables are released afterwards.
static atr_status plan_device(Plan& p)
{
    HostTables& h = p.host;
    Geometry& g = p.geo;
    size_t& B = p.device_bytes;
    B = 0;
    atr_status st;
    if ((st = upload(p, &p.d_cs, h.cs, B)) != ATR_OK) return st;
    if ((st = upload(p, &p.d_excl, h.excl, B)) != ATR_OK) return st;
    if (p.prestage) {
        if ((st = upload(p, &p.d_axis_cs, h.axis_cs, B)) != ATR_OK) return st;
        if ((st = upload(p, &p.d_shifts, h.shifts, B)) != ATR_OK) return st;
        if ((st = dalloc(p, (void**)&p.d_aligned, sizeof(float) * p.n * (size_t)p.aln_x * p.aln_y, B)) != ATR_OK) return st;
        if ((st = dalloc(p, (void**)&p.d_row_stats, sizeof(double) * 2 * p.n * (size_t)p.aln_y, B)) != ATR_OK) return st;
        if ((st = dalloc(p, (void**)&p.d_tilt_stats, sizeof(double) * 2 * p.n, B)) != ATR_OK) return st;
        if (p.fill && (st = upload(p, &p.d_fill_aln, h.fill_aln, B)) != ATR_OK) return st;
        if (p.n_patches > 0 && (st = upload(p, &p.d_local, h.local, B)) != ATR_OK) return st;
    }
    if (p.binned) {
        const int kx = p.bin_x / 2 + 1;
        if ((st = dalloc(p, (void**)&p.d_binned, sizeof(float) * p.n * (size_t)p.bin_x * p.bin_y, B)) != ATR_OK) return st;
        if ((st = dalloc(p, (void**)&p.d_bin_t, sizeof(float2) * p.n * (size_t)kx * p.aln_y, B)) != ATR_OK) return st;
        if ((st = dalloc(p, (void**)&p.d_bin_u, sizeof(float2) * p.n * (size_t)kx * p.bin_y, B)) != ATR_OK) return st;
        FftDesc* fds[4] = {&p.fft_bin_ax, &p.fft_bin_ay, &p.fft_bin_by, &p.fft_bin_bx};
        for (int i = 0; i < 4; i++) {
            if ((st = upload(p, &p.d_tw_bin[i], h.tw_bin[i], B)) != ATR_OK) return st;
            fds[i]->tw = p.d_tw_bin[i];
        }
    }
    size_t sb = storage_bytes(p.storage);
    if ((st = dalloc(p, &p.d_sino, (size_t)p.vol_y * p.n * p.sino_stride * sb, B)) != ATR_OK) return st;
    if ((st = dalloc(p, (void**)&p.d_sflag, sizeof(int) * p.vol_y, B)) != ATR_OK) return st;

    if (p.method == ATR_METHOD_SART) {
        if ((st = upload(p, &p.d_bp_lohi, g.bp_lohi, B)) != ATR_OK) return st;
        if ((st = upload(p, &p.d_lists, h.lists, B)) != ATR_OK) return st;
        if ((st = upload(p, &p.d_fp_lists, p.subset_fp, B)) != ATR_OK) return st;
        if ((st = upload(p, &p.d_fill_vol, h.fill_vol, B)) != ATR_OK) return st;
        size_t vox = (size_t)p.vol_x * p.vol_y * p.vol_z;
        if ((st = dalloc(p, (void**)&p.d_vol, sizeof(float) * vox, B)) != ATR_OK) return st;
        if ((st = dalloc(p, &p.d_diff, (size_t)p.vol_y * p.max_subset * p.sino_stride * sb, B)) != ATR_OK) return st;
        if ((st = dalloc(p, (void**)&p.d_vflag, sizeof(int) * p.vol_y, B)) != ATR_OK) return st;
        if ((st = dalloc(p, (void**)&p.d_vstats, sizeof(double) * 2 * p.vol_y, B)) != ATR_OK) return st;
        if ((st = sart_tables_device(p)) != ATR_OK) return st;
    } else {
        if ((st = upload(p, &p.d_wbp_lohi, g.wbp_lohi, B)) != ATR_OK) return st;
        if ((st = upload(p, &p.d_wbp_tiles, h.wbp_tiles, B)) != ATR_OK) return st;
        if ((st = upload(p, &p.d_lists, p.all_tilts, B)) != ATR_OK) return st;
        if ((st = dalloc(p, (void**)&p.d_wbp_grid,
                         (size_t)p.wbp_slab * 4 * p.vol_z * p.vol_x * sizeof(float), B)) != ATR_OK)
            return st;
        std::vector<BpArgs>& args = p.host.bp_args;
        args.clear();
        for (int y0 = 0; y0 < p.vol_y; y0 += p.wbp_slab) {
            BpArgs a{};
            a.sino = (const char*)p.d_sino + (size_t)y0 * p.n * p.sino_stride * storage_bytes(p.storage);
            a.rows_per_slice = p.n;
            a.stride = p.sino_stride;
            a.proj = p.proj;
            a.list = p.d_lists;
            a.n_list = (int)p.all_tilts.size();
            a.cs = p.d_cs;
            a.lohi = p.d_wbp_lohi;
            a.tiles = p.d_wbp_tiles;
            a.n_ztiles = (2 * p.vol_z + BP_TZ - 1) / BP_TZ;
            a.n_xtiles = (2 * p.vol_x + BP_TX - 1) / BP_TX;
            a.relax = 1.0f;
            a.positivity = 0;
            a.scale = 0.5f;
            a.offset = (float)(p.proj / 2) + 0.5f;
            a.sflag = p.d_sflag + y0;
            a.vol = p.d_wbp_grid;
            a.vol_x = 2 * p.vol_x;
            a.vol_z = 2 * p.vol_z;
            a.vflag = nullptr;
            args.push_back(a);
        }
        if ((st = upload(p, &p.d_bp_args, args, B)) != ATR_OK) return st;
        if ((st = upload(p, &p.d_filter, h.filter, B)) != ATR_OK) return st;
        if ((st = upload(p, &p.d_tw_proj, h.tw_proj, B)) != ATR_OK) return st;
        if ((st = upload(p, &p.d_tw_x2, h.tw_x2, B)) != ATR_OK) return st;
        if ((st = upload(p, &p.d_tw_z2, h.tw_z2, B)) != ATR_OK) return st;
        if ((st = upload(p, &p.d_tw_z, h.tw_z, B)) != ATR_OK) return st;
        if ((st = upload(p, &p.d_tw_x, h.tw_x, B)) != ATR_OK) return st;
        p.fft_proj.tw = p.d_tw_proj;
        p.fft_x2.tw = p.d_tw_x2;
        p.fft_z2.tw = p.d_tw_z2;
        p.fft_z.tw = p.d_tw_z;
        p.fft_x.tw = p.d_tw_x;
        size_t kxn = p.vol_x / 2 + 1;
        if ((st = dalloc(p, (void**)&p.d_wbp_t, (size_t)p.wbp_slab * 2 * p.vol_z * kxn * sizeof(float2), B)) != ATR_OK) return st;
        if ((st = dalloc(p, (void**)&p.d_wbp_u, (size_t)p.wbp_slab * p.vol_z * kxn * sizeof(float2), B)) != ATR_OK) return st;
    }
    // a stream-ordered upload still reads the host tables: the caller drops
    // them once the stream has passed (plan_drop_host_tables)
    if (!p.upload_stream_set) p.host = HostTables();
    return ATR_OK;
}

// Frees the plan's device memory and events; no device synchronisation.
static void plan_free(Plan& p)
{
    void** ptrs[] = {(void**)&p.d_cs, (void**)&p.d_bp_lohi, (void**)&p.d_bp_tiles,
                     (void**)&p.d_wbp_lohi, (void**)&p.d_fp_range, (void**)&p.d_fp_geo_off,
                     (void**)&p.d_fp_len, (void**)&p.d_fp_geo, (void**)&p.d_fp_desc,
                     (void**)&p.d_fp_pairs, (void**)&p.d_lists, (void**)&p.d_fp_lists,
                     (void**)&p.d_fill_aln, (void**)&p.d_fill_vol, (void**)&p.d_filter,
                     (void**)&p.d_tw_proj, (void**)&p.d_tw_x2, (void**)&p.d_tw_z2,
                     (void**)&p.d_tw_z, (void**)&p.d_tw_x, (void**)&p.d_axis_cs,
                     (void**)&p.d_shifts, (void**)&p.d_excl, (void**)&p.d_aligned,
                     (void**)&p.d_row_stats, (void**)&p.d_tilt_stats, (void**)&p.d_sino,
                     (void**)&p.d_diff, (void**)&p.d_vol, (void**)&p.d_sflag, (void**)&p.d_vflag,
                     (void**)&p.d_vstats, (void**)&p.d_wbp_t, (void**)&p.d_wbp_u,
                     (void**)&p.d_bp_args, (void**)&p.d_fp_args, (void**)&p.d_wbp_grid,
                     (void**)&p.d_wbp_tiles, (void**)&p.d_local, (void**)&p.d_binned,
                     (void**)&p.d_bin_t, (void**)&p.d_bin_u, (void**)&p.d_tw_bin[0],
                     (void**)&p.d_tw_bin[1], (void**)&p.d_tw_bin[2], (void**)&p.d_tw_bin[3]};
    for (void** q : ptrs) {
        plan_mfree(p, *q);
        *q = nullptr;
    }
    for (size_t& b : p.tab_bytes) b = 0;
    p.device_bytes = 0;
    for (auto& e : p.ev_pool) (void)hipEventDestroy(e);
    for (auto& e : p.ev_pending) {
        (void)hipEventDestroy(e.second.first);
        (void)hipEventDestroy(e.second.second);
    }
    p.ev_pool.clear();
    p.ev_pending.clear();
}

// Plan construction in two halves, for callers that build the host tables
// on one thread and allocate on another (serve.cpp).
atr_status plan_prepare(const atr_series_desc* d, Plan& p)
{
    atr_status st = plan_sizes(d, p);
    if (st == ATR_OK) st = plan_host(p);
    return st;
}

atr_status plan_upload(Plan& p)
{
    atr_status st = plan_device(p);
    if (st != ATR_OK) plan_free(p);
    return st;
}

void plan_release(Plan& p) { plan_free(p); }

// ---- profiling -----------------------------------------------------------
static hipEvent_t take_event(Plan& p)
{
    if (!p.ev_pool.empty()) {
        hipEvent_t e = p.ev_pool.back();
        p.ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

static void fold_events(Plan& p)
{
    for (auto& e : p.ev_pending) {
        float ms = 0.0f;
        if (hipEventSynchronize(e.second.second) == hipSuccess &&
            hipEventElapsedTime(&ms, e.second.first, e.second.second) == hipSuccess) {
            p.launches[e.first] += 1;
            p.total_ms[e.first] += ms;
        }
        p.ev_pool.push_back(e.second.first);
        p.ev_pool.push_back(e.second.second);
    }
    p.ev_pending.clear();
}

template <class F>
static void timed(Plan& p, int kernel, hipStream_t s, F&& fn)
{
    if (!p.profiling) {
        fn();
        return;
    }
    hipEvent_t a = take_event(p), b = take_event(p);
    if (a == nullptr || b == nullptr) {
        fn();
        return;
    }
    (void)hipEventRecord(a, s);
    fn();
    (void)hipEventRecord(b, s);
    p.ev_pending.push_back({kernel, {a, b}});
    if (p.ev_pending.size() > 4096) fold_events(p);
}

// ---- execution -----------------------------------------------------------
// Pre-stage (alignment + fill statistics) and Fourier binning: what the
// weighting / filter kernels read afterwards is stage_source(p, in).
static void stage_prepare(Plan& p, const float* in, hipStream_t s)
{
    if (p.prestage || p.binned) {
        timed(p, K_PRESTAGE, s, [&] {
            if (p.prestage) {
                launch_prestage(p, in, s);
                launch_tilt_stats(p, s);
            }
            if (p.binned) launch_bin(p, in, s);
        });
    }
}

// SART before the iterations: pre-stage, weighting into the slice-major
// sinogram, the volume state (zero, or the caller's SLICE volume).
atr_status sart_prepare(Plan& p, const float* in, float* out, int accumulate, hipStream_t s)
{
    stage_prepare(p, in, s);
    launch_zero_int(p.d_sflag, p.vol_y, s);
    launch_zero_int(p.d_vflag, p.vol_y, s);
    timed(p, K_WEIGHT, s, [&] { launch_weight_sart(p, in, s); });
    if (accumulate && p.layout == ATR_LAYOUT_SLICE) launch_load_volume(p, out, s);
    else launch_zero(p.d_vol, (size_t)p.vol_x * p.vol_y * p.vol_z, s);
    return ATR_OK;
}

// SART after the iterations: empty fill (CTomoSart.cpp:86-89) + layout.
void sart_finish(Plan& p, float* out, hipStream_t s)
{
    timed(p, K_OUTPUT, s, [&] {
        launch_vol_stats(p, s);
        launch_output(p, out, s);
    });
}

atr_status plan_execute(Plan& p, const float* in, float* out, int accumulate, hipStream_t s)
{
    if (in == nullptr || out == nullptr) return fail(ATR_ERR_INVALID_ARG, "NULL buffer");
    (void)hipGetLastError();
    if (p.method == ATR_METHOD_SART) {
        sart_prepare(p, in, out, accumulate, s);
        // initial back-projection of every projection, relax 1 (CTomoSart.cpp:68-69)
        if (!p.all_tilts.empty()) timed(p, K_BP, s, [&] { launch_bp(p, 0, s); });
        for (int it = 0; it < p.iters; it++) {
            for (int sb = 0; sb < p.n_subsets; sb++) {
                timed(p, K_FP, s, [&] { launch_fp_diff(p, sb, s); });
                if (p.subset_off[sb + 1] > p.subset_off[sb])
                    timed(p, K_BP, s, [&] { launch_bp(p, 1 + sb, s); });
            }
        }
        sart_finish(p, out, s);
    } else {
        stage_prepare(p, in, s);
        launch_zero_int(p.d_sflag, p.vol_y, s);
        timed(p, K_WEIGHT, s, [&] { launch_filter_wbp(p, in, s); });
        for (int y0 = 0, k = 0; y0 < p.vol_y; y0 += p.wbp_slab, k++) {
            int ny = std::min(p.wbp_slab, p.vol_y - y0);
            timed(p, K_WBP_BP, s, [&] {
                launch_wbp_bp(p, k, ny, s);
                launch_wbp_rowfft(p, ny, s);
            });
            timed(p, K_WBP_COL, s, [&] { launch_wbp_colfft(p, y0, ny, s); });
            timed(p, K_WBP_ROW, s, [&] { launch_wbp_rowc2r(p, y0, ny, out, s); });
        }
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "kernel launch");
    return ATR_OK;
}

// ---- autotuner -------------------------------------------------------------
// Times the candidate tile shapes on the plan's own geometry and keeps the
// fastest: back-projection z rows per thread x tilts per LDS stage, then
// forward-projection band rows.  Choices are cached per (sizes, tilts,
// subsets, storage), so later plans of the same shape skip the timing.
struct TuneKey {
    int vol_x, vol_z, proj, n, n_subsets, storage;
    bool operator<(const TuneKey& o) const
    {
        return std::tie(vol_x, vol_z, proj, n, n_subsets, storage) <
               std::tie(o.vol_x, o.vol_z, o.proj, o.n, o.n_subsets, o.storage);
    }
};
struct TuneChoice {
    int tz, chunk, bz, pf;
};
static std::mutex g_tune_mu;
static std::map<TuneKey, TuneChoice> g_tune_cache;

template <class F>
static float time_ms(hipStream_t s, int reps, F&& fn)
{
    hipEvent_t a = nullptr, b = nullptr;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    fn();  // warm-up
    (void)hipEventRecord(a, s);
    for (int r = 0; r < reps; r++) fn();
    (void)hipEventRecord(b, s);
    float ms = 0.0f;
    (void)hipEventSynchronize(b);
    (void)hipEventElapsedTime(&ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms / reps;
}

static atr_status plan_autotune(Plan& p, hipStream_t s, int* tz_out, int* chunk_out, int* bz_out,
                                int* pf_out)
{
    if (p.method != ATR_METHOD_SART) {
        if (tz_out) *tz_out = 0;
        if (chunk_out) *chunk_out = 0;
        if (bz_out) *bz_out = 0;
        if (pf_out) *pf_out = 0;
        return ATR_OK;
    }
    if (p.batch_refs > 0)
        return fail(ATR_ERR_INVALID_ARG,
                    "autotune: the plan belongs to a batch (destroy the batch first)");
    const TuneKey key{p.vol_x, p.vol_z, p.proj, p.n, p.n_subsets, p.storage};
    TuneChoice best{p.bp_tz, p.bp_chunk, p.geo.fp_bz, p.fp_pf};
    bool cached = false;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            best = it->second;
            cached = true;
        }
    }
    atr_status st;
    if (!cached) {
        // defined contents for the timing runs (the next execute starts over)
        const size_t sb = p.storage == ATR_STORAGE_F32 ? 4 : 2;
        ATR_HIP(hipMemsetAsync(p.d_sino, 0, (size_t)p.vol_y * p.n * p.sino_stride * sb, s));
        ATR_HIP(hipMemsetAsync(p.d_diff, 0, (size_t)p.vol_y * p.max_subset * p.sino_stride * sb, s));
        ATR_HIP(hipMemsetAsync(p.d_vol, 0, sizeof(float) * p.vol_x * p.vol_y * p.vol_z, s));
        ATR_HIP(hipMemsetAsync(p.d_sflag, 0, sizeof(int) * p.vol_y, s));
        ATR_HIP(hipMemsetAsync(p.d_vflag, 0, sizeof(int) * p.vol_y, s));
        ATR_HIP(hipStreamSynchronize(s));
        (void)hipGetLastError();
        std::vector<int> bzs;
        fp_band_candidates(p.vol_x, p.vol_z, bzs);
        const bool has_bp = p.subset_off[1] > p.subset_off[0];
        // A candidate whose launch fails (an error, not a time) is dropped: a
        // failed launch would otherwise time as ~0 ms and win.
        float best_bp = 1e30f;
        bool bp_ok = false;
        const int shapes[4][2] = {{16, 16}, {16, 8}, {8, 16}, {8, 8}};
        for (auto& sh : shapes) {
            if ((st = sart_tables(p, sh[0], sh[1], bzs[0], 0)) != ATR_OK) return st;
            const float ms = time_ms(s, 3, [&] { launch_bp(p, has_bp ? 1 : 0, s); });
            if (hipGetLastError() != hipSuccess) continue;
            if (ms < best_bp) {
                best_bp = ms;
                best.tz = sh[0];
                best.chunk = sh[1];
                bp_ok = true;
            }
        }
        // forward projection: the band heights of the direct-staging variant,
        // then those of the band-prefetching one where VolX allows it
        std::vector<std::pair<int, int>> fps;
        for (int bz : bzs) fps.push_back({bz, 0});
        if (fp_pf_supported(p.vol_x)) {
            std::vector<int> pfz;
            fp_band_candidates(p.vol_x, p.vol_z, pfz, FP_PF_DMA);
            for (int bz : pfz) fps.push_back({bz, FP_PF_DMA});
        }
        if (fp_rp_supported(p.vol_x))
            for (int bz : bzs) fps.push_back({bz, FP_PF_REG});
        float best_fp = 1e30f;
        bool fp_ok = false;
        for (auto& c : fps) {
            if ((st = sart_tables(p, best.tz, best.chunk, c.first, c.second)) != ATR_OK) return st;
            const float ms = time_ms(s, 3, [&] { launch_fp_diff(p, 0, s); });
            if (hipGetLastError() != hipSuccess) continue;
            if (ms < best_fp) {
                best_fp = ms;
                best.bz = c.first;
                best.pf = c.second;
                fp_ok = true;
            }
        }
        if (!bp_ok || !fp_ok) {
            // nothing launched: restore the default shape and report it
            (void)sart_tables(p, BP_TZ, BP_CHUNK, bzs[0], 0);
            return fail(ATR_ERR_HIP, "autotune: no candidate tile shape launched");
        }
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_cache[key] = best;
    }
    if ((st = sart_tables(p, best.tz, best.chunk, best.bz, best.pf)) != ATR_OK) return st;
    if (tz_out) *tz_out = best.tz;
    if (chunk_out) *chunk_out = best.chunk;
    if (bz_out) *bz_out = best.bz;
    if (pf_out) *pf_out = p.fp_pf;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "autotune");
    return ATR_OK;
}

atr_status desc_volume(const atr_series_desc* d, int* vol_x, int* vol_y, int* vol_z)
{
    Plan p;
    atr_status st = plan_sizes(d, p);
    if (st != ATR_OK) return st;
    *vol_x = p.vol_x;
    *vol_y = p.vol_y;
    *vol_z = p.vol_z;
    return ATR_OK;
}

}  // namespace atr

// ============================================================================
// C ABI
// ============================================================================
using namespace atr;

extern "C" {

void atr_series_desc_init(atr_series_desc* d)
{
    if (!d) return;
    std::memset(d, 0, sizeof(*d));
    d->random_fill = 1;
    d->bin = 1.0f;
    d->vol_z = 100;
    d->method = ATR_METHOD_SART;
    d->sart_iters = 20;   // -Sart 20 3 (CAtInput.cpp:69-70)
    d->sart_subsets = 1;
    d->sart_relax = -1.0f;
    d->subset_order = ATR_SUBSETS_CONTIGUOUS;
    d->positivity = 1;
    d->storage = ATR_STORAGE_F32;
    d->out_layout = ATR_LAYOUT_XZY;
}

atr_status atr_plan_create(const atr_series_desc* desc, atr_plan** plan)
{
    if (plan == nullptr) return fail(ATR_ERR_INVALID_ARG, "plan is NULL");
    *plan = nullptr;
    atr_plan* h = new (std::nothrow) atr_plan();
    if (!h) return fail(ATR_ERR_OUT_OF_MEMORY, "host allocation failed");
    atr_status st = plan_sizes(desc, h->p);
    if (st == ATR_OK) st = plan_host(h->p);
    if (st == ATR_OK) st = plan_device(h->p);
    if (st != ATR_OK) {
        plan_free(h->p);
        delete h;
        return st;
    }
    *plan = h;
    return ATR_OK;
}

atr_status atr_estimate_device_bytes(const atr_series_desc* desc, size_t* bytes)
{
    if (bytes == nullptr) return fail(ATR_ERR_INVALID_ARG, "bytes is NULL");
    Plan p;
    atr_status st = plan_sizes(desc, p);
    if (st != ATR_OK) return st;
    *bytes = plan_bytes_estimate(p);
    return ATR_OK;
}

atr_status atr_aligned_size(int raw_x, int raw_y, float tilt_axis_deg, int* aln_x, int* aln_y)
{
    if (!aln_x || !aln_y) return fail(ATR_ERR_INVALID_ARG, "atr_aligned_size: NULL argument");
    if (raw_x < 1 || raw_y < 1 || !std::isfinite(tilt_axis_deg))
        return fail(ATR_ERR_INVALID_ARG, "atr_aligned_size: bad image size or axis");
    aligned_size(raw_x, raw_y, tilt_axis_deg, aln_x, aln_y);
    return ATR_OK;
}

atr_status atr_plan_get_volume_size(const atr_plan* plan, int* vx, int* vy, int* vz)
{
    if (!plan) return fail(ATR_ERR_INVALID_ARG, "plan is NULL");
    if (vx) *vx = plan->p.vol_x;
    if (vy) *vy = plan->p.vol_y;
    if (vz) *vz = plan->p.vol_z;
    return ATR_OK;
}

size_t atr_plan_device_bytes(const atr_plan* plan) { return plan ? plan->p.device_bytes : 0; }

atr_status atr_execute(atr_plan* plan, const float* proj_dev, float* vol_dev, int accumulate,
                       void* stream)
{
    if (!plan) return fail(ATR_ERR_INVALID_ARG, "plan is NULL");
    return plan_execute(plan->p, proj_dev, vol_dev, accumulate, (hipStream_t)stream);
}

atr_status atr_execute_host(atr_plan* plan, const float* proj_host, float* vol_host)
{
    if (!plan || !proj_host || !vol_host) return fail(ATR_ERR_INVALID_ARG, "NULL argument");
    Plan& p = plan->p;
    size_t in_elems = p.in_tilt_stride * (p.n - 1) + p.in_row_stride * (p.in_y - 1) + p.in_x;
    size_t out_elems = (size_t)p.vol_x * p.vol_y * p.vol_z;
    float *d_in = nullptr, *d_out = nullptr;
    hipStream_t s = nullptr;
    atr_status st = ATR_OK;
    hipError_t e = hipMalloc((void**)&d_in, in_elems * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&d_out, out_elems * sizeof(float));
    if (e == hipSuccess) e = hipStreamCreate(&s);
    if (e == hipSuccess) e = hipMemcpyAsync(d_in, proj_host, in_elems * sizeof(float), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        st = plan_execute(p, d_in, d_out, 0, s);
        if (st == ATR_OK)
            e = hipMemcpyAsync(vol_host, d_out, out_elems * sizeof(float), hipMemcpyDeviceToHost, s);
        if (st == ATR_OK && e == hipSuccess) e = hipStreamSynchronize(s);
    }
    if (e != hipSuccess) st = hip_fail(e, "atr_execute_host");
    if (s) (void)hipStreamDestroy(s);
    if (d_in) (void)hipFree(d_in);
    if (d_out) (void)hipFree(d_out);
    return st;
}

atr_status atr_plan_set_tile_shape(atr_plan* plan, int bp_tz, int bp_chunk, int fp_rows,
                                   int fp_prefetch)
{
    if (!plan) return fail(ATR_ERR_INVALID_ARG, "plan is NULL");
    Plan& p = plan->p;
    if (p.method != ATR_METHOD_SART) return fail(ATR_ERR_INVALID_ARG, "tile shapes are SART settings");
    if ((bp_tz != 8 && bp_tz != 16) || (bp_chunk != 8 && bp_chunk != 16))
        return fail(ATR_ERR_INVALID_ARG, "bp_tz and bp_chunk must be 8 or 16");
    const int pf = fp_prefetch;
    if (pf < 0 || pf > FP_PF_REG)
        return fail(ATR_ERR_INVALID_ARG, "fp_prefetch must be 0, 1 or 2");
    if (pf == FP_PF_DMA && !fp_pf_supported(p.vol_x))
        return fail(ATR_ERR_INVALID_ARG, "the band-prefetching forward projector needs VolX a "
                                         "multiple of 256 (VolX " + std::to_string(p.vol_x) + ")");
    if (pf == FP_PF_REG && !fp_rp_supported(p.vol_x))
        return fail(ATR_ERR_INVALID_ARG, "the register-prefetching forward projector needs VolX "
                                         "1024 or 2048 (VolX " + std::to_string(p.vol_x) + ")");
    const int lim = fp_band_rows(p.vol_x, pf);
    if (fp_rows < 1 || fp_rows > lim)
        return fail(ATR_ERR_INVALID_ARG,
                    "fp band rows outside [1, " + std::to_string(lim) + "] (the LDS limit at VolX " +
                        std::to_string(p.vol_x) + ")");
    if (p.batch_refs > 0)
        return fail(ATR_ERR_INVALID_ARG,
                    "set_tile_shape: the plan belongs to a batch (destroy the batch first)");
    return sart_tables(p, bp_tz, bp_chunk, fp_rows, pf);
}

atr_status atr_fp_band_candidates(int vol_x, int vol_z, int* rows, int cap, int* n)
{
    if (n == nullptr || vol_x < 2 || vol_z < 2)
        return fail(ATR_ERR_INVALID_ARG, "atr_fp_band_candidates: bad size or NULL count");
    std::vector<int> c;
    fp_band_candidates(vol_x, vol_z, c);
    *n = (int)c.size();
    if (rows)
        for (int i = 0; i < std::min(cap, (int)c.size()); i++) rows[i] = c[i];
    return ATR_OK;
}

atr_status atr_plan_autotune(atr_plan* plan, void* stream, int* bp_tz, int* bp_chunk,
                             int* fp_band_rows_out, int* fp_prefetch)
{
    if (!plan) return fail(ATR_ERR_INVALID_ARG, "plan is NULL");
    return plan_autotune(plan->p, (hipStream_t)stream, bp_tz, bp_chunk, fp_band_rows_out,
                         fp_prefetch);
}

atr_status atr_plan_set_profiling(atr_plan* plan, int enable)
{
    if (!plan) return fail(ATR_ERR_INVALID_ARG, "plan is NULL");
    plan->p.profiling = enable != 0;
    return ATR_OK;
}

atr_status atr_plan_kernel_stats(const atr_plan* plan, int kernel, int* launches, double* total_ms)
{
    if (!plan || kernel < 0 || kernel >= K_COUNT) return fail(ATR_ERR_INVALID_ARG, "bad argument");
    Plan& p = const_cast<atr_plan*>(plan)->p;
    fold_events(p);
    if (launches) *launches = p.launches[kernel];
    if (total_ms) *total_ms = p.total_ms[kernel];
    return ATR_OK;
}

atr_status atr_plan_reset_stats(atr_plan* plan)
{
    if (!plan) return fail(ATR_ERR_INVALID_ARG, "plan is NULL");
    fold_events(plan->p);
    for (int k = 0; k < K_COUNT; k++) {
        plan->p.launches[k] = 0;
        plan->p.total_ms[k] = 0.0;
    }
    return ATR_OK;
}

void atr_plan_destroy(atr_plan* plan)
{
    if (!plan) return;
    (void)hipDeviceSynchronize();
    plan_free(plan->p);
    delete plan;
}

const char* atr_last_error_string(void) { return g_err.c_str(); }

const char* atr_version(void) { return "atr_recon 0.1 (gfx950)"; }

// ---- device helpers --------------------------------------------------------
atr_status atr_device_count(int* count)
{
    if (!count) return fail(ATR_ERR_INVALID_ARG, "count is NULL");
    *count = 0;
    hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess) {
        *count = 0;
        return hip_fail(e, "hipGetDeviceCount");
    }
    return ATR_OK;
}
atr_status atr_set_device(int device)
{
    ATR_HIP(hipSetDevice(device));
    return ATR_OK;
}
atr_status atr_malloc(void** ptr, size_t bytes)
{
    if (!ptr) return fail(ATR_ERR_INVALID_ARG, "ptr is NULL");
    ATR_HIP(hipMalloc(ptr, bytes));
    return ATR_OK;
}
atr_status atr_free(void* ptr)
{
    ATR_HIP(hipFree(ptr));
    return ATR_OK;
}
atr_status atr_memcpy_h2d(void* dst, const void* src, size_t bytes)
{
    ATR_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return ATR_OK;
}
atr_status atr_memcpy_d2h(void* dst, const void* src, size_t bytes)
{
    ATR_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return ATR_OK;
}
atr_status atr_memset(void* ptr, int value, size_t bytes)
{
    ATR_HIP(hipMemset(ptr, value, bytes));
    return ATR_OK;
}
atr_status atr_copy_kernel(void* dst, const void* src, size_t bytes, void* stream)
{
    if ((bytes % 16) != 0 || ((uintptr_t)dst % 16) != 0 || ((uintptr_t)src % 16) != 0)
        return fail(ATR_ERR_INVALID_ARG, "atr_copy_kernel needs 16-byte aligned sizes and pointers");
    launch_copy(src, dst, bytes, (hipStream_t)stream);
    ATR_HIP(hipGetLastError());
    return ATR_OK;
}
atr_status atr_synchronize(void)
{
    ATR_HIP(hipDeviceSynchronize());
    return ATR_OK;
}
atr_status atr_stream_create(void** stream)
{
    if (!stream) return fail(ATR_ERR_INVALID_ARG, "stream is NULL");
    hipStream_t s = nullptr;
    ATR_HIP(hipStreamCreate(&s));
    *stream = s;
    return ATR_OK;
}
atr_status atr_stream_destroy(void* stream)
{
    ATR_HIP(hipStreamDestroy((hipStream_t)stream));
    return ATR_OK;
}
atr_status atr_stream_synchronize(void* stream)
{
    ATR_HIP(hipStreamSynchronize((hipStream_t)stream));
    return ATR_OK;
}
atr_status atr_mem_info(size_t* free_bytes, size_t* total_bytes)
{
    size_t f = 0, t = 0;
    ATR_HIP(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return ATR_OK;
}
atr_status atr_event_create(void** ev)
{
    if (!ev) return fail(ATR_ERR_INVALID_ARG, "ev is NULL");
    hipEvent_t e = nullptr;
    ATR_HIP(hipEventCreate(&e));
    *ev = e;
    return ATR_OK;
}
atr_status atr_event_destroy(void* ev)
{
    ATR_HIP(hipEventDestroy((hipEvent_t)ev));
    return ATR_OK;
}
atr_status atr_event_record(void* ev, void* stream)
{
    ATR_HIP(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream));
    return ATR_OK;
}
atr_status atr_event_elapsed_ms(void* a, void* b, float* ms)
{
    if (!ms) return fail(ATR_ERR_INVALID_ARG, "ms is NULL");
    ATR_HIP(hipEventSynchronize((hipEvent_t)b));
    ATR_HIP(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
    return ATR_OK;
}

}  // extern "C"
