// atr_internal.h -- plan structure, geometry tables and kernel launchers of
// libatr_recon.so (MI355X / gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <functional>
#include <vector>

#include "../../include/aretomo_recon.h"
#include "fft_lds.h"

namespace atr {

// Sentinels of the reference (GBackProj.cu:45,55-59; GForProj.cu:161,181).
constexpr float kEmpty = -1e10f;
constexpr float kSentinel = -1e30f;

// ---- tile shapes ---------------------------------------------------------
// SART back-projection: a workgroup owns BP_TX x BP_TZ voxels of one slice;
// lane = x, each thread walks BP_TZ rows of z.
constexpr int BP_TX = 256;
constexpr int BP_TZ = 16;
constexpr int BP_CHUNK = 16;                 // tilts staged in LDS at once
constexpr int BP_WIN = BP_TX + BP_TZ + 16;   // staged pairs per tilt
constexpr int bp_win(int tz) { return BP_TX + tz + 16; }
// Forward projection: a workgroup of FP_THREADS owns one slice and a group of
// tilts; wave w takes the (tilt, 64-ray tile) pairs w, w+16, ... (FP_MAXP at
// most), so every wave sees every tilt; the slice is streamed through LDS in
// z-bands of interpolation pairs.
constexpr int FP_THREADS = 1024;
constexpr int FP_WAVES = FP_THREADS / 64;
constexpr int FP_WAVE = 64;
constexpr int FP_MAXP = 16;
constexpr int FP_LDS_BUDGET = 150 * 1024;
constexpr int FP_BZ_MAX = 17;                // cell rows per band at most
constexpr int FP_STAGE_ROWS = 9;              // rows a thread loads before forming pairs
constexpr int FP_RP_ISSUE = 8;                // register prefetch: pair slots before its loads
// Band-prefetching variant: the last wave streams the next band's volume rows
// into LDS with direct global->LDS loads while the other waves sample the
// current band; the raw rows sit after the pair rows.  Needs VolX % 256 == 0
// (a wave's load instruction covers 256 floats of one row).
// (the 160 KiB of a CU less 256 bytes for the static LDS of the batched
// work-queue kernel, which runs these bands too)
constexpr int FP_LDS_PF = 160 * 1024 - 256;
// Register-prefetching variant (fp_pf = 2): every thread holds the next
// band's rows of its columns in registers while the current band is
// sampled; VolX 1024 (one column per thread, up to 18 rows) and 2048 (two
// columns, up to 9 rows).
constexpr int FP_PF_DMA = 1, FP_PF_REG = 2;
inline int fp_waves(int pf) { return pf == FP_PF_DMA ? FP_WAVES - 1 : FP_WAVES; }
inline bool fp_pf_supported(int vol_x) { return vol_x >= 256 && vol_x % 256 == 0; }
inline bool fp_rp_supported(int vol_x) { return vol_x == 1024 || vol_x == 2048; }
inline bool fp_mode_supported(int pf, int vol_x)
{
    return pf == 0 || (pf == FP_PF_DMA && fp_pf_supported(vol_x)) ||
           (pf == FP_PF_REG && fp_rp_supported(vol_x));
}

enum Kernel {
    K_BP = 0,
    K_FP = 1,
    K_WEIGHT = 2,
    K_WBP_BP = 3,
    K_PRESTAGE = 4,
    K_OUTPUT = 5,
    K_WBP_COL = 6,
    K_WBP_ROW = 7,
    K_COUNT = 8
};

// Per-slice SART tilt list entry: tilt index and the row of the input
// sinogram/diff buffer that holds it.
struct TiltRow {
    int tilt;
    int row;
};

// FP64 geometry evaluated once per plan on the host.
struct Geometry {
    std::vector<float> cos_sin;        // [2N] float, as CTomoBase.cpp:55-65
    std::vector<int2> bp_lohi;         // [N][VolZ] x-interval of GBackProj inclusion
    // [N][ztiles][xtiles] {window start, class 0 empty / 1 full / 2 partial,
    // FP32 bits of u - start at the tile's first voxel, 0}
    std::vector<int4> bp_tiles;
    std::vector<int2> wbp_lohi;        // [N][2VolZ] x2-interval of GBackProjWbp inclusion
    std::vector<int2> fp_range;        // [N][P] first valid sample and count of each ray
    std::vector<int> fp_len;           // [N] ray length L
    std::vector<int> fp_geo_off;       // [N] offset into fp_geo
    std::vector<float4> fp_geo;        // [sum L] {row byte offset bits, wz, q, 1 - wz}
    std::vector<int2> fp_band;         // [N][nbands] sample range whose rows lie in a z-band
    std::vector<int4> fp_desc;         // [N][P/64][nbands] first geometry entry, count, first k, full
    int fp_bz = 0, fp_nbands = 0;      // z rows per band
};

struct BpArgs;
struct FpArgs;

// Device memory source of a plan: hipMalloc / hipFree when a plan has none,
// the server's arena (which never returns memory to HIP while it serves, so
// releasing a plan synchronises nothing) when it has one.
struct DevAlloc {
    virtual ~DevAlloc() = default;
    virtual hipError_t alloc(void** ptr, size_t bytes) = 0;
    virtual void release(void* ptr) = 0;
};

// Host copies of the tables a plan uploads, built before any device work
// (plan_prepare) and released once uploaded (plan_upload).
struct HostTables {
    std::vector<float2> cs;
    std::vector<unsigned char> excl;
    std::vector<double> axis_cs;
    std::vector<float> shifts, local, filter;
    std::vector<int> fill_aln, fill_vol, sched;
    std::vector<float2> tw_bin[4], tw_proj, tw_x2, tw_z2, tw_z, tw_x;
    std::vector<TiltRow> lists;
    std::vector<int4> wbp_tiles;
    std::vector<BpArgs> bp_args;   // kernel arguments of every step
    std::vector<FpArgs> fp_args;
};

struct Plan {
    DevAlloc* alloc = nullptr;   // not owned; NULL = hipMalloc / hipFree
    // table uploads ordered on this stream (hipMemcpyAsync; the host tables
    // must then live until the stream has passed them), else synchronous
    bool upload_stream_set = false;
    hipStream_t upload_stream = nullptr;
    HostTables host;
    atr_series_desc desc{};
    // sizes
    int n = 0;            // tilts
    int in_x = 0, in_y = 0;   // input image size
    size_t in_tilt_stride = 0, in_row_stride = 0;
    int aln_x = 0, aln_y = 0; // aligned image (= input when no pre-stage)
    bool binned = false;      // Fourier binning of the aligned images (-AtBin > 1)
    int bin_x = 0, bin_y = 0; // binned image size
    int n_patches = 0;        // local alignment patches per tilt (0 = global only)
    int proj = 0;         // effective projection width (ProjX/2*2)
    int vol_x = 0, vol_y = 0, vol_z = 0;
    int sino_stride = 0;  // elements per sinogram row on the device
    bool prestage = false;
    bool fill = false;
    int method = 0, storage = 0, layout = 0;
    // SART
    int iters = 0, n_subsets = 1, sart_start = 0, sart_n = 0, order = 0, positivity = 1;
    float relax = 1.0f;
    int max_subset = 0;
    std::vector<int> subset_off;      // [n_subsets + 1] offsets into subset_tilts
    std::vector<TiltRow> subset_tilts;// BP lists (excluded tilts removed)
    std::vector<int> subset_fp;       // FP tilt lists (all tilts of the subset)
    std::vector<int> subset_fp_off;
    std::vector<int> subset_pair_off; // [n_subsets] offsets into d_fp_pairs
    std::vector<TiltRow> all_tilts;   // initial BP list
    // geometry
    Geometry geo;
    // device tables
    float2* d_cs = nullptr;
    int2* d_bp_lohi = nullptr;
    int4* d_bp_tiles = nullptr;
    int2* d_wbp_lohi = nullptr;
    int2* d_fp_range = nullptr;
    int* d_fp_geo_off = nullptr;
    int* d_fp_len = nullptr;
    float4* d_fp_geo = nullptr;
    int4* d_fp_desc = nullptr;
    int* d_fp_pairs = nullptr;        // per subset: FP wave schedules (fp_pair_schedule)
    TiltRow* d_lists = nullptr;       // all_tilts then every subset's BP list
    int* d_fp_lists = nullptr;        // every subset's FP list
    int* d_fill_aln = nullptr;        // [aln_x*aln_y]
    int* d_fill_vol = nullptr;        // [vol_x*vol_z]
    float* d_filter = nullptr;        // [proj/2+1] ramp/Hamming weights
    float2* d_tw_proj = nullptr;      // twiddles for the row filter
    float2* d_tw_x2 = nullptr;        // 2 VolX
    float2* d_tw_z2 = nullptr;        // 2 VolZ
    float2* d_tw_z = nullptr;         // VolZ
    float2* d_tw_x = nullptr;         // VolX
    // device work buffers
    double* d_axis_cs = nullptr;      // [N][2] cos/sin of the tilt axis (FP64)
    float* d_shifts = nullptr;        // [N][2] global shifts
    unsigned char* d_excl = nullptr;  // [N] excluded tilts
    float* d_aligned = nullptr;       // [N][aln_y][aln_x] pre-stage output with sentinels
    float* d_local = nullptr;         // [N][5][n_patches] patch centres, shifts, good flags
    float* d_binned = nullptr;        // [N][bin_y][bin_x] Fourier-binned images
    float2* d_bin_t = nullptr;        // [N][aln_y][bin_x/2+1] row spectra
    float2* d_bin_u = nullptr;        // [N][bin_y][bin_x/2+1] cropped, column-inverted
    float2* d_tw_bin[4] = {nullptr, nullptr, nullptr, nullptr};
    FftDesc fft_bin_ax{}, fft_bin_ay{}, fft_bin_by{}, fft_bin_bx{};
    double* d_row_stats = nullptr;    // [N][aln_y][2] per-row (sum, count) of non-empty pixels
    double* d_tilt_stats = nullptr;   // [N][2] (sum, count) of non-empty aligned pixels
    void* d_sino = nullptr;           // [vol_y][N][sino_stride] weighted projections
    void* d_diff = nullptr;           // [vol_y][max_subset][sino_stride]
    float* d_vol = nullptr;           // [vol_y][vol_z][vol_x] SART state
    int* d_sflag = nullptr;           // [vol_y] sinogram slice holds sentinels
    int* d_vflag = nullptr;           // [vol_y] volume slice holds sentinels
    double* d_vstats = nullptr;       // [vol_y][2] fill statistics of volume slices
    int bp_tz = BP_TZ, bp_chunk = BP_CHUNK;   // back-projection tile shape (autotuned)
    int fp_pf = 0;                    // forward projector: 0 direct staging, FP_PF_DMA, FP_PF_REG (autotuned)
    // batches holding this plan: their item arrays and captured graphs point
    // at its tables, so the tables must not be rebuilt while this is > 0
    int batch_refs = 0;
    size_t tab_bytes[9] = {0};        // sizes of the tile-shape dependent tables
    BpArgs* d_bp_args = nullptr;      // SART step arguments: initial BP, then per subset
    FpArgs* d_fp_args = nullptr;      // per subset
    float* d_wbp_grid = nullptr;      // [slab][2VolZ][2VolX] the 2x-grid back-projection
    int4* d_wbp_tiles = nullptr;      // [N][2VolZ/16][2VolX/256] as bp_tiles
    float2* d_wbp_t = nullptr;        // [slab][2VolZ][VolX/2+1]
    float2* d_wbp_u = nullptr;        // [slab][VolZ][VolX/2+1]
    int wbp_slab = 0;
    size_t device_bytes = 0;
    // FFT descriptors
    FftDesc fft_proj{}, fft_x2{}, fft_z2{}, fft_z{}, fft_x{};
    // profiling
    bool profiling = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
    int launches[K_COUNT] = {0};
    double total_ms[K_COUNT] = {0};
    hipStream_t stream = nullptr;
};

// ---- kernel arguments ----------------------------------------------------
// SART back-projection of one list of tilts into one series' volume.
struct BpArgs {
    const void* sino;          // [VolY][rows_per_slice][stride] storage type
    int rows_per_slice, stride, proj;
    const TiltRow* list;
    int n_list;
    const float2* cs;
    const int2* lohi;          // [N][VolZ]
    const int4* tiles;         // [N][ztiles][xtiles] {window start, class, u0 bits, 0}
    int n_ztiles, n_xtiles;
    float relax;
    int positivity;
    float scale, offset;       // u = scale (fx cos + fz sin) + offset (offset enters via the tile table)
    const int* sflag;          // NULL: the sinogram holds no empty samples
    float* vol;
    int vol_x, vol_z;
    int* vflag;
};
// Forward projection + difference of one list of tilts of one series.
struct FpArgs {
    const float* vol;
    int vol_x, vol_z;
    const int* vflag;
    const int* fp_list;
    int n_list, group;
    const int* sched;          // fp_pair_schedule blocks, one per group
    const float2* cs;
    const int2* fp_range;
    const float4* geo;
    const int4* fp_desc;
    int nb, bz, proj;
    const void* sino;
    int n, stride;
    void* diff;
    int diff_rows;
    int pf;                    // band-prefetching variant (fp_waves(pf) sampling waves)
};
// Entries of the batched (work-queue) launches; item0 = first queue index.
// shape: the series' back-projection tile shape, bp_shape(bp_tz, bp_chunk).
struct BpItem {
    BpArgs a;
    int item0, n_items_y;      // items of one slice: BP_ITEM_TILES tiles each
    int shape;
};

// The entries of one persistent work-queue launch, laid out largest work
// item first: the items the launch claims last are then the cheapest ones,
// so it ends with a short tail instead of the big slices of whichever series
// came last.  Items are independent, so the order changes no result.  Sets
// item0, copies the entries to out, returns the number of queue items.
template <class Item>
struct QueueEntry {
    Item item;
    int n_items;       // queue items of the entry
    double item_cost;  // work of one of them (any unit common to the launch)
};
template <class Item>
int lay_out_entries(std::vector<QueueEntry<Item>>& v, Item* out)
{
    std::stable_sort(v.begin(), v.end(), [](const QueueEntry<Item>& a, const QueueEntry<Item>& b) {
        return a.item_cost > b.item_cost;
    });
    int total = 0;
    for (size_t i = 0; i < v.size(); i++) {
        v[i].item.item0 = total;
        total += v[i].n_items;
        out[i] = v[i].item;
    }
    return total;
}
// A batched back-projection work item is this many consecutive tiles of
// one slice: a tile is a few microseconds of work, and each item costs the
// queue's claim and two barriers.
constexpr int BP_ITEM_TILES = 4;
inline int bp_items_per_slice(const BpArgs& a)
{
    return (a.n_ztiles * a.n_xtiles + BP_ITEM_TILES - 1) / BP_ITEM_TILES;
}
// 0: 16 z rows x 16 tilts per stage, 1: 16 x 8, 2: 8 x 16, 3: 8 x 8
constexpr int BP_SHAPES = 4;
inline int bp_shape(int tz, int chunk) { return (tz == 8 ? 2 : 0) + (chunk == 8 ? 1 : 0); }
// Batched back-projection: one persistent launch per tile shape present.
struct BpLaunch {
    int first;       // first entry of its queue in the item array
    int n_entries, total, shape;
};
// Writes the queues to out (room for ent.size() entries), each laid out
// largest item first, and returns their launches in order.
std::vector<BpLaunch> lay_out_bp(const std::vector<QueueEntry<BpItem>>& ent, BpItem* out);

struct FpItem {
    FpArgs a;
    int item0, n_groups;       // items of one slice = tilt groups
};

// Batched forward projection: one persistent launch per staging variant
// present in a set of queue entries (each variant's kernel holds only its
// own code, as the per-series kernels do), then one launch of the
// empty-voxel variant over every entry, whose blocks leave at once on slices
// without empty voxels.  Variants 0-3: band prefetch by DMA, 4-7: direct
// staging (VolX 512, 1024, 2048, any), 8-9: register prefetch (VolX 1024,
// 2048).
constexpr int FP_VARIANT_EMPTY = 10;
constexpr int FP_VARIANTS = 11;
int fp_variant(const FpArgs& a);
struct FpLaunch {
    int first;       // first entry of its queue in the item array
    int n_entries, total, variant;
    size_t shm;      // dynamic LDS bytes
};
// Writes the queues to out (room for 2 ent.size() entries), each laid out
// largest item first, and returns their launches in order.
std::vector<FpLaunch> lay_out_fp(const std::vector<QueueEntry<FpItem>>& ent, FpItem* out);

// Re-projector launch arguments (passed by value): tilts [first, first +
// n_range) back-projected with host FP64 cos / sin / cos(a - t) tables, then
// a ray of `ray` samples at (ray_cos, ray_sin) per output pixel, empty below
// min_count non-empty samples.
constexpr int REPROJ_MAX_TILTS = 10;
struct ReprojArgs {
    int proj_x, proj_y, vol_x, vol_z;
    int first, n_range, ray;
    float min_count;
    double cos_t[REPROJ_MAX_TILTS], sin_t[REPROJ_MAX_TILTS], weight[REPROJ_MAX_TILTS];
    double ray_cos, ray_sin;
};
void launch_reproj(const ReprojArgs& a, const float* projs, float* vol_slab, int slab_rows,
                   float* out, hipStream_t s);

// What the weighting / filter kernels read: the binned images, the
// pre-stage output (with the random fill), or the caller's input.
struct StageSource {
    const float* src;
    size_t tilt_stride, row_stride;
    bool fill;
};
inline StageSource stage_source(const Plan& p, const float* in)
{
    if (p.binned) return {p.d_binned, (size_t)p.bin_x * p.bin_y, (size_t)p.bin_x, false};
    if (p.prestage) return {p.d_aligned, (size_t)p.aln_x * p.aln_y, (size_t)p.aln_x, p.fill};
    return {in, p.in_tilt_stride, p.in_row_stride, false};
}

// ---- plan construction in two halves (api.cpp) -----------------------------
// plan_prepare: validation, sizes and every host table (FP64 geometry, fill
// tables, schedules); no device work.  plan_upload: device allocations
// through p.alloc and the uploads; on failure what was allocated is freed.
// plan_release: frees the device memory, without synchronising.
atr_status plan_prepare(const atr_series_desc* d, Plan& p);
atr_status plan_upload(Plan& p);
void plan_release(Plan& p);
atr_status plan_execute(Plan& p, const float* in, float* out, int accumulate, hipStream_t s);

// ---- SART stages shared by single-plan and batched execution (api.cpp) ----
atr_status sart_prepare(Plan& p, const float* in, float* out, int accumulate, hipStream_t s);
void sart_finish(Plan& p, float* out, hipStream_t s);

// ---- error handling ------------------------------------------------------
void set_error(const std::string& msg);
atr_status fail(atr_status st, const std::string& msg);
atr_status hip_fail(hipError_t e, const char* what);
#define ATR_HIP(call)                                         \
    do {                                                      \
        hipError_t _e = (call);                               \
        if (_e != hipSuccess) return ::atr::hip_fail(_e, #call); \
    } while (0)

// ---- host geometry (geometry.cpp) ---------------------------------------
void cos_sin_table(const float* tilts, int n, std::vector<float>& cs);
void build_bp_intervals(const std::vector<float>& cs, int n, int vol_x, int vol_z, int proj,
                        std::vector<int2>& lohi);
void build_bp_tiles(const std::vector<float>& cs, const std::vector<int2>& lohi, int n, int vol_x,
                    int vol_z, double scale, double offset, int tz, std::vector<int4>& tiles);
void build_wbp_intervals(const std::vector<float>& cs, int n, int vol_x, int vol_z, int proj,
                         std::vector<int2>& lohi);
bool build_fp_tables(const std::vector<float>& cs, int n, int vol_x, int vol_z, int proj,
                     int bz, Geometry& g);
int fp_band_rows(int vol_x, int pf = 0);
int fp_balanced_rows(int bz_max, int vol_z);
void fp_band_candidates(int vol_x, int vol_z, std::vector<int>& rows, int pf = 0);
int fp_group(int proj, int n_list, int n_waves = FP_WAVES);
// fn(0) .. fn(n - 1) on up to 16 host threads (returns when all are done).
void parallel_for(int n, const std::function<void(int)>& fn);
void fp_pair_schedule(const Geometry& g, const int* list, int n_list, int proj, int group,
                      std::vector<int>& tab, int n_waves = FP_WAVES);
void split_items(int n_items, int n_splits, std::vector<int>& starts, std::vector<int>& sizes);
void fill_table(uint64_t seed, int n, std::vector<int>& table);
void aligned_size(int raw_x, int raw_y, float axis_deg, int* ax, int* ay);
bool make_fft(int n, FftDesc& fd);
std::vector<float2> twiddles(int n);
std::vector<float> ramp_filter(int n);
uint64_t splitmix64(uint64_t x);

// ---- kernel launchers (kernels_*.hip) -----------------------------------
void launch_prestage(const Plan& p, const float* in, hipStream_t s);
void launch_tilt_stats(const Plan& p, hipStream_t s);
void launch_bin(const Plan& p, const float* in, hipStream_t s);
// Volume dimensions a plan for *d would produce (validates the descriptor).
atr_status desc_volume(const atr_series_desc* d, int* vol_x, int* vol_y, int* vol_z);
int bin_col_block(int ay);
void launch_weight_sart(const Plan& p, const float* in, hipStream_t s);
void launch_filter_wbp(const Plan& p, const float* in, hipStream_t s);
void launch_zero(float* ptr, size_t n, hipStream_t s);
void launch_zero_int(int* ptr, int n, hipStream_t s);
int wbp_col_block(int vol_z);
BpArgs bp_args(const Plan& p, const void* sino, int rows, const TiltRow* list, int n_list,
               float relax, bool check_sent);
FpArgs fp_args(const Plan& p, const int* fp_list, int n_list, const int* sched);
size_t fp_lds_bytes(int bz, int vol_x, int pf = 0);
// step 0: initial back-projection, 1 + s: subset s (arguments in p.d_bp_args)
void launch_bp(const Plan& p, int step, hipStream_t s);
// items: the launch's queue (the array lay_out_bp wrote, at l.first)
void launch_bp_batch(int storage, const BpLaunch& l, const BpItem* items, int* counter, int grid,
                     hipStream_t s);
// persistent grid of a launch: CUs x resident blocks per CU
int bp_batch_grid(int storage, const BpLaunch& l);
// items: the launch's queue (the array lay_out_fp wrote, at l.first)
void launch_fp_batch(int storage, const FpLaunch& l, const FpItem* items, int* counter, int grid,
                     hipStream_t s);
// persistent grid of a launch: CUs x resident blocks per CU
int fp_batch_grid(int storage, const FpLaunch& l);
void launch_fp_diff(const Plan& p, int subset, hipStream_t s);
void launch_vol_stats(const Plan& p, hipStream_t s);
void launch_output(const Plan& p, float* out, hipStream_t s);
// slab k of the WBP: back-projection onto the 2x grid (arguments p.d_bp_args + k)
void launch_wbp_bp(const Plan& p, int slab, int ny, hipStream_t s);
void launch_wbp_rowfft(const Plan& p, int ny, hipStream_t s);
void launch_wbp_colfft(const Plan& p, int y0, int ny, hipStream_t s);
void launch_wbp_rowc2r(const Plan& p, int y0, int ny, float* out, hipStream_t s);
void launch_load_volume(const Plan& p, const float* in, hipStream_t s);
void launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s);

}  // namespace atr

struct atr_plan {
    atr::Plan p;
};
