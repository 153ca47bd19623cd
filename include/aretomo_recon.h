/*
 * aretomo_recon.h -- C ABI of the MI355X (gfx950) reconstruction library
 * libatr_recon.so: AreTomo3's tomographic reconstruction hot path
 * (weighting + ramp/Hamming filter, weighted back-projection, SART) plus the
 * stack-correction pre-stage that feeds it.
 *
 * Reference interfaces replaced (paths relative to the AreTomo3 tree):
 *   atr_plan_create / atr_execute with ATR_LAYOUT_SLICE and proj_y == 1
 *       -> CTomoWbp::Setup/DoIt   AreTomo/Recon/CReconInc.h:173-192
 *       -> CTomoSart::Setup/DoIt  AreTomo/Recon/CReconInc.h:194-236
 *   atr_plan_create / atr_execute(_host) on a whole series
 *       -> CDoWbpRecon::DoIt      AreTomo/Recon/CReconInc.h:259-279,
 *                                 AreTomo/Recon/CDoWbpRecon.cpp:28-47
 *       -> CDoSartRecon::DoIt     AreTomo/Recon/CReconInc.h:281-310,
 *                                 AreTomo/Recon/CDoSartRecon.cpp:30-57
 *   tilt_axis_deg / shift_xy in the descriptor (pre-stage inside the plan)
 *       -> CCorrTomoStack::Set1..Set4 + DoIt
 *                                 AreTomo/Correct/CCorrTomoStack.cpp:75-222
 *   ATR_LAYOUT_XYZ / ATR_LAYOUT_XYZ_NOFLIP
 *       -> CTiltSeries::FlipVol   DataUtil/CTiltSeries.cpp:286-305
 *
 * Conventions
 *   - All entry points return atr_status and never abort; the message of the
 *     last failure on the calling thread is atr_last_error_string().
 *   - A plan owns its device scratch and geometry tables and holds no global
 *     device state, so many plans may run concurrently on one device on
 *     different streams.  One plan must not be executed concurrently with
 *     itself.
 *   - Device pointers are HIP device pointers; `stream` is a hipStream_t
 *     passed as void* (NULL = the default stream).  Work is stream ordered.
 */
#ifndef ARETOMO_RECON_H
#define ARETOMO_RECON_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum atr_status {
    ATR_OK = 0,
    ATR_ERR_INVALID_ARG = 1,
    ATR_ERR_OUT_OF_MEMORY = 2,
    ATR_ERR_HIP = 3,
    ATR_ERR_UNSUPPORTED_SIZE = 4
} atr_status;

typedef enum atr_method { ATR_METHOD_WBP = 0, ATR_METHOD_SART = 1 } atr_method;

/* Storage type of the weighted projections on the device.  Voxel sums are
 * always accumulated in FP32. */
typedef enum atr_storage {
    ATR_STORAGE_F32 = 0,
    ATR_STORAGE_F16 = 1,
    ATR_STORAGE_BF16 = 2
} atr_storage;

/* SART subset composition.  CONTIGUOUS = CSplitItems blocks
 * (AreTomo/Util/CSplitItems.cpp:25-48), the reference behaviour.
 * INTERLEAVED = subset s holds tilts start+s, start+s+n_subsets, ... */
typedef enum atr_subset_order {
    ATR_SUBSETS_CONTIGUOUS = 0,
    ATR_SUBSETS_INTERLEAVED = 1
} atr_subset_order;

/* Output volume layout.
 *   XZY          frame y = slice y, rows z flipped (row VolZ-1-z): the
 *                CTiltSeries returned by CDo*Recon (CDoWbpRecon.cpp:107-123)
 *   XYZ          -FlipVol 1: frame z (physical), rows y
 *   XYZ_NOFLIP   -FlipVol 2: frame VolZ-1-z, rows y
 *   SLICE        frame y, rows z unflipped: the gfVolXZ of CTomo*::DoIt */
typedef enum atr_layout {
    ATR_LAYOUT_XZY = 0,
    ATR_LAYOUT_XYZ = 1,
    ATR_LAYOUT_XYZ_NOFLIP = 2,
    ATR_LAYOUT_SLICE = 3
} atr_layout;

typedef struct atr_series_desc {
    /* input projections: n_tilts images of proj_x * proj_y floats */
    int proj_x;
    int proj_y;
    int n_tilts;
    /* element strides of the input (0 = dense [tilt][y][x]); the padded
     * sinogram of CTomo*::DoIt is proj_y = 1, tilt_stride = (proj_x/2+1)*2 */
    size_t tilt_stride;
    size_t row_stride;
    const float* tilt_deg;      /* [n_tilts]   CAlignParam::GetTilts */
    /* pre-stage (CCorrTomoStack): NULL tilt_axis_deg = input already aligned */
    const float* tilt_axis_deg; /* [n_tilts]   CAlignParam::GetTiltAxis */
    const float* shift_xy;      /* [2*n_tilts] CAlignParam::GetShift, NULL = 0 */
    int random_fill;            /* fill empty pixels (forced on, as -AtBin FFT crop) */
    const unsigned char* exclude; /* [n_tilts] gbNoProjs (CTomoBase.cpp:48-51), NULL = none */
    int vol_z;
    int method;                 /* atr_method */
    /* SART (CTomoSart::Setup, CDoSartRecon::DoIt) */
    int sart_iters;
    int sart_subsets;
    int sart_start_tilt;
    int sart_n_tilts;           /* <= 0: n_tilts - sart_start_tilt */
    float sart_relax;           /* <= 0: 1/sart_subsets (CTomoSart.cpp:71) */
    int subset_order;           /* atr_subset_order */
    int positivity;             /* 1 = clamp at 0 as GBackProj.cu:56 (reference) */
    int storage;                /* atr_storage */
    int out_layout;             /* atr_layout */
    uint64_t fill_seed;         /* seeds the empty-fill index tables */
    /* local alignment (CLocalAlignParam, Correct/GCorrPatchShift.cu:18-48):
     * per tilt n_patches centres x, then centres y, shifts x, shifts y and
     * good flags, i.e. [n_tilts][5][n_patches]; 0 / NULL = global only.
     * Requires the pre-stage (tilt_axis_deg != NULL). */
    int n_patches;
    const float* local_params;
    /* -AtBin (AreTomo/CAreTomoMain.cpp:737-753, Correct/CBinStack.cpp:54-106):
     * Fourier binning of the aligned images to ((int)(size/bin))/2*2 before
     * the reconstruction; 1 = none.  The random fill is then forced on
     * (Correct/CCorrTomoStack.cpp:107).  vol_z is the binned thickness. */
    float bin;
} atr_series_desc;

typedef struct atr_plan atr_plan;

/* Fills *desc with the reference defaults (SART 20 iterations, positivity
 * on, contiguous subsets, relax 1/subsets, FP32, XZY, random fill on). */
void atr_series_desc_init(atr_series_desc* desc);

/* Validates the descriptor, computes the FP64 geometry tables and allocates
 * the device scratch on the current device. */
atr_status atr_plan_create(const atr_series_desc* desc, atr_plan** plan);

/* Size queries: the output volume holds vol_x * vol_y * vol_z floats. */
atr_status atr_plan_get_volume_size(const atr_plan* plan, int* vol_x, int* vol_y,
                                    int* vol_z);
/* Device bytes owned by the plan (tables + scratch). */
size_t atr_plan_device_bytes(const atr_plan* plan);
/* Device bytes a plan for *desc would allocate (serving admission). */
atr_status atr_estimate_device_bytes(const atr_series_desc* desc, size_t* bytes);
/* Aligned image size of the pre-stage for a raw image and the tilt axis of
 * frame 0 (CCorrectUtil::CalcAlignedSize, Correct/CCorrectUtil.cpp:11-21). */
atr_status atr_aligned_size(int raw_x, int raw_y, float tilt_axis_deg, int* aln_x, int* aln_y);

/* Runs the reconstruction on `stream`.  proj_dev: device input laid out as
 * described in the plan's descriptor; vol_dev: device output in out_layout.
 * accumulate != 0 (SART only): the SART starts from the volume already in
 * vol_dev (ATR_LAYOUT_SLICE), as CTomoSart::DoIt does with gfVolXZ;
 * otherwise from zeros, as CDoSartRecon does (CDoSartRecon.cpp:271-279). */
atr_status atr_execute(atr_plan* plan, const float* proj_dev, float* vol_dev,
                       int accumulate, void* stream);

/* Host-memory variant (drop-in for CDoWbpRecon/CDoSartRecon::DoIt): copies
 * proj_host in, reconstructs, copies the volume out, synchronises. */
atr_status atr_execute_host(atr_plan* plan, const float* proj_host, float* vol_host);

/* Per-kernel timing of the last executions (hipEvents on the plan stream);
 * enabled with atr_plan_set_profiling(plan, 1).  kernel: 0 = SART BP,
 * 1 = FP+diff, 2 = weighting/filter, 3 = WBP BP (2x grid + row FFT),
 * 4 = pre-stage, 5 = output/fill, 6 = WBP z-decimation, 7 = WBP row C2R. */
/* Autotuner (no reference counterpart; the reference's tile shapes are fixed
 * launch constants, Recon/GBackProj.cu:79-82, Recon/GForProj.cu:106-110):
 * times the candidate tile shapes of a SART plan on its own geometry -- back-
 * projection z rows per thread {16, 8} x tilts per LDS stage {16, 8}, then
 * forward-projection band rows, for the direct-staging forward projector,
 * (VolX a multiple of 256) the band-prefetching one (fp_prefetch = 1: one
 * wave streams the next z-band into LDS while the others sample) and (VolX
 * 1024, 2048) the register-prefetching one (fp_prefetch = 2: every
 * thread loads its columns of the next z-band into registers while the band
 * is sampled) -- and keeps the fastest.  Choices are cached per shape within the process.  Every voxel
 * sums its tilts in the same order whatever the shape; the FP32 rounding of
 * the back-projection coordinate (taken relative to the tile's window) and
 * the order in which a forward ray adds up its z-bands depend on the shape,
 * at the 1e-7 level (the two forward projectors give the same result for
 * the same band rows).  Outputs may be NULL.  A candidate whose launch fails
 * is dropped, never cached.  A plan that belongs to a batch (atr_batch_create
 * until atr_batch_destroy) cannot be retuned: ATR_ERR_INVALID_ARG. */
atr_status atr_plan_autotune(atr_plan* plan, void* stream, int* bp_tz, int* bp_chunk,
                             int* fp_band_rows, int* fp_prefetch);
/* Sets the tile shapes by hand: bp_tz, bp_chunk in {8, 16}; fp_prefetch 0, 1
 * (VolX a multiple of 256) or 2 (VolX 1024 or 2048); fp_band_rows in
 * [1, the largest band that fits in LDS for that forward projector].  Refused for a plan that
 * belongs to a batch, as atr_plan_autotune. */
atr_status atr_plan_set_tile_shape(atr_plan* plan, int bp_tz, int bp_chunk, int fp_band_rows,
                                   int fp_prefetch);
/* The forward-projection band heights (cell rows) the autotuner times for a
 * slice of vol_x x vol_z, without a GPU: every one fits the LDS, none
 * appears twice.  *n receives their number (at most 3); rows may be NULL. */
atr_status atr_fp_band_candidates(int vol_x, int vol_z, int* rows, int cap, int* n);
atr_status atr_plan_set_profiling(atr_plan* plan, int enable);
atr_status atr_plan_kernel_stats(const atr_plan* plan, int kernel, int* launches,
                                 double* total_ms);
atr_status atr_plan_reset_stats(atr_plan* plan);

void atr_plan_destroy(atr_plan* plan);

const char* atr_last_error_string(void);
const char* atr_version(void);

/* ---- synthetic tilt series (seeded ellipsoid phantoms) -----------------
 * Analytic parallel-beam line integrals of n_ellipsoids random ellipsoids
 * (centres uniform in the volume, semi-axes uniform [axis_min, axis_max],
 * uniform random orientation, density uniform [0.5, 1], additive), tilted
 * about y, optionally mapped to raw images with a tilt axis and per-tilt
 * shifts (the inverse of the pre-stage), plus Gaussian noise of
 * sigma = noise_frac * mean.  Output [n_tilts][proj_y][proj_x] host floats.
 * shift_xy_out (may be NULL) receives the per-tilt shifts drawn uniformly in
 * [-max_shift, max_shift]. */
atr_status atr_phantom_tilt_series(uint64_t seed, int proj_x, int proj_y, int n_tilts,
                                   const float* tilt_deg, int vol_z, int n_ellipsoids,
                                   float axis_min, float axis_max, float tilt_axis_deg,
                                   float max_shift, float noise_frac, float* shift_xy_out,
                                   float* proj_out, int n_threads);

/* ---- batches: many SART series per launch ----------------------------------
 * No single reference counterpart: the reference reconstructs one series at a
 * time (CDoSartRecon::DoIt, AreTomo/Recon/CDoSartRecon.cpp:30-57, called per
 * series from AreTomo/CAreTomoMain.cpp:809-848).  A batch runs every
 * back-projection and forward-projection step of all its series as a few
 * persistent work-queue launches (one per forward-projection staging variant
 * and per back-projection tile shape among the series, plus the empty-voxel
 * forward projector); with ATR_BATCH_GRAPH the iteration sequence
 * is captured once into a HIP graph and replayed.  Each series gives the
 * same volume as atr_execute on its plan alone.  Plans must be SART plans
 * with one storage type, must outlive the batch, and must not be executed
 * while the batch runs. */
typedef struct atr_batch atr_batch;
#define ATR_BATCH_GRAPH 1
atr_status atr_batch_create(atr_plan* const* plans, int n_plans, int flags, atr_batch** batch);
/* proj_dev[i] / vol_dev[i]: input and output of plans[i], as atr_execute. */
atr_status atr_batch_execute(atr_batch* batch, const float* const* proj_dev,
                             float* const* vol_dev, void* stream);
atr_status atr_batch_set_profiling(atr_batch* batch, int enable);
atr_status atr_batch_kernel_stats(const atr_batch* batch, int kernel, int* launches,
                                  double* total_ms);
atr_status atr_batch_reset_stats(atr_batch* batch);
void atr_batch_destroy(atr_batch* batch);

/* ---- serving: a device-resident request loop ------------------------------
 * No single reference counterpart: the reference processes the series of a
 * session one after another per GPU (AreTomo/CAreTomoMain.cpp:809-848, queue
 * in CProcessThread.cpp).  A server admits SART requests (optionally with a WBP
 * reconstruction of the same series) in arrival order while their device
 * bytes fit the HBM budget, keeps each series' data on the GPU from
 * pre-stage to output, and advances all admitted series together: every
 * round is the batched launches of one SART step (as in a batch) carrying
 * one SART step of each series whose turn it is (ATR_SCHEDULE_*).  Requests
 * admitted mid-run join at the next round.  Plans are built on a host thread
 * of the server ahead of admission; input uploads, WBP reconstructions and
 * copy-out run on their own streams beside the rounds; device memory comes
 * from an arena held until atr_server_destroy.  Each result equals what the
 * series gives alone. */
typedef struct atr_server atr_server;
/* Which admitted series take a SART step in a round: shortest-first
 * (default) steps the series with the least remaining work (and those within
 * twice it), any other one at least every 17th round; fair share gives every
 * series the same GPU work per round (deficit round robin on the voxel-tilt
 * updates of a step: a series whose steps cost m times those of the cheapest
 * one steps every m-th round); every-step advances each series by one step
 * every round. */
#define ATR_SCHEDULE_FAIR_SHARE 0
#define ATR_SCHEDULE_EVERY_STEP 1
#define ATR_SCHEDULE_SHORTEST_FIRST 2
typedef struct atr_server_config {
    size_t hbm_budget_bytes; /* 0 = 90% of the free device memory at creation */
    int max_active;          /* series reconstructed together at most (default 8), 0 = no limit */
    int storage;             /* projection storage of every request (default F16) */
    int schedule;            /* ATR_SCHEDULE_* (default shortest-first) */
} atr_server_config;
/* Times on the host clock, ms since the server started.  Per request:
 * queue = admit - submit (waiting for budget or a slot), GPU = compute_done -
 * admit (input upload, WBP and every SART round; the WBP volume is copied out
 * behind the WBP kernels, beside the rounds), copy-out = done - compute_done
 * (the SART volume to host memory). */
typedef struct atr_request_info {
    double submit_ms, admit_ms, done_ms;
    double latency_ms;                   /* done - submit */
    int status;
    double prepared_ms;                  /* host tables of its plans built */
    double compute_done_ms;              /* its last SART kernel finished */
} atr_request_info;
void atr_server_config_init(atr_server_config* cfg);
atr_status atr_server_create(const atr_server_config* cfg, atr_server** server);
/* desc: a SART series (its storage is replaced by the server's); the
 * descriptor and its arrays are copied.  proj_host, vol_host (and wbp_host
 * when with_wbp) must stay valid until atr_server_wait returns for the id;
 * pinned memory from atr_host_alloc lets the copies overlap the work. */
atr_status atr_server_submit(atr_server* server, const atr_series_desc* desc, int with_wbp,
                             const float* proj_host, float* vol_host, float* wbp_host,
                             uint64_t* id);
/* Blocks until the request is done; info may be NULL. */
atr_status atr_server_wait(atr_server* server, uint64_t id, atr_request_info* info);
atr_status atr_server_stats(atr_server* server, uint64_t* rounds, uint64_t* max_batch);
/* Finishes every submitted request, then frees the server. */
void atr_server_destroy(atr_server* server);
atr_status atr_host_alloc(void** ptr, size_t bytes);
atr_status atr_host_free(void* ptr);

/* ---- device helpers (plumbing for hosts without a HIP binding) --------- */
atr_status atr_device_count(int* count);
atr_status atr_set_device(int device);
atr_status atr_malloc(void** ptr, size_t bytes);
atr_status atr_free(void* ptr);
atr_status atr_memcpy_h2d(void* dst, const void* src, size_t bytes);
atr_status atr_memcpy_d2h(void* dst, const void* src, size_t bytes);
atr_status atr_memset(void* ptr, int value, size_t bytes);
/* Plain streaming copy kernel (16-byte aligned), the reference workload for
 * calibrating HBM byte counters. */
atr_status atr_copy_kernel(void* dst, const void* src, size_t bytes, void* stream);
atr_status atr_synchronize(void);
atr_status atr_stream_create(void** stream);
atr_status atr_stream_destroy(void* stream);
atr_status atr_stream_synchronize(void* stream);
atr_status atr_mem_info(size_t* free_bytes, size_t* total_bytes);
/* Wall time of the stream between two markers, milliseconds. */
atr_status atr_event_create(void** ev);
atr_status atr_event_destroy(void* ev);
atr_status atr_event_record(void* ev, void* stream);
atr_status atr_event_elapsed_ms(void* start, void* stop, float* ms);

/* ---- ProjAlign re-projector (AreTomo/ProjAlign/CCalcReproj.cpp,
 * GReproj.cu) --------------------------------------------------------------
 * Predicts projection proj_idx from its lower-tilt neighbours: tilts within
 * 20.5 deg whose cos stretch cos t / cos a is at most 1.2 (skip[i] != 0
 * excludes tilt i; at most 10; else the neighbour toward 0 deg) are back-
 * projected, nearest sample weighted by cos t * cos(a - t), onto a
 * 2 proj_x x vol_z slice per image row and forward-projected at angle a
 * (-1e30 where a ray meets fewer than 80 % non-empty voxels). */
typedef struct atr_reproj atr_reproj;
/* CCalcReproj::Setup(piTomoStkSize, iVolZ) (CCalcReproj.cpp:36-45). */
atr_status atr_reproj_create(int proj_x, int proj_y, int n_tilts, int vol_z, atr_reproj** reproj);
/* CCalcReproj::DoIt(ppfProjs, pfTiltAngles, pbSkipProjs, iProjIdx, pfReproj)
 * (CCalcReproj.cpp:47-65): proj_dev [n_tilts][proj_y][proj_x] and
 * reproj_dev [proj_y][proj_x] on the device; tilts / skip (may be NULL) on
 * the host.  Stream-ordered. */
atr_status atr_reproj_execute(atr_reproj* reproj, const float* proj_dev, const float* tilts,
                              const unsigned char* skip, int proj_idx, float* reproj_dev,
                              void* stream);
/* No counterpart in the reference: atr_reproj_execute without a handle, for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  It allocates
 * nothing: the back-projected slab lives in a fixed 64 MiB device array of
 * the library (one per GPU, held until the process ends), so one slab row,
 * 2 proj_x x vol_z floats, must fit 64 MiB (2 proj_x vol_z <= 16777216),
 * else ATR_ERR_INVALID_ARG.  Same kernels and results as
 * atr_reproj_execute.  Every call on one GPU uses the same array: calls that
 * overlap on one GPU (other streams or threads) race. */
atr_status atr_reproj_execute_noalloc(int proj_x, int proj_y, int n_tilts, int vol_z,
                                      const float* proj_dev, const float* tilts,
                                      const unsigned char* skip, int proj_idx, float* reproj_dev,
                                      void* stream);
/* The tilt range of CCalcReproj::mFindProjRange (CCalcReproj.cpp:73-124). */
atr_status atr_reproj_find_range(const float* tilts, const unsigned char* skip, int n_tilts,
                                 int proj_idx, int* first, int* last);
void atr_reproj_destroy(atr_reproj* reproj);

/* ---- Sample thickness (Recon/CCalcVolThick.cpp) ---------------------------
 * The other in-repo caller of the reconstruction: a WBP of the series
 * Fourier-binned by max(20 / pixel_size, 20), flipped to xyz, the standard
 * deviation of the central 40 % tile of each z section, an 11-point moving
 * average and the edge search of CCalcVolThick::mDetectEdges.  Edges are in
 * aligned (unbinned) pixels; thickness is CAlignParam::m_iThickness and
 * offset_z CAlignParam::m_iOffsetZ. */
typedef struct atr_thickness {
    int low_edge, high_edge;  /* m_aiSampleEdges */
    int thickness;            /* high_edge - low_edge */
    int offset_z;             /* sample centre - volume centre */
    float binning;            /* m_fBinning */
    float pixel_size;         /* m_fPixSize: Angstrom per aligned pixel */
    int vol_z;                /* sections of the binned volume */
    int n_profile;            /* vol_z - 1 profile values */
} atr_thickness;
/* CCalcVolThick::DoIt (CCalcVolThick.cpp:47-90): desc describes the raw
 * series (size, strides, tilts, tilt axis, shifts, exclude, fill seed); its
 * method, bin, vol_z, layout and storage are replaced.  proj_host: host
 * input.  profile (may be NULL) receives the unsmoothed std profile,
 * out->n_profile values. */
atr_status atr_calc_vol_thick(const atr_series_desc* desc, float pixel_size, const float* proj_host,
                              atr_thickness* out, float* profile);
/* mSmooth + mDetectEdges (CCalcVolThick.cpp:93-110,185-255) on a std profile
 * of n >= 10 sections; smoothed (may be NULL) receives the smoothed profile. */
atr_status atr_thickness_edges(const float* std_profile, int n, float binning, atr_thickness* out,
                               float* smoothed);

#ifdef __cplusplus
}
#endif
#endif
