/*
 * oracle.h -- FP64 CPU restatement of AreTomo3's tomographic reconstruction
 * hot path (AreTomo/Recon + the AreTomo/Correct pre-stage).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls
 * this code.  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg
 * of bench.py use it, and only as the checker / the reported CPU baseline.
 *
 * Every function names the reference file:line it restates.  Arithmetic is
 * carried out in double; the reference's own float constants are kept
 * (fRad = 0.017453f in Recon/CTomoBase.cpp:56, 6.2831852f in
 * Recon/GRWeight.cu:20, 3.141592654f/180 in Correct/GCorrPatchShift.cu:104),
 * so the geometry is the reference's geometry evaluated without FP32
 * rounding.  The library is built with -ffp-contract=off so that every
 * double expression is evaluated exactly as written.
 *
 * Layout conventions (same as the reference):
 *   sinogram  : n_tilts rows of pad_x doubles, row t = tilt t (pad_x >= proj_x)
 *   vol slice : vol_z rows of vol_x doubles (x fastest), z = 0 at the bottom
 *   stacks    : [tilt][y][x]
 */
#ifndef ATR_ORACLE_H
#define ATR_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- small helpers ---------------------------------------------------- */
uint64_t orc_splitmix64(uint64_t x);
void orc_fill_table(uint64_t seed, int n, int* table);
void orc_cos_sin(const float* tilts_deg, int n, float* cos_sin);
void orc_split_items(int n_items, int n_splits, int* starts, int* sizes);
int  orc_subset_tilts(int start_tilt, int n_tilts, int n_subsets, int order,
                      int subset, int* tilts);
void orc_fft(double* re, double* im, int n, int sign);
void orc_aligned_size(int raw_x, int raw_y, float tilt_axis_deg,
                      int* aln_x, int* aln_y);

/* ---- per-slice kernels ------------------------------------------------ */
void orc_weight_projs(double* sino, int n_tilts, int pad_x, int proj_x,
                      const float* cos_sin, int vol_z);
void orc_rweight(double* sino, int n_tilts, int pad_x, int fft_size);
void orc_back_proj(const double* sino, int pad_x, int proj_x,
                   const float* cos_sin, const unsigned char* no_proj,
                   const int* tilts, int n_list, int positivity, double relax,
                   double* vol, int vol_x, int vol_z);
void orc_for_proj(const double* vol, int vol_x, int vol_z,
                  const float* cos_sin, const int* tilts, int n_list,
                  int proj_x, int pad_x, double* fwd);
void orc_diff_proj(const double* raw, double* fwd, const int* tilts,
                   int n_list, int proj_x, int pad_x);
void orc_fill_empty(double* img, int w, int h, int pad_x, const int* table);
void orc_back_proj_wbp(const double* sino, int pad_x, int proj_x,
                       const float* cos_sin, const unsigned char* no_proj,
                       int n_tilts, double* img2, int vol_x2, int vol_z2);
void orc_wbp_crop(const double* img2, int vol_x, int vol_z, double* vol);
void orc_fourier_crop(const double* img, int ax, int ay, int bx, int by, double* out);

/* ---- per-slice operators (CTomoSart / CTomoWbp) ------------------------ */
void orc_tomo_sart(double* sino, int n_tilts, int pad_x, int proj_x,
                   const float* cos_sin, const unsigned char* no_proj,
                   double* vol, int vol_x, int vol_z, int start_tilt,
                   int n_sart_tilts, int iters, int n_subsets, double relax,
                   int order, int positivity, const int* vol_fill_table);
void orc_tomo_wbp(double* sino, int n_tilts, int pad_x, int proj_x,
                  const float* cos_sin, const unsigned char* no_proj,
                  double* vol, int vol_x, int vol_z);

/* ---- pre-stage and per-series drivers --------------------------------- */
void orc_corr_proj(const float* raw, int raw_x, int raw_y, float tilt_axis_deg,
                   float shift_x, float shift_y, int n_patches,
                   const float* local_params, double* out, int out_x,
                   int out_y, int out_pad_x, int random_fill,
                   const int* fill_table);
void orc_prestage(const float* raw, int n_tilts, int raw_x, int raw_y,
                  const float* tilt_axis_deg, const float* shifts_xy,
                  int n_patches, const float* local_params,
                  int random_fill, uint64_t fill_seed, double* aligned,
                  int aln_x, int aln_y);
void orc_bin_size(int ax, int ay, float bin, int* bx, int* by);
void orc_bin_stack(const double* in, int n, int ax, int ay, int bx, int by, double* out);
void orc_recon_slices(const double* aligned, int n_tilts, int proj_x,
                      int proj_y, const float* tilts_deg,
                      const unsigned char* no_proj, int method, int vol_z,
                      int start_tilt, int n_sart_tilts, int iters,
                      int n_subsets, double relax, int order, int positivity,
                      uint64_t fill_seed, const int* slices, int n_slices,
                      double* out);
void orc_reproj_range(const float* tilts, const unsigned char* skip, int n,
                      int proj_idx, int* range);
void orc_reproj(const float* projs, const float* tilts, int n_tilts, int proj_x,
                int proj_y, int vol_z, int proj_idx, const int* range, double* out);
void orc_thick_profile(const float* vol, int nx, int ny, int nz, float* out);
void orc_thick_smooth(float* ccs, int n);
void orc_thick_edges(const float* ccs, int n, float binning, int* out);
void orc_flip_vol(const double* vol_xzy, int vol_x, int vol_z, int vol_y,
                  int flip, double* vol_xyz);
int  orc_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
