/*
 * aretomo_files.h -- the on-disk formats at both ends of the reconstruction:
 * MRC stacks in, MRC volumes out, the AreTomo alignment file (.aln) and the
 * tilt-angle file (_TLT.txt / .rawtlt).  Host code only (no device work), in
 * the same library as aretomo_recon.h; every entry point returns atr_status
 * and leaves a message for atr_last_error_string() on failure.
 */
#ifndef ARETOMO_FILES_H
#define ARETOMO_FILES_H
#include <stddef.h>
#include <stdint.h>

#include "aretomo_recon.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- MRC ---------------------------------------------------------------- */
/* Header summary of an MRC file (LibSrc/Mrcfile/Header/CLoadMainHeader.cpp:18-38:
 * byte order detected from implausible sizes or a negative mode). */
typedef struct atr_mrc_info {
    int nx, ny, nz;     /* section width, height, number of sections */
    int mode;           /* 1 int16, 2 float32, 6 uint16 are readable */
    float pixel_size;   /* xlen / mx, 0 when mx is 0 */
    int nsymbt;         /* extended-header bytes after the 1024-byte header */
    int swapped;        /* 1 when the file is of the other byte order */
} atr_mrc_info;

/* Reads the main header.  Replaces Mrc::CLoadMrc::OpenFile +
 * CLoadMainHeader::GetSize/GetMode (DataUtil/CTsPackage.cpp:307-321). */
atr_status atr_mrc_read_info(const char* path, atr_mrc_info* info);

/* Reads sections [first, first + count) as float32 (modes 1, 2, 6 are
 * converted as DataUtil/CTsPackage.cpp:322-356 does; other modes give
 * ATR_ERR_UNSUPPORTED_SIZE).  out holds count * nx * ny floats. */
atr_status atr_mrc_read_sections(const char* path, int first, int count, float* out);

/* Writes a float32 MRC stack of nz sections of nx * ny, as
 * DataUtil/CAsyncSaveVol.cpp:94-117 saves a volume through
 * LibSrc/Mrcfile/App/CSaveMrc.cpp: 1024-byte MRC2014 header (mode 2,
 * m = n, lengths = m * pixel_size, angles 90, map 1 2 3, EXTTYPE "AGAR",
 * NVERSION 20140, "MAP" + little-endian stamp), no extended header (the
 * reference's SetImgSize resets the 32-float one it asked for), then the
 * sections.  Statistics fields stay 0 as in the reference unless
 * with_stats != 0 (min, max, mean of the data). */
atr_status atr_mrc_write(const char* path, const float* data, int nx, int ny, int nz,
                         float pixel_size, int with_stats);

/* Output-file suffix of volume n (DataUtil/CAsyncSaveVol.cpp:94-98):
 * 0 "_Vol.mrc", 1 "_EVN_Vol.mrc", 2 "_ODD_Vol.mrc", 3 "_2ND_Vol.mrc",
 * 4 "_3RD_Vol.mrc"; NULL otherwise. */
const char* atr_volume_suffix(int nth_vol);

/* ---- .aln alignment file ------------------------------------------------- */
/* Parsed .aln (AreTomo/MrcUtil/CLoadAlignFile.cpp:34-291, format written by
 * CSaveAlignFile.cpp:67-127).  Section indices are 1-based after loading:
 * a zero-based (older) file is converted as CLoadAlignFile::mToOneBased. */
typedef struct atr_aln atr_aln;
typedef struct atr_aln_info {
    int raw_size[3];     /* "# RawSize = x y z": the raw stack before dark removal */
    int n_tilts;         /* aligned tilts = raw_size[2] - n_dark */
    int n_patches;       /* "# NumPatches" */
    int n_dark;          /* "# DarkFrame" lines */
    float alpha_offset;  /* "# AlphaOffset" */
    float beta_offset;   /* "# BetaOffset" */
    int thickness;       /* "# Thickness" (CAlignParam::m_iThickness) */
} atr_aln_info;

/* Loads and validates the file: ATR_ERR_INVALID_ARG if it cannot be opened,
 * has no RawSize, or has fewer well-formed global / local rows than the
 * header implies (the reference prints an error and keeps zeros there). */
atr_status atr_aln_read(const char* path, atr_aln** aln);
atr_status atr_aln_get_info(const atr_aln* aln, atr_aln_info* info);
/* Global rows, n_tilts each (any pointer may be NULL): section index,
 * tilt axis (ROT), shift (TX, TY interleaved), tilt angle (TILT). */
atr_status atr_aln_get_global(const atr_aln* aln, int* sec, float* tilt_axis, float* shift_xy,
                              float* tilt);
/* Local rows in the layout atr_series_desc.local_params takes:
 * [n_tilts][5][n_patches] = centre x, centre y, shift x, shift y, good
 * (1 when (int)(good + 0.5) != 0, CLoadAlignFile.cpp:272-275). */
atr_status atr_aln_get_local(const atr_aln* aln, float* local_params);
/* Dark frames, n_dark each: frame index, section index, tilt. */
atr_status atr_aln_get_dark(const atr_aln* aln, int* frame, int* sec, float* tilt);
void atr_aln_destroy(atr_aln* aln);

/* Writes a .aln in the reference layout (CSaveAlignFile.cpp:67-127):
 * header tags, one global row per tilt (GMAG, SMEAN, SFIT, SCALE 1, BASE 0)
 * and, when n_patches > 0, the local rows.  dark_* may be NULL when
 * info->n_dark is 0; local_params may be NULL when n_patches is 0. */
atr_status atr_aln_write(const char* path, const atr_aln_info* info, const int* sec,
                         const float* tilt_axis, const float* shift_xy, const float* tilt,
                         const float* local_params, const int* dark_frame, const int* dark_sec,
                         const float* dark_tilt);

/* ---- tilt-angle file ------------------------------------------------------ */
/* Reads up to n_sections lines of "tilt [acq_index [dose]]"
 * (DataUtil/CTsPackage.cpp:358-388): lines without a number are skipped,
 * reading stops at n_sections.  acq_index / dose (may be NULL) keep their
 * input values where a line has fewer fields.  ATR_ERR_INVALID_ARG if the
 * file cannot be opened or holds fewer than n_sections angles; *n_read gets
 * the count either way. */
atr_status atr_tilt_file_read(const char* path, int n_sections, float* tilts, int* acq_index,
                              float* dose, int* n_read);

#ifdef __cplusplus
}
#endif
#endif
