// files.cpp -- MRC stacks and volumes, the .aln alignment file and the
// tilt-angle file (host code; declarations in include/aretomo_files.h).
//
// MRC follows the MRC2014 layout that LibSrc/Mrcfile reads and writes: a
// 1024-byte main header of 56 little-endian 32-bit words (some split into
// 16-bit halves) plus ten 80-byte labels, nsymbt bytes of extended header,
// then the sections.  The header is assembled field by field at its byte
// offset, so the file bytes do not depend on this compiler's struct layout.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/aretomo_files.h"
#include "atr_internal.h"

using atr::fail;

namespace {

// ---- byte helpers -----------------------------------------------------------
inline uint32_t bswap32(uint32_t v)
{
    return (v >> 24) | ((v >> 8) & 0xff00u) | ((v << 8) & 0xff0000u) | (v << 24);
}
inline uint16_t bswap16(uint16_t v) { return (uint16_t)((v >> 8) | (v << 8)); }

struct HeaderWriter {
    unsigned char b[1024];
    HeaderWriter() { std::memset(b, 0, sizeof(b)); }
    void i32(int off, int32_t v) { std::memcpy(b + off, &v, 4); }
    void f32(int off, float v) { std::memcpy(b + off, &v, 4); }
    void i16(int off, int16_t v) { std::memcpy(b + off, &v, 2); }
};

struct HeaderReader {
    unsigned char b[1024];
    bool swap = false;
    int32_t i32(int off) const
    {
        uint32_t v;
        std::memcpy(&v, b + off, 4);
        return (int32_t)(swap ? bswap32(v) : v);
    }
    float f32(int off) const
    {
        uint32_t v;
        std::memcpy(&v, b + off, 4);
        if (swap) v = bswap32(v);
        float f;
        std::memcpy(&f, &v, 4);
        return f;
    }
};

// MRC2014 word offsets (bytes)
enum {
    H_NX = 0, H_NY = 4, H_NZ = 8, H_MODE = 12, H_NXSTART = 16, H_MX = 28, H_MY = 32,
    H_MZ = 36, H_XLEN = 40, H_YLEN = 44, H_ZLEN = 48, H_ALPHA = 52, H_BETA = 56,
    H_GAMMA = 60, H_MAPC = 64, H_MAPR = 68, H_MAPS = 72, H_AMIN = 76, H_AMAX = 80,
    H_AMEAN = 84, H_ISPG = 88, H_NSYMBT = 92, H_DVID = 96, H_EXTTYPE = 104,
    H_NVERSION = 108, H_NINT = 128, H_NREAL = 130, H_SUB = 132, H_TYPE = 160,
    H_XTILT = 184, H_YTILT = 188, H_ZTILT = 192, H_MAP = 208, H_MACHST = 212,
    H_NLABL = 220
};

struct FileCloser {
    void operator()(FILE* f) const
    {
        if (f) std::fclose(f);
    }
};
using File = std::unique_ptr<FILE, FileCloser>;

atr_status read_header(FILE* f, HeaderReader& h, atr_mrc_info* info)
{
    if (std::fread(h.b, 1, 1024, f) != 1024)
        return fail(ATR_ERR_INVALID_ARG, "MRC file shorter than its 1024-byte header");
    // LibSrc/Mrcfile/Header/CLoadMainHeader.cpp:27-37
    h.swap = false;
    const int max_size = 32 * 1024;
    int nx = h.i32(H_NX), ny = h.i32(H_NY), nz = h.i32(H_NZ), mode = h.i32(H_MODE);
    if (mode < 0 || nx < 0 || ny < 0 || nz < 0 || nx > max_size || ny > max_size || nz > max_size)
        h.swap = true;
    info->nx = h.i32(H_NX);
    info->ny = h.i32(H_NY);
    info->nz = h.i32(H_NZ);
    info->mode = h.i32(H_MODE);
    int mx = h.i32(H_MX);
    info->pixel_size = mx != 0 ? h.f32(H_XLEN) / (float)mx : 0.0f;
    info->nsymbt = h.i32(H_NSYMBT);
    info->swapped = h.swap ? 1 : 0;
    if (info->nx < 1 || info->ny < 1 || info->nz < 1 || info->nsymbt < 0)
        return fail(ATR_ERR_INVALID_ARG, "MRC header with non-positive sizes");
    return ATR_OK;
}

int mode_bytes(int mode)
{
    switch (mode) {
    case 1: return 2;
    case 2: return 4;
    case 6: return 2;
    default: return 0;
    }
}

// ---- .aln ------------------------------------------------------------------
const char* kRawSize = "RawSize";
const char* kNumPatches = "NumPatches";
const char* kDarkFrame = "DarkFrame";
const char* kAlphaOffset = "AlphaOffset";
const char* kBetaOffset = "BetaOffset";
const char* kThickness = "Thickness";

// Text after the first '=' of a tag line, or nullptr.
const char* after_eq(const std::string& s)
{
    size_t p = s.find('=');
    return p == std::string::npos ? nullptr : s.c_str() + p + 1;
}

}  // namespace

struct atr_aln {
    atr_aln_info info{};
    std::vector<int> sec;
    std::vector<float> axis, shift, tilt;
    std::vector<float> local;  // [n][5][np]
    std::vector<int> dark_frame, dark_sec;
    std::vector<float> dark_tilt;
};

extern "C" {

atr_status atr_mrc_read_info(const char* path, atr_mrc_info* info)
{
    if (!path || !info) return fail(ATR_ERR_INVALID_ARG, "atr_mrc_read_info: NULL argument");
    File f(std::fopen(path, "rb"));
    if (!f) return fail(ATR_ERR_INVALID_ARG, std::string("cannot open MRC file ") + path);
    HeaderReader h;
    return read_header(f.get(), h, info);
}

atr_status atr_mrc_read_sections(const char* path, int first, int count, float* out)
{
    if (!path || !out) return fail(ATR_ERR_INVALID_ARG, "atr_mrc_read_sections: NULL argument");
    File f(std::fopen(path, "rb"));
    if (!f) return fail(ATR_ERR_INVALID_ARG, std::string("cannot open MRC file ") + path);
    HeaderReader h;
    atr_mrc_info info;
    atr_status st = read_header(f.get(), h, &info);
    if (st != ATR_OK) return st;
    const int bytes = mode_bytes(info.mode);
    if (bytes == 0)
        return fail(ATR_ERR_UNSUPPORTED_SIZE, "MRC mode " + std::to_string(info.mode) +
                                                  " (readable: 1 int16, 2 float32, 6 uint16)");
    if (first < 0 || count < 0 || first + count > info.nz)
        return fail(ATR_ERR_INVALID_ARG, "MRC section range outside the stack");
    const size_t pix = (size_t)info.nx * info.ny;
    const long long off = 1024LL + info.nsymbt + (long long)first * (long long)(pix * bytes);
    if (fseeko(f.get(), (off_t)off, SEEK_SET) != 0)
        return fail(ATR_ERR_INVALID_ARG, "MRC seek failed");
    std::vector<unsigned char> buf(pix * bytes);
    for (int s = 0; s < count; s++) {
        if (std::fread(buf.data(), 1, buf.size(), f.get()) != buf.size())
            return fail(ATR_ERR_INVALID_ARG, "MRC file shorter than its header says");
        float* dst = out + (size_t)s * pix;
        if (info.mode == 2) {
            for (size_t i = 0; i < pix; i++) {
                uint32_t v;
                std::memcpy(&v, buf.data() + 4 * i, 4);
                if (h.swap) v = bswap32(v);
                std::memcpy(dst + i, &v, 4);
            }
        } else {
            for (size_t i = 0; i < pix; i++) {
                uint16_t v;
                std::memcpy(&v, buf.data() + 2 * i, 2);
                if (h.swap) v = bswap16(v);
                dst[i] = info.mode == 1 ? (float)(int16_t)v : (float)v;
            }
        }
    }
    return ATR_OK;
}

atr_status atr_mrc_write(const char* path, const float* data, int nx, int ny, int nz,
                         float pixel_size, int with_stats)
{
    if (!path || !data) return fail(ATR_ERR_INVALID_ARG, "atr_mrc_write: NULL argument");
    if (nx < 1 || ny < 1 || nz < 1) return fail(ATR_ERR_INVALID_ARG, "atr_mrc_write: empty stack");
    const size_t pix = (size_t)nx * ny, total = pix * nz;
    // LibSrc/Mrcfile/Header/CSaveMainHeader.cpp mSetDefault + SetImgSize
    HeaderWriter h;
    h.i32(H_NX, nx);
    h.i32(H_NY, ny);
    h.i32(H_NZ, nz);
    h.i32(H_MODE, 2);
    h.i32(H_MX, nx);
    h.i32(H_MY, ny);
    h.i32(H_MZ, nz);
    h.f32(H_XLEN, nx * pixel_size);
    h.f32(H_YLEN, ny * pixel_size);
    h.f32(H_ZLEN, nz * pixel_size);
    h.f32(H_ALPHA, 90.0f);
    h.f32(H_BETA, 90.0f);
    h.f32(H_GAMMA, 90.0f);
    h.i32(H_MAPC, 1);
    h.i32(H_MAPR, 2);
    h.i32(H_MAPS, 3);
    if (with_stats) {
        float lo = data[0], hi = data[0];
        double sum = 0.0;
        for (size_t i = 0; i < total; i++) {
            lo = std::min(lo, data[i]);
            hi = std::max(hi, data[i]);
            sum += data[i];
        }
        h.f32(H_AMIN, lo);
        h.f32(H_AMAX, hi);
        h.f32(H_AMEAN, (float)(sum / (double)total));
    }
    h.i16(H_DVID, -16224);
    std::memcpy(h.b + H_EXTTYPE, "AGAR", 4);
    h.i32(H_NVERSION, 20140);
    h.i16(H_SUB, 1);
    h.i16(H_TYPE, 5);
    h.f32(H_XTILT, 90.0f);
    h.f32(H_YTILT, 90.0f);
    h.f32(H_ZTILT, 90.0f);
    std::memcpy(h.b + H_MAP, "MAP", 4);  // with its terminating zero
    const unsigned char stamp[4] = {0x44, 0x44, 0x00, 0x00};
    std::memcpy(h.b + H_MACHST, stamp, 4);
    File f(std::fopen(path, "wb"));
    if (!f) return fail(ATR_ERR_INVALID_ARG, std::string("cannot create MRC file ") + path);
    if (std::fwrite(h.b, 1, 1024, f.get()) != 1024)
        return fail(ATR_ERR_INVALID_ARG, "MRC header write failed");
    const size_t chunk = (size_t)1 << 24;
    for (size_t i = 0; i < total; i += chunk) {
        size_t n = std::min(chunk, total - i);
        if (std::fwrite(data + i, sizeof(float), n, f.get()) != n)
            return fail(ATR_ERR_INVALID_ARG, "MRC data write failed");
    }
    if (std::fclose(f.release()) != 0) return fail(ATR_ERR_INVALID_ARG, "MRC close failed");
    return ATR_OK;
}

const char* atr_volume_suffix(int nth_vol)
{
    static const char* const names[5] = {"_Vol.mrc", "_EVN_Vol.mrc", "_ODD_Vol.mrc",
                                         "_2ND_Vol.mrc", "_3RD_Vol.mrc"};
    return (nth_vol >= 0 && nth_vol < 5) ? names[nth_vol] : nullptr;
}

atr_status atr_aln_read(const char* path, atr_aln** out)
{
    if (!path || !out) return fail(ATR_ERR_INVALID_ARG, "atr_aln_read: NULL argument");
    *out = nullptr;
    File f(std::fopen(path, "rt"));
    if (!f) return fail(ATR_ERR_INVALID_ARG, std::string("cannot open alignment file ") + path);
    // CLoadAlignFile.cpp:51-66: lines shorter than 4 characters are skipped,
    // '#' lines are header, every other line is data (global rows, then local)
    std::vector<std::string> header, data;
    char buf[256];
    while (std::fgets(buf, sizeof(buf), f.get())) {
        if (std::strlen(buf) < 4) continue;
        if (buf[0] == '#') header.emplace_back(buf + 1);
        else data.emplace_back(buf);
    }
    auto a = std::make_unique<atr_aln>();
    atr_aln_info& I = a->info;
    bool have_raw = false;
    for (const std::string& ln : header) {
        const char* v = after_eq(ln);
        if (ln.find(kRawSize) != std::string::npos && v) {
            I.raw_size[0] = I.raw_size[1] = I.raw_size[2] = 0;
            std::sscanf(v, "%d %d %d", &I.raw_size[0], &I.raw_size[1], &I.raw_size[2]);
            have_raw = true;
        } else if (ln.find(kDarkFrame) != std::string::npos && v) {
            int fr = 0, sc = 0;
            float t = 0.0f;
            std::sscanf(v, "%d %d %f", &fr, &sc, &t);
            a->dark_frame.push_back(fr);
            a->dark_sec.push_back(sc);
            a->dark_tilt.push_back(t);
        } else if (ln.find(kNumPatches) != std::string::npos && v) {
            std::sscanf(v, "%d", &I.n_patches);
        } else if (ln.find(kAlphaOffset) != std::string::npos && v) {
            std::sscanf(v, "%f", &I.alpha_offset);
        } else if (ln.find(kBetaOffset) != std::string::npos && v) {
            std::sscanf(v, "%f", &I.beta_offset);
        } else if (ln.find(kThickness) != std::string::npos && v) {
            std::sscanf(v, "%d", &I.thickness);
        }
    }
    if (!have_raw) return fail(ATR_ERR_INVALID_ARG, "alignment file has no RawSize line");
    I.n_dark = (int)a->dark_frame.size();
    // CDarkFrames::GetNumAlnTilts: raw sections minus dark ones
    I.n_tilts = std::max(0, I.raw_size[2] - I.n_dark);
    if (I.n_tilts <= 0) return fail(ATR_ERR_INVALID_ARG, "alignment file lists no aligned tilt");
    if (I.n_patches < 0) I.n_patches = 0;
    const int n = I.n_tilts, np = I.n_patches;
    if ((size_t)n > data.size())
        return fail(ATR_ERR_INVALID_ARG, "alignment file has " + std::to_string(data.size()) +
                                             " data rows, " + std::to_string(n) + " global rows expected");
    a->sec.assign(n, 0);
    a->axis.assign(n, 0.0f);
    a->shift.assign(2 * (size_t)n, 0.0f);
    a->tilt.assign(n, 0.0f);
    // CLoadAlignFile.cpp:224-251
    int good_rows = 0;
    for (int i = 0; i < n; i++) {
        int sec;
        float ax, gmag, tx, ty, smean, sfit, scale, base, tilt;
        int k = std::sscanf(data[i].c_str(), "%d %f %f %f %f %f %f %f %f %f", &sec, &ax, &gmag, &tx,
                            &ty, &smean, &sfit, &scale, &base, &tilt);
        if (k != 10) continue;
        a->sec[i] = sec;
        a->axis[i] = ax;
        a->shift[2 * i] = tx;
        a->shift[2 * i + 1] = ty;
        a->tilt[i] = tilt;
        good_rows++;
    }
    if (good_rows != n)
        return fail(ATR_ERR_INVALID_ARG, "alignment file: " + std::to_string(good_rows) + " of " +
                                             std::to_string(n) + " global rows readable");
    // CLoadAlignFile.cpp:256-284: rows "t p x y sx sy good"
    if (np > 0) {
        const size_t want = (size_t)n * np;
        a->local.assign(want * 5, 0.0f);
        size_t read = 0;
        for (size_t i = 0; i < want && n + i < data.size(); i++) {
            int t, p;
            float x, y, sx, sy, g;
            int k = std::sscanf(data[n + i].c_str(), "%d %d %f %f %f %f %f", &t, &p, &x, &y, &sx,
                                &sy, &g);
            if (k != 7) continue;
            if (t < 0 || t >= n || p < 0 || p >= np)
                return fail(ATR_ERR_INVALID_ARG, "alignment file: local row outside the table");
            float* q = a->local.data() + (size_t)t * 5 * np;
            q[p] = x;
            q[np + p] = y;
            q[2 * np + p] = sx;
            q[3 * np + p] = sy;
            q[4 * np + p] = ((int)(g + 0.5f) == 0) ? 0.0f : 1.0f;
            read++;
        }
        if (read < want)
            return fail(ATR_ERR_INVALID_ARG, "alignment file: " + std::to_string(read) + " of " +
                                                 std::to_string(want) + " local rows readable");
    }
    // CLoadAlignFile::mToOneBased with CAlignParam/CDarkFrames::bZeroBased
    bool aln0 = std::find(a->sec.begin(), a->sec.end(), 0) != a->sec.end();
    bool dark0 = false;
    if (I.n_dark > 0) {
        dark0 = a->dark_frame[0] == a->dark_sec[0] ||
                std::find(a->dark_sec.begin(), a->dark_sec.end(), 0) != a->dark_sec.end();
    }
    if (aln0 || dark0) {
        for (int& s : a->sec) s += 1;
        for (int i = 0; i < I.n_dark; i++) a->dark_sec[i] = a->dark_frame[i] + 1;
    }
    *out = a.release();
    return ATR_OK;
}

atr_status atr_aln_get_info(const atr_aln* aln, atr_aln_info* info)
{
    if (!aln || !info) return fail(ATR_ERR_INVALID_ARG, "atr_aln_get_info: NULL argument");
    *info = aln->info;
    return ATR_OK;
}

atr_status atr_aln_get_global(const atr_aln* aln, int* sec, float* tilt_axis, float* shift_xy,
                              float* tilt)
{
    if (!aln) return fail(ATR_ERR_INVALID_ARG, "atr_aln_get_global: NULL alignment");
    const size_t n = aln->info.n_tilts;
    if (sec) std::copy(aln->sec.begin(), aln->sec.end(), sec);
    if (tilt_axis) std::copy(aln->axis.begin(), aln->axis.end(), tilt_axis);
    if (shift_xy) std::copy(aln->shift.begin(), aln->shift.begin() + 2 * n, shift_xy);
    if (tilt) std::copy(aln->tilt.begin(), aln->tilt.end(), tilt);
    return ATR_OK;
}

atr_status atr_aln_get_local(const atr_aln* aln, float* local_params)
{
    if (!aln || !local_params) return fail(ATR_ERR_INVALID_ARG, "atr_aln_get_local: NULL argument");
    if (aln->info.n_patches <= 0) return fail(ATR_ERR_INVALID_ARG, "alignment file has no local rows");
    std::copy(aln->local.begin(), aln->local.end(), local_params);
    return ATR_OK;
}

atr_status atr_aln_get_dark(const atr_aln* aln, int* frame, int* sec, float* tilt)
{
    if (!aln) return fail(ATR_ERR_INVALID_ARG, "atr_aln_get_dark: NULL alignment");
    if (frame) std::copy(aln->dark_frame.begin(), aln->dark_frame.end(), frame);
    if (sec) std::copy(aln->dark_sec.begin(), aln->dark_sec.end(), sec);
    if (tilt) std::copy(aln->dark_tilt.begin(), aln->dark_tilt.end(), tilt);
    return ATR_OK;
}

void atr_aln_destroy(atr_aln* aln) { delete aln; }

atr_status atr_aln_write(const char* path, const atr_aln_info* info, const int* sec,
                         const float* tilt_axis, const float* shift_xy, const float* tilt,
                         const float* local_params, const int* dark_frame, const int* dark_sec,
                         const float* dark_tilt)
{
    if (!path || !info || !sec || !tilt_axis || !shift_xy || !tilt)
        return fail(ATR_ERR_INVALID_ARG, "atr_aln_write: NULL argument");
    const int n = info->n_tilts, np = info->n_patches;
    if (n < 1 || np < 0 || info->n_dark < 0) return fail(ATR_ERR_INVALID_ARG, "atr_aln_write: bad counts");
    if (info->n_dark > 0 && (!dark_frame || !dark_sec || !dark_tilt))
        return fail(ATR_ERR_INVALID_ARG, "atr_aln_write: dark frames without arrays");
    if (np > 0 && !local_params) return fail(ATR_ERR_INVALID_ARG, "atr_aln_write: no local table");
    File f(std::fopen(path, "wt"));
    if (!f) return fail(ATR_ERR_INVALID_ARG, std::string("cannot create alignment file ") + path);
    FILE* o = f.get();
    // header: CSaveAlignFile.cpp:67-105
    std::fprintf(o, "# AreTomo Alignment / Priims bprmMn\n");
    std::fprintf(o, "# %s = %d %d %d\n", kRawSize, info->raw_size[0], info->raw_size[1],
                 info->raw_size[2]);
    std::fprintf(o, "# %s = %d\n", kNumPatches, np);
    for (int i = 0; i < info->n_dark; i++)
        std::fprintf(o, "# %s =  %4d %4d %8.2f\n", kDarkFrame, dark_frame[i], dark_sec[i], dark_tilt[i]);
    std::fprintf(o, "# %s = %8.2f\n", kAlphaOffset, info->alpha_offset);
    std::fprintf(o, "# %s = %8.2f\n", kBetaOffset, info->beta_offset);
    std::fprintf(o, "# %s = %d\n", kThickness, info->thickness);
    // global rows: CSaveAlignFile.cpp:107-123
    std::fprintf(o, "# SEC     ROT         GMAG       TX          TY      SMEAN     SFIT    "
                    "SCALE     BASE     TILT\n");
    for (int i = 0; i < n; i++)
        std::fprintf(o, "%5d  %9.4f  %9.5f  %9.3f  %9.3f  %7.2f  %7.2f  %7.2f  %7.2f  %8.2f\n", sec[i],
                     tilt_axis[i], 1.0f, shift_xy[2 * i], shift_xy[2 * i + 1], 1.0f, 1.0f, 1.0f,
                     0.0f, tilt[i]);
    // local rows: CSaveAlignFile.cpp:125-140, table [t][5][np] in memory
    if (np > 0) {
        std::fprintf(o, "# Local Alignment\n");
        for (int t = 0; t < n; t++) {
            const float* q = local_params + (size_t)t * 5 * np;
            for (int p = 0; p < np; p++)
                std::fprintf(o, "%4d %3d %8.2f  %8.2f  %8.2f  %8.2f  %4.1f\n", t, p, q[p], q[np + p],
                             q[2 * np + p], q[3 * np + p], q[4 * np + p]);
        }
    }
    if (std::fclose(f.release()) != 0) return fail(ATR_ERR_INVALID_ARG, "alignment file close failed");
    return ATR_OK;
}

atr_status atr_tilt_file_read(const char* path, int n_sections, float* tilts, int* acq_index,
                              float* dose, int* n_read)
{
    if (n_read) *n_read = 0;
    if (!path || !tilts || n_sections < 1)
        return fail(ATR_ERR_INVALID_ARG, "atr_tilt_file_read: NULL argument or no sections");
    File f(std::fopen(path, "rt"));
    if (!f) return fail(ATR_ERR_INVALID_ARG, std::string("cannot open tilt file ") + path);
    char buf[256];
    int count = 0;
    while (count < n_sections && std::fgets(buf, sizeof(buf), f.get())) {
        float t = 0.0f, d = 0.0f;
        int a = 0;
        int k = std::sscanf(buf, "%f %d %f", &t, &a, &d);
        if (k < 1) continue;
        tilts[count] = t;
        if (k >= 2 && acq_index) acq_index[count] = a;
        if (k >= 3 && dose) dose[count] = d;
        count++;
    }
    if (n_read) *n_read = count;
    if (count < n_sections)
        return fail(ATR_ERR_INVALID_ARG, "tilt file holds " + std::to_string(count) + " angles, " +
                                             std::to_string(n_sections) + " sections");
    return ATR_OK;
}

}  // extern "C"
