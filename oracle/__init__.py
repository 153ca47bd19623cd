"""FP64 CPU oracle for AreTomo3's reconstruction hot path (ctypes over liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the reported CPU
baseline.  Only ``tests/``, ``__graft_entry__.smoke()`` and the cpu_baseline
leg of ``bench.py`` import this package.  The product (``aretomo3_amd``)
never imports, links or calls it.

Each wrapper names the C function in ``oracle.c``; the C functions cite the
reference file:line they restate.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_c_int, _c_u64, _c_f, _c_d, _c_vp = (ctypes.c_int, ctypes.c_uint64, ctypes.c_float,
                                      ctypes.c_double, ctypes.c_void_p)


def build():
    """Compile liboracle.so from oracle.c (gcc, OpenMP)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "orc_splitmix64": (ctypes.c_uint64, [_c_u64]),
            "orc_fill_table": (None, [_c_u64, _c_int, _ip]),
            "orc_cos_sin": (None, [_fp, _c_int, _fp]),
            "orc_split_items": (None, [_c_int, _c_int, _ip, _ip]),
            "orc_subset_tilts": (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int, _ip]),
            "orc_fft": (None, [_dp, _dp, _c_int, _c_int]),
            "orc_aligned_size": (None, [_c_int, _c_int, _c_f, ctypes.POINTER(_c_int),
                                        ctypes.POINTER(_c_int)]),
            "orc_weight_projs": (None, [_dp, _c_int, _c_int, _c_int, _fp, _c_int]),
            "orc_rweight": (None, [_dp, _c_int, _c_int, _c_int]),
            "orc_back_proj": (None, [_dp, _c_int, _c_int, _fp, _c_vp, _ip, _c_int, _c_int,
                                     _c_d, _dp, _c_int, _c_int]),
            "orc_for_proj": (None, [_dp, _c_int, _c_int, _fp, _ip, _c_int, _c_int, _c_int, _dp]),
            "orc_diff_proj": (None, [_dp, _dp, _ip, _c_int, _c_int, _c_int]),
            "orc_fill_empty": (None, [_dp, _c_int, _c_int, _c_int, _ip]),
            "orc_back_proj_wbp": (None, [_dp, _c_int, _c_int, _fp, _c_vp, _c_int, _dp,
                                         _c_int, _c_int]),
            "orc_wbp_crop": (None, [_dp, _c_int, _c_int, _dp]),
            "orc_fourier_crop": (None, [_dp, _c_int, _c_int, _c_int, _c_int, _dp]),
            "orc_reproj_range": (None, [_fp, _c_vp, _c_int, _c_int, _ip]),
            "orc_reproj": (None, [_fp, _fp, _c_int, _c_int, _c_int, _c_int, _c_int, _ip, _dp]),
            "orc_thick_profile": (None, [_fp, _c_int, _c_int, _c_int, _fp]),
            "orc_thick_smooth": (None, [_fp, _c_int]),
            "orc_thick_edges": (None, [_fp, _c_int, _c_f, _ip]),
            "orc_bin_size": (None, [_c_int, _c_int, _c_f, ctypes.POINTER(_c_int),
                                    ctypes.POINTER(_c_int)]),
            "orc_bin_stack": (None, [_dp, _c_int, _c_int, _c_int, _c_int, _c_int, _dp]),
            "orc_tomo_sart": (None, [_dp, _c_int, _c_int, _c_int, _fp, _c_vp, _dp, _c_int,
                                     _c_int, _c_int, _c_int, _c_int, _c_int, _c_d, _c_int,
                                     _c_int, _ip]),
            "orc_tomo_wbp": (None, [_dp, _c_int, _c_int, _c_int, _fp, _c_vp, _dp, _c_int,
                                    _c_int]),
            "orc_corr_proj": (None, [_fp, _c_int, _c_int, _c_f, _c_f, _c_f, _c_int, _c_vp,
                                     _dp, _c_int, _c_int, _c_int, _c_int, _c_vp]),
            "orc_prestage": (None, [_fp, _c_int, _c_int, _c_int, _fp, _c_vp, _c_int, _c_vp,
                                    _c_int, _c_u64, _dp, _c_int, _c_int]),
            "orc_recon_slices": (None, [_dp, _c_int, _c_int, _c_int, _fp, _c_vp, _c_int,
                                        _c_int, _c_int, _c_int, _c_int, _c_int, _c_d, _c_int,
                                        _c_int, _c_u64, _ip, _c_int, _dp]),
            "orc_flip_vol": (None, [_dp, _c_int, _c_int, _c_int, _c_int, _dp]),
            "orc_max_threads": (_c_int, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _mask(no_proj):
    if no_proj is None:
        return None, None
    m = np.ascontiguousarray(no_proj, dtype=np.uint8)
    return m, m.ctypes.data_as(ctypes.c_void_p)


def splitmix64(x):
    return int(lib().orc_splitmix64(x))


def fill_table(seed, n):
    t = np.empty(n, np.int32)
    lib().orc_fill_table(seed, n, t)
    return t


def cos_sin(tilts):
    tilts = _f32(tilts)
    out = np.empty(2 * len(tilts), np.float32)
    lib().orc_cos_sin(tilts, len(tilts), out)
    return out


def split_items(n_items, n_splits):
    st = np.empty(n_splits, np.int32)
    sz = np.empty(n_splits, np.int32)
    lib().orc_split_items(n_items, n_splits, st, sz)
    return st, sz


def subset_tilts(start, n_tilts, n_subsets, order, subset):
    buf = np.empty(n_tilts + 1, np.int32)
    n = lib().orc_subset_tilts(start, n_tilts, n_subsets, order, subset, buf)
    return buf[:n].copy()


def fft(x, sign=-1):
    x = np.asarray(x, np.complex128)
    re, im = _f64(x.real), _f64(x.imag)
    lib().orc_fft(re, im, len(x), sign)
    return re + 1j * im


def aligned_size(raw_x, raw_y, axis):
    ax, ay = _c_int(), _c_int()
    lib().orc_aligned_size(raw_x, raw_y, axis, ctypes.byref(ax), ctypes.byref(ay))
    return ax.value, ay.value


def weight_projs(sino, proj_x, cs, vol_z):
    s = _f64(sino).copy()
    lib().orc_weight_projs(s, s.shape[0], s.shape[1], proj_x, _f32(cs), vol_z)
    return s


def rweight(sino, fft_size):
    s = _f64(sino).copy()
    lib().orc_rweight(s, s.shape[0], s.shape[1], fft_size)
    return s


def back_proj(sino, proj_x, cs, tilts, vol, positivity=1, relax=1.0, no_proj=None):
    s = _f64(sino)
    v = _f64(vol).copy()
    t = _i32(tilts)
    m, mp = _mask(no_proj)
    lib().orc_back_proj(s, s.shape[1], proj_x, _f32(cs), mp, t, len(t), positivity, relax,
                        v, v.shape[1], v.shape[0])
    return v


def for_proj(vol, cs, tilts, proj_x, pad_x, n_tilts):
    v = _f64(vol)
    t = _i32(tilts)
    fwd = np.zeros((n_tilts, pad_x), np.float64)
    lib().orc_for_proj(v, v.shape[1], v.shape[0], _f32(cs), t, len(t), proj_x, pad_x, fwd)
    return fwd


def diff_proj(raw, fwd, tilts, proj_x):
    f = _f64(fwd).copy()
    t = _i32(tilts)
    lib().orc_diff_proj(_f64(raw), f, t, len(t), proj_x, f.shape[1])
    return f


def fill_empty(img, w, table):
    im = _f64(img).copy()
    lib().orc_fill_empty(im, w, im.shape[0], im.shape[1], _i32(table))
    return im


def back_proj_wbp(sino, proj_x, cs, vol_x2, vol_z2, no_proj=None):
    s = _f64(sino)
    out = np.empty((vol_z2, vol_x2), np.float64)
    m, mp = _mask(no_proj)
    lib().orc_back_proj_wbp(s, s.shape[1], proj_x, _f32(cs), mp, s.shape[0], out, vol_x2,
                            vol_z2)
    return out


def wbp_crop(img2, vol_x, vol_z):
    out = np.empty((vol_z, vol_x), np.float64)
    lib().orc_wbp_crop(_f64(img2), vol_x, vol_z, out)
    return out


def fourier_crop(img, bx, by):
    img = _f64(img)
    ay, ax = img.shape
    out = np.empty((by, bx), np.float64)
    lib().orc_fourier_crop(img, ax, ay, bx, by, out)
    return out


def bin_size(ax, ay, bin_):
    bx, by = _c_int(), _c_int()
    lib().orc_bin_size(ax, ay, bin_, ctypes.byref(bx), ctypes.byref(by))
    return bx.value, by.value


def bin_stack(aligned, bin_):
    """Fourier binning of an aligned stack [n][ay][ax] (CBinStack::DoFFT)."""
    a = _f64(aligned)
    n, ay, ax = a.shape
    bx, by = bin_size(ax, ay, float(bin_))
    out = np.empty((n, by, bx), np.float64)
    lib().orc_bin_stack(a, n, ax, ay, bx, by, out)
    return out


def tomo_sart(sino, proj_x, cs, vol, start_tilt, n_sart_tilts, iters, n_subsets,
              relax=-1.0, order=0, positivity=1, fill_tbl=None, no_proj=None):
    s = _f64(sino).copy()
    v = _f64(vol).copy()
    m, mp = _mask(no_proj)
    if fill_tbl is None:
        fill_tbl = np.zeros(v.size, np.int32)
    lib().orc_tomo_sart(s, s.shape[0], s.shape[1], proj_x, _f32(cs), mp, v, v.shape[1],
                        v.shape[0], start_tilt, n_sart_tilts, iters, n_subsets, relax, order,
                        positivity, _i32(fill_tbl))
    return v, s


def tomo_wbp(sino, proj_x, cs, vol_x, vol_z, no_proj=None):
    s = _f64(sino).copy()
    v = np.zeros((vol_z, vol_x), np.float64)
    m, mp = _mask(no_proj)
    lib().orc_tomo_wbp(s, s.shape[0], s.shape[1], proj_x, _f32(cs), mp, v, vol_x, vol_z)
    return v, s


def corr_proj(raw, axis, shift, out_x, out_y, out_pad_x, random_fill=0, table=None,
              local=None):
    """local: patch table of this tilt, [5][n_patches] (cx, cy, sx, sy, good)."""
    raw = _f32(raw)
    out = np.zeros((out_y, out_pad_x), np.float64)
    tp = None
    if table is not None:
        table = _i32(table)
        tp = table.ctypes.data_as(ctypes.c_void_p)
    npat, lp = 0, None
    if local is not None:
        local = _f32(local).reshape(5, -1)
        npat, lp = local.shape[1], local.ctypes.data_as(ctypes.c_void_p)
    lib().orc_corr_proj(raw, raw.shape[1], raw.shape[0], axis, shift[0], shift[1], npat, lp,
                        out, out_x, out_y, out_pad_x, random_fill, tp)
    return out


def prestage(raw, axes, shifts, random_fill=1, seed=0, local=None):
    """raw: [n][raw_y][raw_x] float32 -> aligned [n][aln_y][aln_x] float64.
    local: per-tilt patch tables [n][5][n_patches] (cx, cy, sx, sy, good)."""
    raw = _f32(raw)
    n, ry, rx = raw.shape
    ax, ay = aligned_size(rx, ry, float(np.float32(axes[0])))
    out = np.empty((n, ay, ax), np.float64)
    sh = None
    if shifts is not None:
        sh = _f32(shifts)
    npat, lp = 0, None
    if local is not None:
        local = _f32(local).reshape(n, 5, -1)
        npat, lp = local.shape[2], local.ctypes.data_as(ctypes.c_void_p)
    lib().orc_prestage(raw, n, rx, ry, _f32(axes),
                       None if sh is None else sh.ctypes.data_as(ctypes.c_void_p),
                       npat, lp, random_fill, seed, out, ax, ay)
    return out


def recon_slices(aligned, tilts, method, vol_z, slices, start_tilt=0, n_sart_tilts=None,
                 iters=20, n_subsets=5, relax=-1.0, order=0, positivity=1, seed=0,
                 no_proj=None):
    """Reference per-series recon (CDoWbpRecon / CDoSartRecon) on selected slices.

    aligned: [n][proj_y][proj_x]; returns [len(slices)][vol_z][vol_x] in the
    reference frame layout (z flipped)."""
    a = _f64(aligned)
    n, py, px = a.shape
    if n_sart_tilts is None:
        n_sart_tilts = n - start_tilt
    sl = _i32(slices)
    vol_x = px // 2 * 2
    out = np.empty((len(sl), vol_z, vol_x), np.float64)
    m, mp = _mask(no_proj)
    lib().orc_recon_slices(a, n, px, py, _f32(tilts), mp, method, vol_z, start_tilt,
                           n_sart_tilts, iters, n_subsets, relax, order, positivity, seed,
                           sl, len(sl), out)
    return out


def reproj_range(tilts, proj_idx, skip=None):
    """CCalcReproj::mFindProjRange -> (first, last) tilt index."""
    t = _f32(tilts)
    m, mp = _mask(skip)
    r = np.zeros(2, np.int32)
    lib().orc_reproj_range(t, mp, len(t), int(proj_idx), r)
    return int(r[0]), int(r[1])


def reproj(projs, tilts, vol_z, proj_idx, skip=None, rng=None):
    """CCalcReproj::DoIt: projs [n][proj_y][proj_x] float32 -> [proj_y][proj_x]."""
    p = _f32(projs)
    n, py, px = p.shape
    if rng is None:
        rng = reproj_range(tilts, proj_idx, skip)
    r = np.asarray(rng, np.int32)
    out = np.empty((py, px), np.float64)
    lib().orc_reproj(p, _f32(tilts), n, px, py, int(vol_z), int(proj_idx), r, out)
    return out


def thick_profile(vol_xyz):
    """CCalcVolThick::mMeasureStd over z = 0 .. nz-2 of an xyz volume [nz][ny][nx]."""
    v = _f32(vol_xyz)
    nz, ny, nx = v.shape
    out = np.zeros(max(nz - 1, 0), np.float32)
    lib().orc_thick_profile(v, nx, ny, nz, out)
    return out


def thick_smooth(profile):
    p = _f32(profile).copy()
    lib().orc_thick_smooth(p, len(p))
    return p


def thick_edges(smoothed, binning):
    """CCalcVolThick::mDetectEdges -> (low edge, high edge, thickness, offset z)."""
    out = np.zeros(4, np.int32)
    lib().orc_thick_edges(_f32(smoothed), len(smoothed), float(binning), out)
    return tuple(int(v) for v in out)


def calc_vol_thick(raw, tilts, axes, shifts, pixel_size, seed=0):
    """CCalcVolThick::DoIt on the CPU: pre-stage with fill, Fourier binning by
    max(20 / pixel, 20), WBP of thickness (int)(0.8 binned x)/2*2, z-flipped xyz
    volume, std profile, smoothing, edges."""
    binning = max(np.float32(20.0) / np.float32(pixel_size), np.float32(20.0))
    aligned = prestage(raw, axes, shifts, 1, seed) if axes is not None else _f64(raw)
    binned = bin_stack(aligned, float(binning))
    n, by, bx = binned.shape
    vol_z = int(bx * 0.8) // 2 * 2
    vol = recon_slices(binned, tilts, 0, vol_z, list(range(by)), seed=seed)
    xyz = flip_vol(vol, 1).astype(np.float32)
    prof = thick_profile(xyz)
    sm = thick_smooth(prof)
    return dict(edges=thick_edges(sm, binning), profile=prof, smoothed=sm, vol_z=vol_z,
                binning=float(binning), volume=xyz)


def flip_vol(vol_xzy, flip=1):
    v = _f64(vol_xzy)
    ny, nz, nx = v.shape
    out = np.empty((nz, ny, nx), np.float64)
    lib().orc_flip_vol(v, nx, nz, ny, flip, out)
    return out


def max_threads():
    return int(lib().orc_max_threads())
