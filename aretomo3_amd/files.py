"""On-disk formats and the reconstruct-from-files driver (AreTomo3 ``-Cmd 2``).

* MRC stacks in (DataUtil/CTsPackage.cpp:307-356) and float MRC volumes out
  (DataUtil/CAsyncSaveVol.cpp:82-126), through ``atr_mrc_*``.
* The ``.aln`` alignment file (AreTomo/MrcUtil/CLoadAlignFile.cpp,
  CSaveAlignFile.cpp:67-127), through ``atr_aln_*``.
* The ``_TLT.txt`` / ``.rawtlt`` tilt-angle file
  (DataUtil/CTsPackage.cpp:358-388), through ``atr_tilt_file_read``.
* :func:`reconstruct_all` / :func:`reconstruct_files`: raw MRC + ``.aln`` ->
  volume MRCs, the reconstruction part of CAreTomoMain::mSkipAlign
  (AreTomo/CAreTomoMain.cpp:138-171) with mRecon (:725-761), mRecon2nd
  (:763-797) and mSartRecon / mWbpRecon (:809-879).

Parsing and writing run in the C library; this module only moves arrays.
"""
import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from ._lib import check, lib
from .recon import Plan


def _path(p):
    return os.fsencode(os.fspath(p))


# ---- MRC ------------------------------------------------------------------
def mrc_info(path):
    """Header summary: dict(nx, ny, nz, mode, pixel_size, nsymbt, swapped)."""
    info = L.MrcInfo()
    check(lib().atr_mrc_read_info(_path(path), ctypes.byref(info)))
    return {k: getattr(info, k) for k, _ in L.MrcInfo._fields_}


def read_mrc(path, first=0, count=None):
    """Sections [first, first+count) as float32 [count][ny][nx]."""
    info = mrc_info(path)
    if count is None:
        count = info["nz"] - first
    out = np.empty((count, info["ny"], info["nx"]), np.float32)
    check(lib().atr_mrc_read_sections(_path(path), int(first), int(count), out.ctypes.data))
    return out


def write_mrc(path, stack, pixel_size=1.0, with_stats=False):
    """Float32 MRC of stack [nz][ny][nx] in the reference volume layout."""
    a = np.ascontiguousarray(stack, np.float32)
    if a.ndim == 2:
        a = a[None]
    nz, ny, nx = a.shape
    check(lib().atr_mrc_write(_path(path), a.ctypes.data, nx, ny, nz, float(pixel_size),
                              1 if with_stats else 0))


def volume_suffix(nth_vol=0):
    s = lib().atr_volume_suffix(int(nth_vol))
    if s is None:
        raise ValueError(f"no volume suffix for index {nth_vol}")
    return s.decode()


# ---- .aln ------------------------------------------------------------------
@dataclass
class Alignment:
    """Contents of a .aln file (CAlignParam + CLocalAlignParam + CDarkFrames)."""
    raw_size: tuple
    sec: np.ndarray            # [n] 1-based MRC section of each aligned tilt
    tilt_axis: np.ndarray      # [n] ROT
    shifts: np.ndarray         # [n][2] TX, TY
    tilts: np.ndarray          # [n] TILT
    local: np.ndarray = None   # [n][5][n_patches] or None
    dark_frame: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    dark_sec: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    dark_tilt: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    alpha_offset: float = 0.0
    beta_offset: float = 0.0
    thickness: int = 0

    @property
    def n_tilts(self):
        return len(self.tilts)

    @property
    def n_patches(self):
        return 0 if self.local is None else self.local.shape[2]


def read_aln(path):
    h = ctypes.c_void_p()
    check(lib().atr_aln_read(_path(path), ctypes.byref(h)))
    try:
        info = L.AlnInfo()
        check(lib().atr_aln_get_info(h, ctypes.byref(info)))
        n, np_, nd = info.n_tilts, info.n_patches, info.n_dark
        sec = np.empty(n, np.int32)
        axis = np.empty(n, np.float32)
        shifts = np.empty((n, 2), np.float32)
        tilts = np.empty(n, np.float32)
        check(lib().atr_aln_get_global(h, sec.ctypes.data, axis.ctypes.data, shifts.ctypes.data,
                                       tilts.ctypes.data))
        local = None
        if np_ > 0:
            local = np.empty((n, 5, np_), np.float32)
            check(lib().atr_aln_get_local(h, local.ctypes.data))
        df = np.empty(nd, np.int32)
        ds = np.empty(nd, np.int32)
        dt = np.empty(nd, np.float32)
        if nd:
            check(lib().atr_aln_get_dark(h, df.ctypes.data, ds.ctypes.data, dt.ctypes.data))
        return Alignment(tuple(info.raw_size), sec, axis, shifts, tilts, local, df, ds, dt,
                         info.alpha_offset, info.beta_offset, info.thickness)
    finally:
        lib().atr_aln_destroy(h)


def write_aln(path, aln):
    info = L.AlnInfo()
    info.raw_size[:] = [int(v) for v in aln.raw_size]
    info.n_tilts, info.n_patches, info.n_dark = aln.n_tilts, aln.n_patches, len(aln.dark_frame)
    info.alpha_offset, info.beta_offset = aln.alpha_offset, aln.beta_offset
    info.thickness = int(aln.thickness)
    keep = [np.ascontiguousarray(aln.sec, np.int32), np.ascontiguousarray(aln.tilt_axis, np.float32),
            np.ascontiguousarray(aln.shifts, np.float32), np.ascontiguousarray(aln.tilts, np.float32)]
    loc = None if aln.local is None else np.ascontiguousarray(aln.local, np.float32)
    dark = [np.ascontiguousarray(aln.dark_frame, np.int32), np.ascontiguousarray(aln.dark_sec, np.int32),
            np.ascontiguousarray(aln.dark_tilt, np.float32)]
    check(lib().atr_aln_write(_path(path), ctypes.byref(info), *[k.ctypes.data for k in keep],
                              None if loc is None else loc.ctypes.data,
                              *[(d.ctypes.data if len(d) else None) for d in dark]))


# ---- tilt-angle file ---------------------------------------------------------
def read_tilt_file(path, n_sections):
    """(tilts, acquisition indices, doses) of the first n_sections angle lines."""
    t = np.zeros(n_sections, np.float32)
    a = np.zeros(n_sections, np.int32)
    d = np.zeros(n_sections, np.float32)
    n = ctypes.c_int()
    check(lib().atr_tilt_file_read(_path(path), int(n_sections), t.ctypes.data, a.ctypes.data,
                                   d.ctypes.data, ctypes.byref(n)))
    return t, a, d


def find_tilt_file(stem):
    """``<stem>_TLT.txt`` if present, else ``<stem>.rawtlt`` (CTsPackage.cpp:360-364)."""
    for suffix in ("_TLT.txt", ".rawtlt"):
        if os.path.exists(os.fspath(stem) + suffix):
            return os.fspath(stem) + suffix
    raise FileNotFoundError(f"no {stem}_TLT.txt or {stem}.rawtlt")


# ---- reconstruct from files (-Cmd 2) -----------------------------------------
def frame_index_from_tilt(tilts, tilt):
    """CAlignParam::GetFrameIdxFromTilt (AreTomo/MrcUtil/CAlignParam.cpp:177-188):
    first frame of smallest |tilt - t| (float arithmetic)."""
    d = np.abs(np.float32(tilt) - np.asarray(tilts, np.float32))
    return int(np.argmin(d))


def binned_vol_z(vol_z, thickness, ext_z=300, at_bin=1.0):
    """CAreTomoMain::mRecon (AreTomo/CAreTomoMain.cpp:734-738): VolZ from -VolZ
    or thickness + ExtZ, divided by the binning, even, at least 100."""
    z = vol_z if vol_z is not None and vol_z >= 0 else thickness + ext_z
    z = int(np.float32(z) / np.float32(at_bin)) // 2 * 2
    return max(z, 100)


def second_vol_z(vol_z, thickness, ext_z, at_bin):
    """CAreTomoMain::mRecon2nd (AreTomo/CAreTomoMain.cpp:775-790): VolZ from
    -VolZ or thickness + ExtZ, at least 100, then divided by the binning and
    made even (no lower bound after binning; at most 16 skips the volume)."""
    z = vol_z if vol_z is not None and vol_z > 0 else thickness + ext_z
    if z <= 100:
        z = 100
    return int(np.float32(z) / np.float32(at_bin)) // 2 * 2


def _recon(proj, aln, vol_z, at_bin, wbp, sart, recon_range, layout, storage, fill_seed):
    n, ry, rx = proj.shape
    kw = dict(tilt_axis=aln.tilt_axis, shifts=aln.shifts, local_params=aln.local, bin=at_bin,
              layout=layout, storage=storage, fill_seed=fill_seed)
    if wbp:
        plan = Plan(rx, ry, aln.tilts, vol_z, L.METHOD_WBP, **kw)
    else:
        # CAreTomoMain::mSartRecon (:813-826): tilt range from -ReconRange,
        # subsets = tilts / (-Sart second value)
        start = frame_index_from_tilt(aln.tilts, recon_range[0])
        end = frame_index_from_tilt(aln.tilts, recon_range[1])
        if start == end:
            raise L.AtrError(L.ATR_ERR_INVALID_ARG, "empty SART tilt range")
        n_sart = end - start + 1
        subsets = max(1, n_sart // int(sart[1]))
        plan = Plan(rx, ry, aln.tilts, vol_z, L.METHOD_SART, sart_iters=int(sart[0]),
                    sart_subsets=subsets, sart_start_tilt=start, sart_n_tilts=n_sart, **kw)
    try:
        return plan.execute_host(proj), plan.vol_y
    finally:
        plan.close()


def _load_selected(mrc_path, aln):
    info = mrc_info(mrc_path)
    if info["nz"] != aln.raw_size[2]:
        raise L.AtrError(L.ATR_ERR_INVALID_ARG,
                         f"{mrc_path} has {info['nz']} sections, the .aln expects {aln.raw_size[2]}")
    raw = read_mrc(mrc_path)
    return np.ascontiguousarray(raw[aln.sec - 1]), info


def reconstruct_all(mrc_path, aln_path, out_stem, vol_z=None, ext_z=300, at_bin=(1.0, 0.0, 0.0),
                    wbp=False, sart=(20, 3), recon_range=(-90.0, 90.0), flip_vol=0,
                    pixel_size=None, storage=L.STORAGE_F32, fill_seed=0, evn_mrc=None,
                    odd_mrc=None, write=True):
    """Every volume AreTomo3 ``-Cmd 2`` writes from a raw series and its .aln.

    * ``_Vol``: the series at ``at_bin[0]`` by WBP or SART (``wbp``),
      CAreTomoMain::mRecon (AreTomo/CAreTomoMain.cpp:725-761);
    * ``_EVN_Vol`` / ``_ODD_Vol``: the even / odd frame-sum series
      (``evn_mrc``, ``odd_mrc``, -SplitSum) with the same alignment and method;
    * ``_2ND_Vol``: WBP at ``at_bin[1]`` and ``_3RD_Vol``: SART at
      ``at_bin[2]`` when those are >= 1, CAreTomoMain::mRecon2nd (:763-797).

    Images are taken by the .aln section indices, so dark frames listed there
    are left out (CRemoveDarkFrames).  The pre-stage applies the per-tilt tilt
    axis, shifts and local patch shifts, -AtBin Fourier-bins, then the
    reconstruction runs; volumes are written as ``out_stem + suffix`` in the
    -FlipVol layout with the binned pixel size.  Returns {suffix index:
    (volume, path or None)} (0 _Vol, 1 _EVN_Vol, 2 _ODD_Vol, 3 _2ND_Vol,
    4 _3RD_Vol)."""
    aln = read_aln(aln_path)
    if mrc_info(mrc_path)["nz"] <= 5:
        # CAreTomoMain::mCheckTiltSeries (AreTomo/CAreTomoMain.cpp:986-995)
        raise L.AtrError(L.ATR_ERR_INVALID_ARG, f"{mrc_path}: too few tilt images (<= 5)")
    bins = tuple(at_bin) if np.ndim(at_bin) else (float(at_bin), 0.0, 0.0)
    bins = tuple(bins) + (0.0,) * (3 - len(bins))
    layout = {0: L.LAYOUT_XZY, 1: L.LAYOUT_XYZ, 2: L.LAYOUT_XYZ_NOFLIP}[int(flip_vol)]
    proj, info = _load_selected(mrc_path, aln)
    n, ry, rx = proj.shape
    pix0 = info["pixel_size"] if pixel_size is None else float(pixel_size)
    _, ay = aligned_size(rx, ry, aln.tilt_axis[0])
    jobs = []  # (suffix index, projections, vol_z, bin, wbp)
    z = binned_vol_z(vol_z, aln.thickness, ext_z, bins[0])
    jobs.append((0, proj, z, bins[0], wbp))
    for idx, path in ((1, evn_mrc), (2, odd_mrc)):
        if path is not None:
            jobs.append((idx, _load_selected(path, aln)[0], z, bins[0], wbp))
    for idx, b, use_wbp in ((3, bins[1], True), (4, bins[2], False)):
        if b >= 1.0:
            z2 = second_vol_z(vol_z, aln.thickness, ext_z, b)
            if z2 > 16:
                jobs.append((idx, proj, z2, b, use_wbp))
    out = {}
    for idx, p, zz, b, use_wbp in jobs:
        vol, vol_y = _recon(p, aln, zz, b, use_wbp, sart, recon_range, layout, storage, fill_seed)
        path = None
        if write:
            pix = pix0
            if b != 1.0:
                # CBinStack::DoFFT (AreTomo/Correct/CBinStack.cpp:66-67):
                # (pixel / binned height) * aligned height, in float
                pix = float(np.float32(np.float32(pix0) / np.float32(vol_y)) * np.float32(ay))
            path = os.fspath(out_stem) + volume_suffix(idx)
            write_mrc(path, vol, pix)
        out[idx] = (vol, path)
    return out


def reconstruct_files(mrc_path, aln_path, out_stem, vol_z=None, ext_z=300, at_bin=1.0, wbp=False,
                      sart=(20, 3), recon_range=(-90.0, 90.0), flip_vol=0, pixel_size=None,
                      storage=L.STORAGE_F32, fill_seed=0, write=True):
    """The main volume of :func:`reconstruct_all` (``_Vol.mrc``):
    returns (volume, path or None)."""
    res = reconstruct_all(mrc_path, aln_path, out_stem, vol_z=vol_z, ext_z=ext_z,
                          at_bin=(float(at_bin), 0.0, 0.0), wbp=wbp, sart=sart,
                          recon_range=recon_range, flip_vol=flip_vol, pixel_size=pixel_size,
                          storage=storage, fill_seed=fill_seed, write=write)
    return res[0]


def aligned_size(raw_x, raw_y, tilt_axis_deg):
    """Size of the aligned image for a raw image and tilt axis
    (CCorrectUtil::CalcAlignedSize, AreTomo/Correct/CCorrectUtil.cpp:11-21)."""
    ax, ay = ctypes.c_int(), ctypes.c_int()
    check(lib().atr_aligned_size(int(raw_x), int(raw_y), float(tilt_axis_deg), ctypes.byref(ax),
                                 ctypes.byref(ay)))
    return ax.value, ay.value
