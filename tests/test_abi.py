"""The C ABI of libatr_recon.so without a GPU: it loads, exports every symbol
include/aretomo_recon.h declares, and validates descriptors on the host."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from aretomo3_amd import _lib as L
from aretomo3_amd import phantom

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("aretomo_recon.h", "aretomo_files.h")]


def header_functions():
    found = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        found.update(re.findall(r"\b(atr_[a-z0-9_]+)\s*\(", text))
    return sorted(found)


def test_header_matches_binding_and_library():
    declared = header_functions()
    assert declared == sorted(L.EXPORTS)
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], check=True,
                         capture_output=True, text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = [s for s in declared if s not in exported]
    assert not missing, missing
    lib = L.lib()
    for s in declared:
        assert getattr(lib, s) is not None


def test_library_is_built_for_gfx950():
    out = subprocess.run(["/opt/rocm/llvm/bin/clang-offload-bundler", "--list", "--type=o",
                          "--input=" + L.LIB_PATH], capture_output=True, text=True)
    if out.returncode == 0 and out.stdout.strip():
        assert "gfx950" in out.stdout
    else:   # fall back to the embedded code object names
        data = open(L.LIB_PATH, "rb").read()
        assert b"gfx950" in data
    assert b"gfx950" in L.lib().atr_version()


def make_desc(**kw):
    d = L.SeriesDesc()
    L.lib().atr_series_desc_init(ctypes.byref(d))
    keep = []
    tilts = np.ascontiguousarray(kw.pop("tilts", np.linspace(-60, 60, 21)), np.float32)
    keep.append(tilts)
    d.tilt_deg = tilts.ctypes.data
    d.n_tilts = len(tilts)
    d.proj_x, d.proj_y, d.vol_z = 64, 16, 32
    for k, v in kw.items():
        if isinstance(v, np.ndarray):
            keep.append(v)
            v = v.ctypes.data
        setattr(d, k, v)
    return d, keep


def estimate(d):
    b = ctypes.c_size_t(0)
    st = L.lib().atr_estimate_device_bytes(ctypes.byref(d), ctypes.byref(b))
    return st, b.value


def test_desc_defaults_are_the_reference_defaults():
    d = L.SeriesDesc()
    L.lib().atr_series_desc_init(ctypes.byref(d))
    assert d.method == L.METHOD_SART
    assert d.sart_iters == 20 and d.sart_subsets >= 1
    assert d.positivity == 1 and d.subset_order == L.SUBSETS_CONTIGUOUS
    assert d.sart_relax <= 0          # 1 / subsets
    assert d.storage == L.STORAGE_F32 and d.out_layout == L.LAYOUT_XZY
    assert d.random_fill == 1
    assert not d.tilt_axis_deg and not d.shift_xy and not d.exclude


def test_estimate_valid_and_monotone():
    d, keep = make_desc()
    st, b32 = estimate(d)
    assert st == L.ATR_OK and b32 > 0
    d.vol_z = 64
    st, b64 = estimate(d)
    assert st == L.ATR_OK and b64 > b32
    d.method = L.METHOD_WBP
    st, bw = estimate(d)
    assert st == L.ATR_OK and bw > 0
    d.method = L.METHOD_SART
    d.storage = L.STORAGE_F16
    st, bh = estimate(d)
    assert st == L.ATR_OK and bh < b64


@pytest.mark.parametrize("field,value,status", [
    ("vol_z", 1, 1),
    ("proj_x", 2, 1),
    ("n_tilts", 0, 1),
    ("method", 7, 1),
    ("storage", 9, 1),
    ("out_layout", 5, 1),
    ("subset_order", 3, 1),
    ("sart_subsets", 50, 1),
    ("sart_iters", -1, 1),
    ("sart_start_tilt", 30, 1),
    ("tilt_deg", None, 1),
])
def test_estimate_rejects_bad_descriptors(field, value, status):
    d, keep = make_desc()
    setattr(d, field, value)
    st, _ = estimate(d)
    assert st == status
    assert L.lib().atr_last_error_string()


def test_rejects_non_finite_angles_and_small_strides():
    t = np.linspace(-60, 60, 21).astype(np.float32)
    t[3] = np.nan
    d, keep = make_desc(tilts=t)
    assert estimate(d)[0] == L.ATR_ERR_INVALID_ARG
    axes = np.full(21, 85.0, np.float32)
    axes[5] = np.inf
    d, keep = make_desc(tilt_axis_deg=axes)
    assert estimate(d)[0] == L.ATR_ERR_INVALID_ARG
    d, keep = make_desc(row_stride=32)
    assert estimate(d)[0] == L.ATR_ERR_INVALID_ARG


def test_wbp_size_limits():
    d, keep = make_desc(method=L.METHOD_WBP)
    d.proj_x = 4 * 7 * 11     # prime factors 7, 11: direct-DFT FFT stages
    assert estimate(d)[0] == L.ATR_OK
    d.proj_x = 600
    assert estimate(d)[0] == L.ATR_OK
    d.proj_x = 4100           # ProjX above 4096
    assert estimate(d)[0] == L.ATR_ERR_UNSUPPORTED_SIZE


def test_local_and_bin_validation():
    d, keep = make_desc()
    lp = np.zeros((d.n_tilts, 5, 4), np.float32)
    d.n_patches, d.local_params = 4, lp.ctypes.data
    assert estimate(d)[0] == L.ATR_ERR_INVALID_ARG          # no pre-stage
    axes = np.full(d.n_tilts, 85.0, np.float32)
    d.tilt_axis_deg = axes.ctypes.data
    assert estimate(d)[0] == L.ATR_OK
    lp[1, 2, 3] = np.nan
    assert estimate(d)[0] == L.ATR_ERR_INVALID_ARG
    lp[1, 2, 3] = 0.0
    d.local_params = None
    assert estimate(d)[0] == L.ATR_ERR_INVALID_ARG
    d.n_patches = 0
    for b, want in ((1.0, L.ATR_OK), (2.0, L.ATR_OK), (1.5, L.ATR_OK), (0.5, L.ATR_ERR_INVALID_ARG),
                    (float("nan"), L.ATR_ERR_INVALID_ARG), (1e6, L.ATR_ERR_INVALID_ARG)):
        d.bin = b
        assert estimate(d)[0] == want, b


def test_null_arguments_fail_cleanly():
    lib = L.lib()
    assert lib.atr_plan_create(None, None) == L.ATR_ERR_INVALID_ARG
    assert lib.atr_execute(None, None, None, 0, None) == L.ATR_ERR_INVALID_ARG
    assert lib.atr_plan_get_volume_size(None, None, None, None) == L.ATR_ERR_INVALID_ARG
    assert lib.atr_plan_device_bytes(None) == 0
    lib.atr_plan_destroy(None)
    d, keep = make_desc()
    assert lib.atr_estimate_device_bytes(ctypes.byref(d), None) == L.ATR_ERR_INVALID_ARG
    assert b"NULL" in lib.atr_last_error_string()


def test_plan_create_validates_before_touching_a_device():
    d, keep = make_desc(vol_z=1)
    h = ctypes.c_void_p()
    assert L.lib().atr_plan_create(ctypes.byref(d), ctypes.byref(h)) == L.ATR_ERR_INVALID_ARG
    assert not h.value


def test_phantom_is_deterministic_and_thread_independent():
    t = phantom.tilt_range(-60, 60, 30)
    a, sa = phantom.tilt_series(3, 40, 8, t, 16, n_ellipsoids=10, axis_min=2, axis_max=6,
                                tilt_axis=85.0, max_shift=2.0, noise_frac=0.1, n_threads=1)
    b, sb = phantom.tilt_series(3, 40, 8, t, 16, n_ellipsoids=10, axis_min=2, axis_max=6,
                                tilt_axis=85.0, max_shift=2.0, noise_frac=0.1, n_threads=4)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(sa, sb)
    assert np.abs(sa).max() <= 2.0 and np.abs(sa).max() > 0
    c, _ = phantom.tilt_series(4, 40, 8, t, 16, n_ellipsoids=10, axis_min=2, axis_max=6)
    assert not np.array_equal(a, c)
    assert np.all(np.isfinite(a))


def test_phantom_line_integral_of_one_ball():
    # a ball of radius r and density rho projects to 2 rho sqrt(r^2 - d^2) at
    # every tilt; with one ellipsoid the peak value is bounded by 2 * r_max * 1.0
    t = phantom.tilt_range(-60, 60, 60)
    p, _ = phantom.tilt_series(9, 64, 64, t, 32, n_ellipsoids=1, axis_min=5, axis_max=5)
    assert p.max() <= 2 * 5 * 1.0 + 1e-3
    assert p.max() >= 2 * 5 * 0.5 - 1e-3
    assert p.min() >= 0.0


def band_candidates(vol_x, vol_z):
    rows = (ctypes.c_int * 8)()
    n = ctypes.c_int()
    L.check(L.lib().atr_fp_band_candidates(vol_x, vol_z, rows, 8, ctypes.byref(n)))
    return list(rows[:n.value])


# largest band height whose bz + 1 rows of VolX + 1 float2 pairs fit 150 KiB,
# capped at 17 (the forward projector's LDS budget), and at least 1
def lds_limit(vol_x):
    return min(17, max(1, (150 * 1024) // ((vol_x + 1) * 8) - 1))


@pytest.mark.parametrize("vol_x,vol_z", [(512, 128), (1024, 256), (2048, 512), (4096, 40),
                                         (6399, 64), (6400, 64), (8192, 256), (9599, 100)])
def test_autotune_band_candidates_fit_the_lds(vol_x, vol_z):
    c = band_candidates(vol_x, vol_z)
    assert 1 <= len(c) <= 3 and len(set(c)) == len(c)
    lim = lds_limit(vol_x)
    assert all(1 <= b <= lim for b in c), (c, lim)
    # every candidate's LDS fits the hardware's 160 KiB as well
    assert all((b + 1) * (vol_x + 1) * 8 <= 160 * 1024 for b in c)
    if vol_x >= 6400:
        assert c == [1]


def test_autotune_band_candidates_at_the_headline_size():
    # VolZ 256 has 255 cell rows: 17 rows split them into 15 equal bands;
    # 2/3 and 1/2 of 17 (11 and 8) need 24 and 32 bands, of at most 11 and 8
    assert band_candidates(1024, 256) == [17, 11, 8]
    with pytest.raises(L.AtrError):
        L.check(L.lib().atr_fp_band_candidates(1024, 1, None, 0, ctypes.byref(ctypes.c_int())))
