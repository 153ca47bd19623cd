"""The re-projector's no-allocation entry point (atr_reproj_execute_noalloc)
against the handle entry point (atr_reproj_execute), through the C ABI: the
same kernels over a different slab split, so the re-projections must agree
bit for bit.  Two calls in a row share the library's fixed scratch array,
the second with a smaller series; a slab row larger than the scratch is
refused before any device work."""
import ctypes

import numpy as np
import pytest

from aretomo3_amd import DeviceArray, phantom, synchronize
from aretomo3_amd import _lib as L
from aretomo3_amd.projalign import CalcReproj

pytestmark = pytest.mark.gpu


def series(seed, px, py, step, vol_z):
    tilts = phantom.tilt_range(-60, 60, step).astype(np.float32)
    projs, _ = phantom.tilt_series(seed, px, py, tilts, vol_z, n_ellipsoids=25, axis_min=3,
                                   axis_max=12, noise_frac=0.05)
    projs[2, :, : px // 6] = -1e30          # an empty band in one projection
    return tilts, np.ascontiguousarray(projs, np.float32)


def noalloc(projs, tilts, skip, idx, vol_z):
    n, py, px = projs.shape
    d_in = DeviceArray.from_numpy(projs)
    d_out = DeviceArray(4 * px * py)
    try:
        L.check(L.lib().atr_reproj_execute_noalloc(px, py, n, vol_z, d_in.ptr, tilts.ctypes.data,
                                                   skip.ctypes.data, idx, d_out.ptr, None))
        synchronize()
        return d_out.to_numpy((py, px))
    finally:
        d_in.free()
        d_out.free()


@pytest.mark.parametrize("px,py,step,vol_z,idx", [
    # one slab row is 2 * 1024 * 2048 floats = 16 MiB: the scratch holds 4
    # rows, so the 13 rows run as 4 slabs (the handle's buffer takes all 13)
    (1024, 13, 3.0, 2048, 30),
    # then a smaller series: 1 slab
    (256, 8, 6.0, 96, 5),
])
def test_noalloc_matches_handle_entry_point(gpu_available, px, py, step, vol_z, idx):
    tilts, projs = series(20261203 + px, px, py, step, vol_z)
    n = len(tilts)
    skip = (np.abs(tilts) >= abs(tilts[idx]) - 1e-6).astype(np.uint8)
    skip[idx] = 1
    print(f"case reproj noalloc px={px} py={py} n={n} vz={vol_z} idx={idx}", flush=True)
    rp = CalcReproj(px, py, n, vol_z)
    want = rp.do_it(projs, tilts, skip, idx)
    rp.close()
    for rep in range(2):   # the second call finds the scratch as the first left it
        got = noalloc(projs, tilts, skip, idx, vol_z)
        np.testing.assert_array_equal(got, want)
    assert (want > -1e29).any()


def test_noalloc_refuses_a_slab_row_beyond_the_scratch(gpu_available):
    tilts = np.linspace(-60, 60, 5).astype(np.float32)
    skip = np.zeros(5, np.uint8)
    d_out = DeviceArray(4 * 4096)
    try:
        # 2 * 4096 * 2049 floats > 64 MiB
        st = L.lib().atr_reproj_execute_noalloc(4096, 1, 5, 2049, d_out.ptr, tilts.ctypes.data,
                                                skip.ctypes.data, 2, d_out.ptr, None)
        assert st == L.ATR_ERR_INVALID_ARG
        assert b"64 MiB" in L.lib().atr_last_error_string()
    finally:
        d_out.free()
