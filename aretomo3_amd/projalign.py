"""The ProjAlign re-projector (AreTomo/ProjAlign/CCalcReproj.cpp, GReproj.cu)
over ``atr_reproj_*``: predicts one projection of a tilt series from its
lower-tilt neighbours, as the projection-matching alignment does.
"""
import ctypes

import numpy as np

from . import _lib as L
from ._lib import check, lib
from .recon import DeviceArray, synchronize


def find_proj_range(tilts, proj_idx, skip=None):
    """CCalcReproj::mFindProjRange: (first, last) reference tilt indices."""
    t = np.ascontiguousarray(tilts, np.float32)
    s = None if skip is None else np.ascontiguousarray(skip, np.uint8)
    a, b = ctypes.c_int(), ctypes.c_int()
    check(lib().atr_reproj_find_range(t.ctypes.data, None if s is None else s.ctypes.data, len(t),
                                      int(proj_idx), ctypes.byref(a), ctypes.byref(b)))
    return a.value, b.value


class CalcReproj:
    """CCalcReproj: Setup(stack size, VolZ) then DoIt per projection."""

    def __init__(self, proj_x, proj_y, n_tilts, vol_z):
        h = ctypes.c_void_p()
        check(lib().atr_reproj_create(int(proj_x), int(proj_y), int(n_tilts), int(vol_z),
                                      ctypes.byref(h)))
        self.handle = h.value
        self.proj_x, self.proj_y, self.n_tilts, self.vol_z = proj_x, proj_y, n_tilts, vol_z

    def execute(self, proj_dev, tilts, skip, proj_idx, out_dev, stream=None):
        """Stream-ordered, device buffers [n][y][x] in, [y][x] out."""
        t = np.ascontiguousarray(tilts, np.float32)
        s = None if skip is None else np.ascontiguousarray(skip, np.uint8)
        pp = proj_dev.ptr if isinstance(proj_dev, DeviceArray) else proj_dev
        op = out_dev.ptr if isinstance(out_dev, DeviceArray) else out_dev
        check(lib().atr_reproj_execute(self.handle, pp, t.ctypes.data,
                                       None if s is None else s.ctypes.data, int(proj_idx), op,
                                       stream))

    def do_it(self, projs, tilts, skip, proj_idx):
        """Host stack [n][y][x] in, host re-projection [y][x] out."""
        projs = np.ascontiguousarray(projs, np.float32)
        assert projs.shape == (self.n_tilts, self.proj_y, self.proj_x)
        d_in = DeviceArray.from_numpy(projs)
        d_out = DeviceArray(4 * self.proj_x * self.proj_y)
        try:
            self.execute(d_in, tilts, skip, proj_idx, d_out)
            synchronize()
            return d_out.to_numpy((self.proj_y, self.proj_x))
        finally:
            d_in.free()
            d_out.free()

    def close(self):
        if getattr(self, "handle", None):
            lib().atr_reproj_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()
