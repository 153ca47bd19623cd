"""Seeded synthetic tilt series (atr_phantom_tilt_series in libatr_recon.so).

Ellipsoid phantoms with analytic line integrals; used to make the synthetic
inputs of the benchmark and of the tests.
"""
import ctypes
import os

import numpy as np

from ._lib import check, lib


def tilt_range(lo, hi, step):
    """Tilt angles lo, lo+step, ..., hi (inclusive) as float32."""
    n = int(round((hi - lo) / step)) + 1
    return (lo + step * np.arange(n)).astype(np.float32)


def tilt_series(seed, proj_x, proj_y, tilts, vol_z, n_ellipsoids=200, axis_min=4.0,
                axis_max=40.0, tilt_axis=float("nan"), max_shift=0.0, noise_frac=0.0,
                n_threads=0):
    """Returns (projections [n][proj_y][proj_x] float32, shifts [n][2] float32).

    tilt_axis NaN: ideal aligned projections (no pre-stage needed).
    Otherwise: raw images that the pre-stage maps back onto the ideal ones,
    with per-tilt shifts drawn uniformly in [-max_shift, max_shift]."""
    tilts = np.ascontiguousarray(tilts, np.float32)
    n = len(tilts)
    out = np.empty((n, proj_y, proj_x), np.float32)
    shifts = np.zeros((n, 2), np.float32)
    if n_threads <= 0:
        n_threads = min(16, os.cpu_count() or 1)
    check(lib().atr_phantom_tilt_series(
        int(seed), int(proj_x), int(proj_y), n, ctypes.c_void_p(tilts.ctypes.data), int(vol_z),
        int(n_ellipsoids), float(axis_min), float(axis_max), float(tilt_axis), float(max_shift),
        float(noise_frac), ctypes.c_void_p(shifts.ctypes.data), ctypes.c_void_p(out.ctypes.data),
        int(n_threads)))
    return out, shifts
