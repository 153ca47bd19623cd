"""Shared helpers for the tests: synthetic series and error norms."""
import numpy as np

from aretomo3_amd import phantom


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))


def small_series(seed=7, px=64, py=12, tilts=None, vol_z=32, n_ell=30, noise=0.05,
                 axis=float("nan"), shift=0.0, amin=3.0, amax=10.0):
    if tilts is None:
        tilts = phantom.tilt_range(-60, 60, 6)
    proj, shifts = phantom.tilt_series(seed, px, py, tilts, vol_z, n_ellipsoids=n_ell,
                                       axis_min=amin, axis_max=amax, tilt_axis=axis,
                                       max_shift=shift, noise_frac=noise)
    return np.asarray(tilts, np.float32), proj, shifts
