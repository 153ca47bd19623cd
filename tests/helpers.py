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


def download_frames(dev, frame_shape, frames):
    """Frames `frames` of a device array of [k][*frame_shape] float32 frames,
    without copying the rest to the host."""
    import ctypes

    from aretomo3_amd import _lib as L
    per = int(np.prod(frame_shape))
    out = np.empty((len(frames),) + tuple(frame_shape), np.float32)
    for i, f in enumerate(frames):
        L.check(L.lib().atr_memcpy_d2h(out[i].ctypes.data, dev.ptr + 4 * per * int(f), 4 * per))
    return out


def fourier_ring_correlation(a, b, n_rings=32):
    """Fourier ring correlation of two stacks of XZ slices [k][z][x], summed
    over the slices, per ring of normalised radius (ring n_rings - 1 ends at
    Nyquist)."""
    fa = np.fft.rfft2(np.asarray(a, np.float64))
    fb = np.fft.rfft2(np.asarray(b, np.float64))
    nz, nx = fa.shape[1], a.shape[2]
    kz = np.fft.fftfreq(nz)[:, None]
    kx = np.fft.rfftfreq(nx)[None, :]
    r = np.sqrt(kz ** 2 + kx ** 2) / 0.5
    ring = np.minimum((r * n_rings).astype(int), n_rings).ravel()
    num = np.zeros(n_rings + 1)
    da = np.zeros(n_rings + 1)
    db = np.zeros(n_rings + 1)
    for k in range(fa.shape[0]):
        num += np.bincount(ring, (fa[k] * np.conj(fb[k])).real.ravel(), n_rings + 1)
        da += np.bincount(ring, (np.abs(fa[k]) ** 2).ravel(), n_rings + 1)
        db += np.bincount(ring, (np.abs(fb[k]) ** 2).ravel(), n_rings + 1)
    return (num / np.sqrt(np.maximum(da * db, 1e-300)))[:n_rings]
