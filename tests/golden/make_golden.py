#!/usr/bin/env python3
"""Writes the small golden fixtures of tests/golden/ (run from the repo root).

The reference holds no golden vectors for its reconstruction path, so these
fixtures are the FP64 oracle's outputs on seeded inputs (the oracle itself is
pinned by tests/test_oracle.py).  Inputs: a seeded ellipsoid phantom mapped to
raw images with tilt axis 85 deg and per-tilt shifts (the pre-stage runs).

  series_small.npz
    raw [11][40][32] f32, tilts [11], axes [11], shifts [11][2], seed, vol_z,
    sart_xzy [32][12][40]  SART 3 iterations x 3 contiguous subsets, reference
                           relax (1/3) and positivity, XZY layout (z flipped)
    sart_il_xzy             same, interleaved subsets, relax 0.3, positivity off
    wbp_xzy   [32][12][40]  WBP
    aligned   [11][32][40]  pre-stage output with random fill
Expected volumes are stored as float32 (the oracle computes in FP64).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from aretomo3_amd import phantom  # noqa: E402


def series_small():
    seed, vol_z = 20261210, 12
    tilts = phantom.tilt_range(-60.0, 60.0, 12.0)
    n = len(tilts)
    raw, shifts = phantom.tilt_series(seed, 32, 40, tilts, vol_z, n_ellipsoids=12,
                                      axis_min=3.0, axis_max=8.0, tilt_axis=85.0,
                                      max_shift=2.0, noise_frac=0.05, n_threads=1)
    axes = np.full(n, 85.0, np.float32)
    aligned = oracle.prestage(raw, axes, shifts, 1, seed)
    ny = aligned.shape[1]
    slices = np.arange(ny, dtype=np.int32)
    sart = oracle.recon_slices(aligned, tilts, 1, vol_z, slices, iters=3, n_subsets=3,
                               relax=-1.0, order=0, positivity=1, seed=seed)
    sart_il = oracle.recon_slices(aligned, tilts, 1, vol_z, slices, iters=3, n_subsets=3,
                                  relax=0.3, order=1, positivity=0, seed=seed)
    wbp = oracle.recon_slices(aligned, tilts, 0, vol_z, slices)
    np.savez_compressed(os.path.join(HERE, "series_small.npz"), raw=raw, tilts=tilts,
                        axes=axes, shifts=shifts, seed=np.uint64(seed), vol_z=vol_z,
                        aligned=aligned.astype(np.float32),
                        sart_xzy=sart.astype(np.float32), sart_il_xzy=sart_il.astype(np.float32),
                        wbp_xzy=wbp.astype(np.float32))


if __name__ == "__main__":
    series_small()
