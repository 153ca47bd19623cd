"""The headline SART series at its full size: 1024 x 1024 raw projections, 61
tilts -60..60 step 2, tilt axis 85 deg with per-tilt shifts (so the
pre-stage runs), volume 1024 x 1024 x 256, 20 iterations x 5 interleaved
subsets, relaxation 0.3, positivity off, FP32 storage.

* Sampled y-slices against the FP64 oracle (relative L2 <= 1e-5).
* Linearity over the whole volume: with positivity off and every empty
  pixel filled from the same seeded index chain, the reconstruction is a
  linear map of the raw projections, so R(a p + b q) = a R(p) + b R(q) up to
  FP32 rounding (relative L2 <= 1e-5); a dropped or doubled term anywhere in
  the volume breaks it.
"""
import numpy as np
import pytest

import oracle
from aretomo3_amd import Plan, phantom
from aretomo3_amd import _lib as L
from helpers import rel_l2

pytestmark = pytest.mark.gpu

PX, VOL_Z, ITERS, SUBSETS, RELAX, SEED = 1024, 256, 20, 5, 0.3, 20261203


def series(seed, n_ell):
    tilts = phantom.tilt_range(-60.0, 60.0, 2.0)
    raw, shifts = phantom.tilt_series(seed, PX, PX, tilts, VOL_Z, n_ellipsoids=n_ell,
                                      axis_min=4.0, axis_max=40.0, tilt_axis=85.0, max_shift=3.0,
                                      noise_frac=0.10, n_threads=16)
    return tilts, raw, shifts


def plan_for(tilts, shifts):
    axes = np.full(len(tilts), 85.0, np.float32)
    return axes, Plan(PX, PX, tilts, VOL_Z, L.METHOD_SART, tilt_axis=axes, shifts=shifts,
                      sart_iters=ITERS, sart_subsets=SUBSETS, relax=RELAX, positivity=False,
                      subset_order=L.SUBSETS_INTERLEAVED, fill_seed=SEED)


def test_headline_sart_sampled_slices(gpu_available):
    tilts, raw, shifts = series(SEED, 200)
    axes, plan = plan_for(tilts, shifts)
    print("case full-size SART sampled slices", flush=True)
    vol = plan.execute_host(raw)
    plan.close()
    assert vol.shape == (PX, VOL_Z, PX)
    sl = [0, 341, 682, 1023]
    aligned = oracle.prestage(raw, axes, shifts, 1, SEED)
    ref = oracle.recon_slices(aligned, tilts, 1, VOL_Z, sl, iters=ITERS, n_subsets=SUBSETS,
                              relax=RELAX, order=1, positivity=0, seed=SEED)
    err = rel_l2(vol[sl], ref)
    print("rel_l2", err, flush=True)
    assert err <= 1e-5


def test_headline_sart_linearity(gpu_available):
    tilts, p, shifts = series(SEED, 200)
    _, q, _ = series(SEED + 1, 120)
    a, b = np.float32(0.75), np.float32(-1.5)
    mix = (a * p + b * q).astype(np.float32)
    _, plan = plan_for(tilts, shifts)
    print("case full-size SART linearity", flush=True)
    rp = plan.execute_host(p)
    rq = plan.execute_host(q)
    rm = plan.execute_host(mix)
    plan.close()
    want = a * rp.astype(np.float64) + b * rq.astype(np.float64)
    err = rel_l2(rm, want)
    print("rel_l2", err, flush=True)
    assert err <= 1e-5
