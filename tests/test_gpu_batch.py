"""Batched SART (atr_batch_*): uneven series in one persistent launch per
step give the volumes they give alone, with and without HIP-graph replay."""
import numpy as np
import pytest

import oracle
from aretomo3_amd import Batch, DeviceArray, Plan, phantom, synchronize
from aretomo3_amd import _lib as L
from helpers import rel_l2

pytestmark = pytest.mark.gpu

# (px, py, tilt step, vol_z, iters, subsets, tilt axis or None, storage)
SERIES = [
    (96, 6, 6.0, 32, 3, 3, None, L.STORAGE_F32),
    (64, 48, 4.0, 24, 2, 5, 85.0, L.STORAGE_F32),
    (130, 4, 10.0, 40, 4, 2, None, L.STORAGE_F32),
]


def make_series(i, spec):
    px, py, step, vz, iters, subsets, axis, storage = spec
    tilts = phantom.tilt_range(-60, 60, step)
    raw, shifts = phantom.tilt_series(100 + i, px, py, tilts, vz, n_ellipsoids=20, axis_min=3,
                                      axis_max=10, tilt_axis=float("nan") if axis is None else axis,
                                      max_shift=0.0 if axis is None else 2.0, noise_frac=0.05)
    kw = dict(sart_iters=iters, sart_subsets=subsets, relax=0.3,
              subset_order=L.SUBSETS_INTERLEAVED, positivity=False, storage=storage,
              fill_seed=7 + i)
    if axis is not None:
        kw.update(tilt_axis=np.full(len(tilts), axis, np.float32), shifts=shifts)
    plan = Plan(px, py, tilts, vz, L.METHOD_SART, **kw)
    return plan, raw, tilts, shifts


def run_alone(plans, raws):
    outs = []
    for plan, raw in zip(plans, raws):
        d_in = DeviceArray.from_numpy(raw)
        d_out = DeviceArray(4 * plan.vol_x * plan.vol_y * plan.vol_z)
        plan.execute(d_in, d_out)
        synchronize()
        outs.append(d_out.to_numpy(plan.volume_shape))
    return outs


def test_batch_matches_alone_and_oracle(gpu_available):
    built = [make_series(i, s) for i, s in enumerate(SERIES)]
    plans = [b[0] for b in built]
    raws = [b[1] for b in built]
    print("case batch vs alone", [p.volume_shape for p in plans], flush=True)
    alone = run_alone(plans, raws)
    batch = Batch(plans)
    d_in = [DeviceArray.from_numpy(r) for r in raws]
    d_out = [DeviceArray(4 * p.vol_x * p.vol_y * p.vol_z) for p in plans]
    batch.execute(d_in, d_out)
    synchronize()
    for i, (p, ref) in enumerate(zip(plans, alone)):
        got = d_out[i].to_numpy(p.volume_shape)
        err = rel_l2(got, ref)
        print("series", i, "batch vs alone rel_l2", err, "identical", np.array_equal(got, ref))
        assert err <= 1e-6
    # one series also against the FP64 oracle
    plan, raw, tilts, _ = built[0]
    px, py, step, vz, iters, subsets, axis, storage = SERIES[0]
    ref = oracle.recon_slices(raw.astype(np.float64), tilts, 1, vz, list(range(py)), iters=iters,
                              n_subsets=subsets, relax=0.3, order=1, positivity=0, seed=7)
    err = rel_l2(d_out[0].to_numpy(plan.volume_shape), ref)
    print("series 0 vs oracle rel_l2", err)
    assert err <= 1e-5
    batch.close()


def test_batch_graph_replays(gpu_available):
    built = [make_series(i, s) for i, s in enumerate(SERIES)]
    plans = [b[0] for b in built]
    raws = [b[1] for b in built]
    print("case batch graph replays", flush=True)
    d_in = [DeviceArray.from_numpy(r) for r in raws]
    eager = Batch(plans)
    d_ref = [DeviceArray(4 * p.vol_x * p.vol_y * p.vol_z) for p in plans]
    eager.execute(d_in, d_ref)
    synchronize()
    ref = [d.to_numpy(p.volume_shape) for d, p in zip(d_ref, plans)]
    eager.close()
    graphed = Batch(plans, graph=True)
    other, other_raw, _, _ = make_series(9, (64, 3, 6.0, 16, 2, 2, None, L.STORAGE_F32))
    d_other_in = DeviceArray.from_numpy(other_raw)
    d_other_out = DeviceArray(4 * other.vol_x * other.vol_y * other.vol_z)
    for rep in range(3):
        d_out = [DeviceArray(4 * p.vol_x * p.vol_y * p.vol_z) for p in plans]
        graphed.execute(d_in, d_out)
        synchronize()
        # eager work of the program between replays
        other.execute(d_other_in, d_other_out)
        synchronize()
        for i, p in enumerate(plans):
            got = d_out[i].to_numpy(p.volume_shape)
            err = rel_l2(got, ref[i])
            print("replay", rep, "series", i, "rel_l2", err)
            assert err <= 1e-6
    graphed.close()


def test_batch_rejects_mixed_plans(gpu_available):
    t = phantom.tilt_range(-60, 60, 10)
    a = Plan(32, 2, t, 8, L.METHOD_SART, sart_iters=1, sart_subsets=1)
    w = Plan(32, 2, t, 8, L.METHOD_WBP)
    h = Plan(32, 2, t, 8, L.METHOD_SART, sart_iters=1, sart_subsets=1, storage=L.STORAGE_F16)
    print("case batch rejects", flush=True)
    for bad in ([a, w], [a, h], [a, a]):
        with pytest.raises(L.AtrError):
            Batch(bad)
