"""Tile shapes and the autotuner: every back-projection shape (z rows per
thread x tilts per LDS stage) and forward-projection band height gives the
oracle's result; the autotuner returns one of them and its plan stays exact."""
import numpy as np
import pytest

import oracle
from aretomo3_amd import Batch, DeviceArray, Plan, phantom, synchronize
from aretomo3_amd import _lib as L
from helpers import rel_l2, small_series

pytestmark = pytest.mark.gpu
TOL32 = 1e-5


def _series():
    tilts, proj, _ = small_series(px=96, py=5, vol_z=40, tilts=phantom.tilt_range(-60, 60, 4))
    ref = oracle.recon_slices(proj.astype(np.float64), tilts, 1, 40, list(range(5)), iters=3,
                              n_subsets=4, relax=0.3, order=1, positivity=0)
    return tilts, proj, ref


@pytest.mark.parametrize("tz,chunk,rows", [(16, 16, 17), (16, 8, 9), (8, 16, 5), (8, 8, 2)])
def test_tile_shapes(gpu_available, tz, chunk, rows):
    tilts, proj, ref = _series()
    print(f"case shape tz={tz} chunk={chunk} rows={rows}", flush=True)
    plan = Plan(96, 5, tilts, 40, L.METHOD_SART, sart_iters=3, sart_subsets=4, relax=0.3,
                subset_order=L.SUBSETS_INTERLEAVED, positivity=False)
    plan.set_tile_shape(tz, chunk, rows)
    err = rel_l2(plan.execute_host(proj), ref)
    print("rel_l2", err)
    assert err <= TOL32


def test_autotune(gpu_available):
    tilts, proj, ref = _series()
    print("case autotune", flush=True)
    plan = Plan(96, 5, tilts, 40, L.METHOD_SART, sart_iters=3, sart_subsets=4, relax=0.3,
                subset_order=L.SUBSETS_INTERLEAVED, positivity=False)
    default = plan.execute_host(proj)
    tz, chunk, rows, pf = plan.autotune()
    print("chosen", tz, chunk, rows, pf)
    assert tz in (8, 16) and chunk in (8, 16) and rows >= 1 and pf == 0   # VolX 96: no prefetch
    tuned = plan.execute_host(proj)
    assert rel_l2(tuned, default) <= 1e-6
    assert rel_l2(tuned, ref) <= TOL32
    # the cached choice is reused by a plan of the same shape
    plan2 = Plan(96, 5, tilts, 40, L.METHOD_SART, sart_iters=3, sart_subsets=4, relax=0.3,
                 subset_order=L.SUBSETS_INTERLEAVED, positivity=False)
    assert plan2.autotune() == (tz, chunk, rows, pf)
    # a non-default shape joins a batch, where it runs as it does alone
    plan2.set_tile_shape(8, 16, rows)
    alone = plan2.execute_host(proj)
    d_in = DeviceArray.from_numpy(proj)
    d_out = DeviceArray(4 * int(np.prod(plan2.volume_shape)))
    batch = Batch([plan2, plan], graph=True)
    d_out1 = DeviceArray(4 * int(np.prod(plan.volume_shape)))
    batch.execute([d_in, d_in], [d_out, d_out1])
    synchronize()
    assert np.array_equal(d_out.to_numpy(plan2.volume_shape), alone)
    assert np.array_equal(d_out1.to_numpy(plan.volume_shape), tuned)
    # while the batch holds the plan its tables stay as they are
    with pytest.raises(L.AtrError, match="batch"):
        plan2.autotune()
    with pytest.raises(L.AtrError, match="batch"):
        plan2.set_tile_shape(16, 16, rows)
    batch.close()
    plan2.set_tile_shape(16, 16, rows)


@pytest.mark.parametrize("px,storage", [(512, L.STORAGE_F32), (768, L.STORAGE_F16),
                                        (1024, L.STORAGE_F32), (2048, L.STORAGE_F16)])
def test_prefetch_forward_projector(gpu_available, px, storage):
    # the band-prefetching forward projector (one wave streams the next band
    # into LDS) and the register-prefetching one (VolX 1024 / 2048: every
    # thread loads its columns of the next band into registers) give the
    # direct-staging projector's volume bit for bit at the same band rows, for
    # widths 512 / 768 / 1024 / 2048 (compile-time and run-time band pitch),
    # with and without empty voxels in the volume
    modes = [0] + [1] * (px % 256 == 0) + [2] * (px in (1024, 2048))
    rows = 5 if px == 2048 else 8     # band rows every mode fits (LDS)
    tilts = phantom.tilt_range(-60, 60, 4)
    proj, _ = phantom.tilt_series(91, px, 3, tilts, 96, n_ellipsoids=60, axis_min=6, axis_max=40,
                                  noise_frac=0.05)
    print(f"case prefetch fp px={px} storage={storage} modes={modes}", flush=True)
    kw = dict(sart_iters=2, sart_subsets=3, relax=0.3, subset_order=L.SUBSETS_INTERLEAVED,
              positivity=False, storage=storage)
    plan = Plan(px, 3, tilts, 96, L.METHOD_SART, **kw)
    out = {}
    for pf in modes:
        plan.set_tile_shape(16, 16, rows, pf)
        out[pf] = plan.execute_host(proj)
    plan.close()
    for pf in modes[1:]:
        assert np.array_equal(out[0], out[pf]), pf
    ref = oracle.recon_slices(proj.astype(np.float64), tilts, 1, 96, [0, 1, 2], iters=2,
                              n_subsets=3, relax=0.3, order=1, positivity=0)
    err = rel_l2(out[modes[-1]], ref)
    print("rel_l2", err)
    assert err <= (TOL32 if storage == L.STORAGE_F32 else 2e-3)
    # with empty voxels: the slice operator continuing from a volume that holds
    # -1e30 voxels runs the sentinel variant beside each projector
    start = np.random.default_rng(5).random((3, 96, px)).astype(np.float32)
    start[1, 40:44, 100:140] = -1e30
    res = {}
    for pf in modes:
        p = Plan(px, 3, tilts, 96, L.METHOD_SART, layout=L.LAYOUT_SLICE, **kw)
        p.set_tile_shape(16, 16, rows, pf)
        d_in = DeviceArray.from_numpy(proj)
        d_vol = DeviceArray.from_numpy(start)
        p.execute(d_in, d_vol, accumulate=True)
        res[pf] = d_vol.to_numpy(start.shape)
        p.close()
    for pf in modes[1:]:
        assert np.array_equal(res[0], res[pf]), pf


def test_prefetch_refused_for_other_widths(gpu_available):
    tilts = phantom.tilt_range(-60, 60, 10)
    plan = Plan(200, 2, tilts, 32, L.METHOD_SART, sart_iters=1, sart_subsets=1)
    print("case prefetch refusal", flush=True)
    with pytest.raises(L.AtrError, match="multiple of 256"):
        plan.set_tile_shape(16, 16, 4, 1)
    with pytest.raises(L.AtrError, match="1024 or 2048"):
        plan.set_tile_shape(16, 16, 4, 2)
    with pytest.raises(L.AtrError, match="0, 1 or 2"):
        plan.set_tile_shape(16, 16, 4, 3)
    plan.close()
