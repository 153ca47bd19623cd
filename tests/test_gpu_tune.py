"""Tile shapes and the autotuner: every back-projection shape (z rows per
thread x tilts per LDS stage) and forward-projection band height gives the
oracle's result; the autotuner returns one of them and its plan stays exact."""
import numpy as np
import pytest

import oracle
from aretomo3_amd import Batch, Plan, phantom
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
    tz, chunk, rows = plan.autotune()
    print("chosen", tz, chunk, rows)
    assert tz in (8, 16) and chunk in (8, 16) and rows >= 2
    tuned = plan.execute_host(proj)
    assert rel_l2(tuned, default) <= 1e-6
    assert rel_l2(tuned, ref) <= TOL32
    # the cached choice is reused by a plan of the same shape
    plan2 = Plan(96, 5, tilts, 40, L.METHOD_SART, sart_iters=3, sart_subsets=4, relax=0.3,
                 subset_order=L.SUBSETS_INTERLEAVED, positivity=False)
    assert plan2.autotune() == (tz, chunk, rows)
    # a non-default back-projection shape cannot join a batch
    plan2.set_tile_shape(8, 16, rows)
    with pytest.raises(L.AtrError):
        Batch([plan2])
