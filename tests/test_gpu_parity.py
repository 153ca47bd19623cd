"""Parity of the HIP path (through the C ABI) with the FP64 oracle.

Tolerance for the FP32-storage path: relative L2 <= 1e-5 against FP64; for
FP16/BF16 projection storage: <= 2e-3.  Sizes are small enough for the oracle
to finish in seconds; the inclusion decisions of both sides come from the same
FP64 expressions, so what remains is FP32 rounding.
"""
import numpy as np
import pytest

import oracle
from aretomo3_amd import Plan
from aretomo3_amd import _lib as L
from helpers import rel_l2, small_series

pytestmark = pytest.mark.gpu
TOL32 = 1e-5


def _oracle_series(proj, tilts, method, vol_z, **kw):
    n, py, px = proj.shape
    return oracle.recon_slices(proj.astype(np.float64), tilts, method, vol_z, list(range(py)), **kw)


@pytest.mark.parametrize("px,py,vol_z,step", [(64, 12, 32, 6), (96, 8, 40, 3), (66, 5, 17, 10)])
def test_sart_matches_oracle(gpu_available, px, py, vol_z, step):
    from aretomo3_amd import phantom
    tilts, proj, _ = small_series(px=px, py=py, vol_z=vol_z, tilts=phantom.tilt_range(-60, 60, step))
    print(f"case sart px={px} py={py} vz={vol_z} n={len(tilts)}", flush=True)
    plan = Plan(px, py, tilts, vol_z, L.METHOD_SART, sart_iters=3, sart_subsets=3)
    gpu = plan.execute_host(proj)
    ref = _oracle_series(proj, tilts, 1, vol_z, iters=3, n_subsets=3)
    err = rel_l2(gpu, ref)
    print("rel_l2", err)
    assert err <= TOL32


def test_sart_options_match_oracle(gpu_available):
    tilts, proj, _ = small_series(px=64, py=6, vol_z=24)
    print("case sart options", flush=True)
    plan = Plan(64, 6, tilts, 24, L.METHOD_SART, sart_iters=4, sart_subsets=5, relax=0.3,
                subset_order=L.SUBSETS_INTERLEAVED, positivity=False, sart_start_tilt=2,
                sart_n_tilts=15)
    gpu = plan.execute_host(proj)
    ref = _oracle_series(proj, tilts, 1, 24, start_tilt=2, n_sart_tilts=15, iters=4,
                         n_subsets=5, relax=0.3, order=1, positivity=0)
    err = rel_l2(gpu, ref)
    print("rel_l2", err)
    assert err <= TOL32


@pytest.mark.parametrize("px,py,vol_z", [(64, 6, 32), (48, 5, 20), (96, 4, 30)])
def test_wbp_matches_oracle(gpu_available, px, py, vol_z):
    tilts, proj, _ = small_series(px=px, py=py, vol_z=vol_z)
    print(f"case wbp px={px} py={py} vz={vol_z}", flush=True)
    plan = Plan(px, py, tilts, vol_z, L.METHOD_WBP)
    gpu = plan.execute_host(proj)
    ref = _oracle_series(proj, tilts, 0, vol_z)
    err = rel_l2(gpu, ref)
    print("rel_l2", err)
    assert err <= TOL32


@pytest.mark.parametrize("method", [L.METHOD_SART, L.METHOD_WBP])
def test_prestage_matches_oracle(gpu_available, method):
    tilts, raw, shifts = small_series(px=64, py=48, vol_z=24, axis=85.0, shift=3.0)
    print(f"case prestage method={method}", flush=True)
    axes = np.full(len(tilts), 85.0, np.float32)
    plan = Plan(64, 48, tilts, 24, method, tilt_axis=axes, shifts=shifts, sart_iters=2,
                sart_subsets=3, fill_seed=11)
    gpu = plan.execute_host(raw)
    aligned = oracle.prestage(raw, axes, shifts, 1, 11)
    ref = oracle.recon_slices(aligned, tilts, method, 24, list(range(aligned.shape[1])),
                              iters=2, n_subsets=3, seed=11)
    err = rel_l2(gpu, ref)
    print("rel_l2", err)
    assert err <= TOL32


@pytest.mark.parametrize("case", ["sart_xzy", "sart_il_xzy", "wbp_xzy"])
def test_golden_fixture(gpu_available, case):
    import os
    g = dict(np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                  "series_small.npz")))
    print(f"case golden {case}", flush=True)
    raw = g["raw"]
    n, py, px = raw.shape
    opts = {
        "sart_xzy": dict(method=L.METHOD_SART, sart_iters=3, sart_subsets=3),
        "sart_il_xzy": dict(method=L.METHOD_SART, sart_iters=3, sart_subsets=3, relax=0.3,
                            subset_order=L.SUBSETS_INTERLEAVED, positivity=False),
        "wbp_xzy": dict(method=L.METHOD_WBP),
    }[case]
    method = opts.pop("method")
    plan = Plan(px, py, g["tilts"], int(g["vol_z"]), method, tilt_axis=g["axes"],
                shifts=g["shifts"], fill_seed=int(g["seed"]), **opts)
    gpu = plan.execute_host(raw)
    err = rel_l2(gpu, g[case])
    print("rel_l2", err)
    assert gpu.shape == g[case].shape
    assert err <= TOL32


@pytest.mark.parametrize("method", [L.METHOD_SART, L.METHOD_WBP])
def test_acquisition_order_and_asymmetric_tilts(gpu_available, method):
    # tilts in a dose-symmetric acquisition order (0, +3, -3, +6, ...) over an
    # asymmetric range, not sorted by angle: the subsets, tables and the WBP
    # tilt list follow the given order exactly as the oracle does
    from aretomo3_amd import phantom
    order = [0.0]
    for k in range(1, 16):
        order += [3.0 * k, -3.0 * k]
    tilts = np.array([t for t in order if -45.0 <= t <= 57.0], np.float32)
    _, proj, _ = small_series(px=72, py=5, vol_z=28, tilts=tilts)
    print(f"case acquisition order method={method} n={len(tilts)}", flush=True)
    plan = Plan(72, 5, tilts, 28, method, sart_iters=3, sart_subsets=4)
    gpu = plan.execute_host(proj)
    ref = _oracle_series(proj, tilts, method, 28, iters=3, n_subsets=4)
    err = rel_l2(gpu, ref)
    print("rel_l2", err)
    assert err <= TOL32


@pytest.mark.parametrize("n_tilts,subsets", [(2, 1), (3, 3)])
def test_sart_few_tilts(gpu_available, n_tilts, subsets):
    # the smallest series: one or two tilts per subset, every subset a single
    # forward-projection workgroup per slice
    tilts = np.linspace(-30.0, 30.0, n_tilts).astype(np.float32)
    _, proj, _ = small_series(px=40, py=3, vol_z=12, tilts=tilts)
    print(f"case few tilts n={n_tilts} subsets={subsets}", flush=True)
    plan = Plan(40, 3, tilts, 12, L.METHOD_SART, sart_iters=4, sart_subsets=subsets)
    gpu = plan.execute_host(proj)
    ref = _oracle_series(proj, tilts, 1, 12, iters=4, n_subsets=subsets)
    err = rel_l2(gpu, ref)
    print("rel_l2", err)
    assert err <= TOL32
