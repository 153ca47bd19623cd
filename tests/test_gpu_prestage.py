"""The full pre-stage on the GPU against the FP64 oracle: local patch shifts
(Correct/GCorrPatchShift.cu:18-48), -AtBin Fourier binning
(Correct/CBinStack.cpp:54-106) and FFT lengths with prime factors above 5.
Tolerance: relative L2 <= 1e-5 (FP32 storage against FP64)."""
import numpy as np
import pytest

import oracle
from aretomo3_amd import Plan
from aretomo3_amd import _lib as L
from helpers import rel_l2, small_series

pytestmark = pytest.mark.gpu
TOL32 = 1e-5


def patch_table(seed, n_tilts, raw_x, raw_y, n_patches, max_shift=2.5):
    """[n][5][n_patches]: centres in the centred raw frame, shifts, good flags."""
    rng = np.random.default_rng(seed)
    lp = np.empty((n_tilts, 5, n_patches), np.float32)
    lp[:, 0] = rng.uniform(-raw_x / 2, raw_x / 2, (n_tilts, n_patches))
    lp[:, 1] = rng.uniform(-raw_y / 2, raw_y / 2, (n_tilts, n_patches))
    lp[:, 2:4] = rng.uniform(-max_shift, max_shift, (n_tilts, 2, n_patches))
    lp[:, 4] = (rng.uniform(0, 1, (n_tilts, n_patches)) < 0.8).astype(np.float32)
    lp[0, 4] = 0.0      # one tilt with no good patch: global shift only
    return lp


def oracle_recon(raw, tilts, method, vol_z, axes=None, shifts=None, local=None, bin_=1.0,
                 seed=0, **kw):
    if axes is not None:
        aligned = oracle.prestage(raw, axes, shifts, 1, seed, local=local)
    else:
        aligned = raw.astype(np.float64)
    if bin_ != 1.0:
        aligned = oracle.bin_stack(aligned, bin_)
    return oracle.recon_slices(aligned, tilts, method, vol_z, list(range(aligned.shape[1])),
                               seed=seed, **kw)


@pytest.mark.parametrize("method", [L.METHOD_SART, L.METHOD_WBP])
def test_local_shifts_match_oracle(gpu_available, method):
    tilts, raw, shifts = small_series(seed=21, px=64, py=48, vol_z=24, axis=85.0, shift=3.0)
    n = len(tilts)
    axes = np.full(n, 85.0, np.float32)
    lp = patch_table(5, n, 64, 48, 12)
    print(f"case local shifts method={method}", flush=True)
    plan = Plan(64, 48, tilts, 24, method, tilt_axis=axes, shifts=shifts, local_params=lp,
                sart_iters=2, sart_subsets=3, fill_seed=13)
    gpu = plan.execute_host(raw)
    ref = oracle_recon(raw, tilts, method, 24, axes, shifts, lp, seed=13, iters=2, n_subsets=3)
    err = rel_l2(gpu, ref)
    # the local shifts must matter at this size
    plain = oracle_recon(raw, tilts, method, 24, axes, shifts, None, seed=13, iters=2,
                         n_subsets=3)
    print("rel_l2", err, "vs no local", rel_l2(plain, ref))
    assert rel_l2(plain, ref) > 1e-2
    assert err <= TOL32


# (raw x, raw y, tilt axis or None, bin, method, vol_z, with local shifts)
BIN_CASES = [
    (64, 48, 85.0, 2.0, L.METHOD_SART, 16, False),
    (64, 48, 85.0, 1.5, L.METHOD_WBP, 20, False),     # binned width 42 = 2*3*7
    (64, 48, None, 2.0, L.METHOD_WBP, 12, False),     # aligned input, no pre-stage
    (96, 64, 86.0, 3.0, L.METHOD_SART, 10, True),
    (80, 60, -84.0, 2.0, L.METHOD_WBP, 16, True),
]


@pytest.mark.parametrize("case", range(len(BIN_CASES)))
def test_fourier_binning_matches_oracle(gpu_available, case):
    rx, ry, axis, bin_, method, vol_z, local = BIN_CASES[case]
    tilts, raw, shifts = small_series(seed=30 + case, px=rx, py=ry, vol_z=2 * vol_z,
                                      axis=float("nan") if axis is None else axis,
                                      shift=0.0 if axis is None else 2.0)
    n = len(tilts)
    kw = dict(sart_iters=2, sart_subsets=2, fill_seed=17 + case)
    axes = lp = None
    if axis is not None:
        axes = np.full(n, axis, np.float32)
        kw.update(tilt_axis=axes, shifts=shifts)
        if local:
            lp = patch_table(case, n, rx, ry, 8)
            kw.update(local_params=lp)
    print(f"case bin {BIN_CASES[case]}", flush=True)
    plan = Plan(rx, ry, tilts, vol_z, method, bin=bin_, **kw)
    gpu = plan.execute_host(raw)
    ref = oracle_recon(raw, tilts, method, vol_z, axes, shifts if axis is not None else None,
                       lp, bin_, seed=17 + case, iters=2, n_subsets=2)
    assert gpu.shape == ref.shape
    ax, ay = (rx, ry) if axis is None else oracle.aligned_size(rx, ry, axis)
    bx, by = oracle.bin_size(ax, ay, bin_)
    assert (plan.vol_x, plan.vol_y) == (bx, by)
    err = rel_l2(gpu, ref)
    print("binned", (bx, by), "rel_l2", err)
    assert err <= TOL32


@pytest.mark.parametrize("px,py,vol_z", [(56, 4, 26), (98, 3, 22), (194, 2, 34)])
def test_wbp_fft_prime_factors(gpu_available, px, py, vol_z):
    # 2*VolX = 112 = 16*7, 196 = 4*7*7, 388 = 4*97; 2*VolZ = 52 = 4*13, 44 = 4*11, 68 = 4*17
    tilts, proj, _ = small_series(px=px, py=py, vol_z=vol_z)
    print(f"case wbp fft px={px} vz={vol_z}", flush=True)
    plan = Plan(px, py, tilts, vol_z, L.METHOD_WBP)
    gpu = plan.execute_host(proj)
    ref = oracle.recon_slices(proj.astype(np.float64), tilts, 0, vol_z, list(range(py)))
    err = rel_l2(gpu, ref)
    print("rel_l2", err)
    assert err <= TOL32


def test_per_tilt_axes_and_many_patches(gpu_available):
    # a tilt axis refined per tilt (84..86 deg; the aligned size follows the
    # first tilt's axis) and a 6 x 6 patch grid, as -Patch 6 6 leaves it
    tilts, raw, shifts = small_series(seed=41, px=72, py=56, vol_z=20, axis=85.0, shift=2.5)
    n = len(tilts)
    axes = np.linspace(84.0, 86.0, n).astype(np.float32)
    lp = patch_table(9, n, 72, 56, 36)
    print("case per-tilt axes, 36 patches", flush=True)
    plan = Plan(72, 56, tilts, 20, L.METHOD_SART, tilt_axis=axes, shifts=shifts, local_params=lp,
                sart_iters=2, sart_subsets=3, fill_seed=3)
    gpu = plan.execute_host(raw)
    ref = oracle_recon(raw, tilts, L.METHOD_SART, 20, axes, shifts, lp, seed=3, iters=2,
                       n_subsets=3)
    assert gpu.shape == ref.shape
    err = rel_l2(gpu, ref)
    print("rel_l2", err)
    assert err <= TOL32
