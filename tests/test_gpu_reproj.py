"""The ProjAlign re-projector on the GPU against the FP64 oracle: same empty
pixels, relative L2 <= 1e-6 elsewhere (coordinates are FP64 on both sides,
so every nearest-sample pick agrees; only the final FP32 store differs)."""
import numpy as np
import pytest

import oracle
from aretomo3_amd import phantom
from aretomo3_amd.projalign import CalcReproj, find_proj_range
from helpers import rel_l2

pytestmark = pytest.mark.gpu


def compare(gpu, ref):
    empty_g, empty_r = gpu < -1e29, ref < -1e29
    assert np.array_equal(empty_g, empty_r), (empty_g.sum(), empty_r.sum())
    assert (~empty_r).sum() > 0
    return rel_l2(gpu[~empty_r], ref[~empty_r])


@pytest.mark.parametrize("px,py,vol_z,step", [(96, 10, 48, 3.0), (130, 6, 40, 2.0), (512, 32, 128, 3.0)])
def test_reproj_matches_oracle(gpu_available, px, py, vol_z, step):
    tilts = phantom.tilt_range(-60, 60, step).astype(np.float32)
    projs, _ = phantom.tilt_series(51, px, py, tilts, vol_z, n_ellipsoids=30, axis_min=3,
                                   axis_max=12, noise_frac=0.05)
    projs[3, :, : px // 5] = -1e30          # an empty band in one projection
    n = len(tilts)
    rp = CalcReproj(px, py, n, vol_z)
    for idx in (0, n // 4, n // 2, 3 * n // 4, n - 1):
        skip = (np.abs(tilts) >= abs(tilts[idx]) - 1e-6).astype(np.uint8)
        skip[idx] = 1
        try:
            rng = find_proj_range(tilts, idx, skip)
        except Exception:
            continue
        print(f"case reproj px={px} py={py} vz={vol_z} idx={idx} range={rng}", flush=True)
        gpu = rp.do_it(projs, tilts, skip, idx)
        ref = oracle.reproj(projs, tilts, vol_z, idx, skip)
        err = compare(gpu, ref)
        print("rel_l2", err)
        assert err <= 1e-6
    rp.close()


def test_reproj_zero_tilt_identity_and_fallback(gpu_available):
    rng = np.random.default_rng(3)
    p = rng.uniform(1, 2, (2, 4, 64)).astype(np.float32)
    t = np.array([0.0, 30.0], np.float32)
    print("case reproj identity", flush=True)
    rp = CalcReproj(64, 4, 2, 16)
    out = rp.do_it(p, t, np.array([0, 1], np.uint8), 0)
    np.testing.assert_array_equal(out[:, :-1], p[0, :, :-1])
    assert np.all(out[:, -1] < -1e29)
    rp.close()
    # 30-degree steps: nothing within 20.5 deg, the neighbour toward 0 is used
    t5 = np.array([-60, -30, 0, 30, 60], np.float32)
    p5 = rng.uniform(1, 2, (5, 3, 48)).astype(np.float32)
    rp = CalcReproj(48, 3, 5, 12)
    skip = np.array([0, 0, 0, 0, 1], np.uint8)
    assert find_proj_range(t5, 4, skip) == (3, 3)
    gpu = rp.do_it(p5, t5, skip, 4)
    ref = oracle.reproj(p5, t5, 12, 4, skip)
    assert compare(gpu, ref) <= 1e-6
    rp.close()
