"""CCalcVolThick on the GPU (atr_calc_vol_thick) against the oracle's
CPU pipeline: the std profile within 1e-4 (the binned WBP volume is FP32 vs
FP64) and the same sample edges, thickness and z offset."""
import numpy as np
import pytest

import oracle
from aretomo3_amd import CalcVolThick, phantom

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("slab,pixel", [(400, 1.0), (240, 0.8)])
def test_calc_vol_thick_matches_oracle(gpu_available, slab, pixel):
    tilts = phantom.tilt_range(-60, 60, 3).astype(np.float32)
    raw, shifts = phantom.tilt_series(61, 1024, 1024, tilts, slab, n_ellipsoids=300, axis_min=10,
                                      axis_max=60, tilt_axis=86.0, max_shift=3.0, noise_frac=0.05)
    axes = np.full(len(tilts), 86.0, np.float32)
    print(f"case thickness slab={slab} pixel={pixel}", flush=True)
    c = CalcVolThick()
    got = c.do_it(raw, tilts, pixel, tilt_axis=axes, shifts=shifts, fill_seed=3)
    ref = oracle.calc_vol_thick(raw, tilts, axes, shifts, pixel, seed=3)
    print("gpu", got, "oracle", ref["edges"])
    prof = c.profile
    assert prof.shape == ref["profile"].shape
    np.testing.assert_allclose(prof, ref["profile"], rtol=1e-4, atol=1e-7)
    e0, e1, th, off = ref["edges"]
    assert (got["low_edge"], got["high_edge"], got["thickness"], got["offset_z"]) == (e0, e1, th, off)
    assert got["vol_z"] == ref["vol_z"] and got["binning"] == pytest.approx(ref["binning"])
    assert c.get_thickness() == th
    # a sample inside the volume: both edges found on their own halves
    assert 0 < got["low_edge"] < got["high_edge"] < got["vol_z"] * got["binning"]
    assert th >= slab * 0.8
