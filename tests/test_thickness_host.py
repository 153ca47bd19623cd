"""The thickness edge search of the library (atr_thickness_edges) against the
oracle restatement of CCalcVolThick::mSmooth + mDetectEdges, bit for bit, on
random and hand-built profiles; argument checks."""
import numpy as np
import pytest

import oracle
from aretomo3_amd import _lib as L
from aretomo3_amd.recon import thickness_edges


def oracle_edges(profile, binning):
    sm = oracle.thick_smooth(profile)
    return oracle.thick_edges(sm, binning), sm


@pytest.mark.parametrize("n", [10, 11, 39, 64, 161])
def test_edges_match_oracle(n):
    rng = np.random.default_rng(n)
    for trial in range(40):
        z = np.arange(n)
        c = rng.uniform(0.3, 0.7) * n
        w = rng.uniform(0.1, 0.4) * n
        prof = (np.exp(-((z - c) / w) ** 4) + rng.normal(0, 0.05, n)).astype(np.float32)
        binning = np.float32(rng.choice([1.0, 20.0, 23.7, 40.0]))
        got, sm = thickness_edges(prof, binning)
        (e0, e1, th, off), sm_ref = oracle_edges(prof, binning)
        np.testing.assert_array_equal(sm, sm_ref)
        assert (got["low_edge"], got["high_edge"], got["thickness"], got["offset_z"]) == (e0, e1, th, off)
        assert got["n_profile"] == n and got["binning"] == binning


def test_edges_box_and_errors():
    p = np.zeros(100, np.float32)
    p[30:70] = 1.0
    got, _ = thickness_edges(p, 20.0)
    (e0, e1, th, off), _ = oracle_edges(p, 20.0)
    assert (got["low_edge"], got["high_edge"], got["thickness"], got["offset_z"]) == (e0, e1, th, off)
    with pytest.raises(L.AtrError):
        thickness_edges(np.zeros(9, np.float32), 20.0)
    with pytest.raises(L.AtrError):
        thickness_edges(np.zeros(20, np.float32), 0.0)
