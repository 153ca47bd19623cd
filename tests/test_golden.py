"""The committed golden fixtures (tests/golden/series_small.npz) against the
oracle (CPU) -- a change to the oracle that moves its outputs fails here."""
import os

import numpy as np

import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "series_small.npz")


def load():
    return dict(np.load(GOLDEN))


def test_golden_prestage():
    g = load()
    a = oracle.prestage(g["raw"], g["axes"], g["shifts"], 1, int(g["seed"]))
    np.testing.assert_array_equal(a.astype(np.float32), g["aligned"])


def test_golden_recon():
    g = load()
    a = oracle.prestage(g["raw"], g["axes"], g["shifts"], 1, int(g["seed"]))
    vz, seed = int(g["vol_z"]), int(g["seed"])
    sl = np.arange(a.shape[1], dtype=np.int32)
    cases = {
        "sart_xzy": dict(method=1, iters=3, n_subsets=3, relax=-1.0, order=0, positivity=1),
        "sart_il_xzy": dict(method=1, iters=3, n_subsets=3, relax=0.3, order=1, positivity=0),
        "wbp_xzy": dict(method=0),
    }
    for name, kw in cases.items():
        m = kw.pop("method")
        v = oracle.recon_slices(a, g["tilts"], m, vz, sl, seed=seed, **kw)
        np.testing.assert_allclose(v, g[name], rtol=1e-6, atol=1e-6 * np.abs(g[name]).max(),
                                   err_msg=name)
