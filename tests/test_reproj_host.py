"""Host side of the re-projector without a GPU: the reference-tilt range
(CCalcReproj::mFindProjRange) of the library equals the oracle's on random
tilt lists and skip masks, and bad arguments fail cleanly."""
import ctypes

import numpy as np
import pytest

import oracle
from aretomo3_amd import _lib as L
from aretomo3_amd import projalign


def test_range_matches_oracle_on_random_series():
    rng = np.random.default_rng(11)
    checked = 0
    for trial in range(300):
        step = rng.choice([1.0, 2.0, 3.0, 5.0, 7.5, 12.0, 25.0])
        lo = -rng.uniform(30, 70)
        t = np.arange(lo, 70.0, step, dtype=np.float32) + np.float32(rng.uniform(-0.5, 0.5))
        t = t[np.abs(t) < 85]
        if len(t) < 3:
            continue
        idx = int(rng.integers(0, len(t)))
        # the projection-matching order: this angle and larger |angles| skipped
        skip = (np.abs(t) >= abs(t[idx]) - 1e-6).astype(np.uint8)
        if trial % 3 == 0:
            skip = (rng.uniform(0, 1, len(t)) < 0.3).astype(np.uint8)
        try:
            got = projalign.find_proj_range(t, idx, skip)
        except L.AtrError:
            # only the fallback off the end of the series is refused
            r0 = oracle.reproj_range(t, idx, skip)
            assert r0[0] < 0 or r0[1] >= len(t)
            continue
        assert got == oracle.reproj_range(t, idx, skip), (t, idx, skip)
        assert 0 <= got[0] <= got[1] < len(t) and got[1] - got[0] <= 9
        checked += 1
    assert checked > 200


def test_reproj_argument_errors():
    lib = L.lib()
    h = ctypes.c_void_p()
    assert lib.atr_reproj_create(1, 4, 5, 8, ctypes.byref(h)) == L.ATR_ERR_INVALID_ARG
    assert lib.atr_reproj_create(8, 4, 1, 8, ctypes.byref(h)) == L.ATR_ERR_INVALID_ARG
    assert lib.atr_reproj_create(8, 4, 5, 8, None) == L.ATR_ERR_INVALID_ARG
    t = np.array([-10, 0, 10], np.float32)
    a, b = ctypes.c_int(), ctypes.c_int()
    assert lib.atr_reproj_find_range(t.ctypes.data, None, 3, 3, ctypes.byref(a),
                                     ctypes.byref(b)) == L.ATR_ERR_INVALID_ARG
    # the fallback (neighbour toward 0 deg) outside the series is an error,
    # not an out-of-range read; inside it is the fallback of the reference
    with pytest.raises(L.AtrError):
        projalign.find_proj_range(np.array([80.0, 20.0], np.float32), 0, np.array([1, 1], np.uint8))
    assert projalign.find_proj_range(np.array([-80.0, -20.0], np.float32), 0,
                                     np.array([1, 1], np.uint8)) == (1, 1)
    lib.atr_reproj_destroy(None)
