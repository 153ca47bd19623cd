"""Reconstruct from files (-Cmd 2): a raw MRC stack with dark sections and a
.aln with local patch shifts go through aretomo3_amd.files.reconstruct_files
(pre-stage, -AtBin Fourier binning, SART or WBP, -FlipVol, MRC output); the
volume matches the FP64 oracle run on the same selected sections (relative
L2 <= 1e-5) and the written MRC holds it with the binned pixel size."""
import numpy as np
import pytest

import oracle
from aretomo3_amd import _lib as L
from aretomo3_amd import files as F
from aretomo3_amd import phantom
from helpers import rel_l2

pytestmark = pytest.mark.gpu
TOL32 = 1e-5


def write_inputs(tmp_path, n_patches=6):
    rx, ry = 64, 48
    tilts = phantom.tilt_range(-50, 50, 10).astype(np.float32)          # 11 aligned tilts
    raw, shifts = phantom.tilt_series(41, rx, ry, tilts, 40, n_ellipsoids=25, axis_min=3,
                                      axis_max=9, tilt_axis=86.0, max_shift=2.0, noise_frac=0.05)
    rng = np.random.default_rng(2)
    n = len(tilts)
    # raw stack of n + 2 sections: dark ones at 1-based sections 1 and 9
    secs = [s for s in range(1, n + 3) if s not in (1, 9)]
    stack = rng.normal(0, 0.01, (n + 2, ry, rx)).astype(np.float32)
    stack[np.array(secs) - 1] = raw
    local = np.empty((n, 5, n_patches), np.float32)
    local[:, 0] = rng.uniform(-rx / 2, rx / 2, (n, n_patches))
    local[:, 1] = rng.uniform(-ry / 2, ry / 2, (n, n_patches))
    local[:, 2:4] = rng.uniform(-2, 2, (n, 2, n_patches))
    local[:, 4] = (rng.uniform(0, 1, (n, n_patches)) < 0.8)
    # round-trip the table through the file precision first (%8.2f)
    local = np.round(local, 2).astype(np.float32)
    shifts = np.round(shifts, 3).astype(np.float32)
    axes = np.full(n, 86.0, np.float32)
    aln = F.Alignment(raw_size=(rx, ry, n + 2), sec=np.array(secs, np.int32), tilt_axis=axes,
                      shifts=shifts, tilts=tilts, local=local,
                      dark_frame=np.array([0, n + 1], np.int32),
                      dark_sec=np.array([1, 9], np.int32),
                      dark_tilt=np.array([-60, 60], np.float32), thickness=30)
    F.write_mrc(tmp_path / "ts.mrc", stack, pixel_size=1.7)
    F.write_aln(tmp_path / "ts.aln", aln)
    return raw, aln


@pytest.mark.parametrize("wbp,flip,at_bin", [(False, 0, 2.0), (True, 1, 2.0), (False, 2, 1.0)])
def test_reconstruct_files_matches_oracle(gpu_available, tmp_path, wbp, flip, at_bin):
    raw, aln = write_inputs(tmp_path)
    print(f"case files wbp={wbp} flip={flip} bin={at_bin}", flush=True)
    vol_z = 24
    vol, path = F.reconstruct_files(tmp_path / "ts.mrc", tmp_path / "ts.aln", tmp_path / "ts",
                                    vol_z=vol_z * at_bin, ext_z=0, at_bin=at_bin, wbp=wbp,
                                    sart=(2, 4), flip_vol=flip)
    z = F.binned_vol_z(vol_z * at_bin, 30, 0, at_bin)
    a = F.read_aln(tmp_path / "ts.aln")
    aligned = oracle.prestage(raw, a.tilt_axis, a.shifts, 1, 0, local=a.local)
    if at_bin != 1.0:
        aligned = oracle.bin_stack(aligned, at_bin)
    n = len(a.tilts)
    subsets = max(1, n // 4)
    ref = oracle.recon_slices(aligned, a.tilts, 0 if wbp else 1, z, list(range(aligned.shape[1])),
                              iters=2, n_subsets=subsets)
    if flip:
        # -FlipVol 1: z reversed; -FlipVol 2: plain xzy -> xyz (CAreTomoMain.cpp:900)
        ref = oracle.flip_vol(ref, 1 if flip == 1 else 0)
    assert vol.shape == ref.shape
    err = rel_l2(vol, ref)
    print("rel_l2", err, "shape", vol.shape)
    assert err <= TOL32
    back = F.read_mrc(path)
    np.testing.assert_array_equal(back, vol)
    info = F.mrc_info(path)
    ay = F.aligned_size(64, 48, 86.0)[1]
    by = aligned.shape[1]
    want = np.float32(np.float32(1.7) / np.float32(by)) * np.float32(ay) if at_bin != 1.0 else 1.7
    assert info["pixel_size"] == pytest.approx(float(want), rel=1e-6)
    assert path.endswith("ts_Vol.mrc")


def test_reconstruct_all_volumes(gpu_available, tmp_path):
    # -AtBin 2 3 4 with even/odd sums: _Vol (SART, bin 2), _EVN_Vol and
    # _ODD_Vol (same alignment), _2ND_Vol (WBP, bin 3), _3RD_Vol (SART, bin 4)
    raw, aln = write_inputs(tmp_path)
    rng = np.random.default_rng(9)
    n = len(aln.tilts)
    full = F.read_mrc(tmp_path / "ts.mrc")
    evn = (full * 0.5 + rng.normal(0, 0.01, full.shape)).astype(np.float32)
    odd = (full - evn).astype(np.float32)
    F.write_mrc(tmp_path / "ts_EVN.mrc", evn, 1.7)
    F.write_mrc(tmp_path / "ts_ODD.mrc", odd, 1.7)
    print("case files all volumes", flush=True)
    res = F.reconstruct_all(tmp_path / "ts.mrc", tmp_path / "ts.aln", tmp_path / "ts", vol_z=60,
                            ext_z=0, at_bin=(2.0, 3.0, 4.0), sart=(2, 4), flip_vol=1,
                            evn_mrc=tmp_path / "ts_EVN.mrc", odd_mrc=tmp_path / "ts_ODD.mrc")
    assert sorted(res) == [0, 1, 2, 3, 4]
    a = F.read_aln(tmp_path / "ts.aln")
    subsets = max(1, n // 4)
    cases = {0: (full, 2.0, F.binned_vol_z(60, 30, 0, 2.0), 1),
             1: (evn, 2.0, F.binned_vol_z(60, 30, 0, 2.0), 1),
             2: (odd, 2.0, F.binned_vol_z(60, 30, 0, 2.0), 1),
             3: (full, 3.0, F.second_vol_z(60, 30, 0, 3.0), 0),
             4: (full, 4.0, F.second_vol_z(60, 30, 0, 4.0), 1)}
    for idx, (stack, b, z, method) in cases.items():
        sel = stack[a.sec - 1]
        aligned = oracle.bin_stack(oracle.prestage(sel, a.tilt_axis, a.shifts, 1, 0, local=a.local), b)
        ref = oracle.recon_slices(aligned, a.tilts, method, z, list(range(aligned.shape[1])),
                                  iters=2, n_subsets=subsets)
        ref = oracle.flip_vol(ref, 1)
        vol, path = res[idx]
        err = rel_l2(vol, ref)
        print(idx, path, vol.shape, "rel_l2", err)
        assert vol.shape == ref.shape and err <= TOL32
        assert path.endswith("ts" + F.volume_suffix(idx))
        np.testing.assert_array_equal(F.read_mrc(path), vol)
