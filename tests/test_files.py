"""On-disk formats without a GPU: MRC read/write checked byte by byte against
the MRC2014 word layout decoded independently with numpy, the .aln reader and
writer against the reference's printf layouts (CSaveAlignFile.cpp:67-127)
and its loader rules (CLoadAlignFile.cpp), and the tilt-file reader
(CTsPackage.cpp:358-388)."""
import os

import numpy as np
import pytest

from aretomo3_amd import _lib as L
from aretomo3_amd import files as F


def header_words(path):
    raw = open(path, "rb").read(1024)
    return raw, np.frombuffer(raw, "<i4"), np.frombuffer(raw, "<f4"), np.frombuffer(raw, "<i2")


def test_mrc_volume_header_is_the_reference_header(tmp_path):
    rng = np.random.default_rng(1)
    vol = rng.standard_normal((5, 7, 9)).astype(np.float32)
    p = tmp_path / "a_Vol.mrc"
    F.write_mrc(p, vol, pixel_size=2.5)
    raw, w, f, h = header_words(p)
    assert os.path.getsize(p) == 1024 + vol.nbytes
    assert list(w[0:4]) == [9, 7, 5, 2]                 # nx ny nz mode
    assert list(w[4:7]) == [0, 0, 0]                    # starts
    assert list(w[7:10]) == [9, 7, 5]                   # mx my mz
    assert np.allclose(f[10:13], [9 * 2.5, 7 * 2.5, 5 * 2.5])
    assert list(f[13:16]) == [90, 90, 90]
    assert list(w[16:19]) == [1, 2, 3]                  # mapc mapr maps
    assert list(f[19:22]) == [0, 0, 0]                  # statistics left 0
    assert w[22] == 0 and w[23] == 0                    # ispg, nsymbt: no extended header
    assert h[48] == -16224                              # dvid
    assert raw[104:108] == b"AGAR" and w[27] == 20140   # EXTTYPE, NVERSION
    assert h[64] == 0 and h[65] == 0 and h[66] == 1     # nint, nreal, sub
    assert h[80] == 5                                   # type
    assert list(f[46:49]) == [90, 90, 90]               # tilt angles
    assert raw[208:212] == b"MAP\x00" and raw[212:216] == b"\x44\x44\x00\x00"
    assert w[55] == 0 and raw[224:] == bytes(800)       # no labels
    data = np.fromfile(p, "<f4", offset=1024).reshape(vol.shape)
    np.testing.assert_array_equal(data, vol)
    np.testing.assert_array_equal(F.read_mrc(p), vol)
    info = F.mrc_info(p)
    assert (info["nx"], info["ny"], info["nz"], info["mode"]) == (9, 7, 5, 2)
    assert info["pixel_size"] == pytest.approx(2.5)
    F.write_mrc(p, vol, pixel_size=1.0, with_stats=True)
    _, _, f, _ = header_words(p)
    assert f[19] == vol.min() and f[20] == vol.max()
    assert f[21] == pytest.approx(vol.astype(np.float64).mean(), rel=1e-6)


def make_mrc(path, data, mode, big_endian=False, ext=b""):
    """An MRC written from the MRC2014 word layout (independent of the library)."""
    e = ">" if big_endian else "<"
    nz, ny, nx = data.shape
    hdr = np.zeros(256, e + "i4")
    hdr[0:4] = [nx, ny, nz, mode]
    hdr[7:10] = [nx, ny, nz]
    hdr[23] = len(ext)
    hb = bytearray(hdr.tobytes())
    hb[40:44] = np.array([nx * 1.5], e + "f4").tobytes()
    dt = {1: "i2", 2: "f4", 6: "u2", 0: "i1"}[mode]
    with open(path, "wb") as fh:
        fh.write(bytes(hb) + ext + data.astype(e + dt).tobytes())


@pytest.mark.parametrize("mode,big", [(1, False), (6, False), (2, True), (1, True), (6, True)])
def test_mrc_reader_modes_and_byte_order(tmp_path, mode, big):
    rng = np.random.default_rng(mode)
    lo, hi = {1: (-3000, 3000), 6: (0, 60000), 2: (-5, 5)}[mode]
    data = rng.uniform(lo, hi, (4, 6, 10))
    if mode != 2:
        data = np.round(data)
    p = tmp_path / "s.mrc"
    make_mrc(p, data, mode, big, ext=b"\x07" * 96)
    info = F.mrc_info(p)
    assert info["swapped"] == int(big) and info["nsymbt"] == 96
    assert info["pixel_size"] == pytest.approx(1.5)
    np.testing.assert_array_equal(F.read_mrc(p), data.astype(np.float32))
    np.testing.assert_array_equal(F.read_mrc(p, 1, 2), data[1:3].astype(np.float32))


def test_mrc_reader_errors(tmp_path):
    p = tmp_path / "b.mrc"
    make_mrc(p, np.zeros((2, 3, 4)), 0)
    with pytest.raises(L.AtrError) as e:
        F.read_mrc(p)
    assert e.value.status == L.ATR_ERR_UNSUPPORTED_SIZE
    make_mrc(p, np.zeros((2, 3, 4)), 2)
    with pytest.raises(L.AtrError):
        F.read_mrc(p, 1, 2)                 # past the last section
    with open(p, "r+b") as fh:            # truncated data
        fh.truncate(1024 + 10)
    with pytest.raises(L.AtrError):
        F.read_mrc(p)
    with pytest.raises(L.AtrError):
        F.mrc_info(tmp_path / "missing.mrc")
    assert F.volume_suffix(0) == "_Vol.mrc" and F.volume_suffix(4) == "_3RD_Vol.mrc"
    assert [F.volume_suffix(i) for i in (1, 2, 3)] == ["_EVN_Vol.mrc", "_ODD_Vol.mrc", "_2ND_Vol.mrc"]
    with pytest.raises(ValueError):
        F.volume_suffix(5)


def sample_alignment(n_patches=3):
    rng = np.random.default_rng(3)
    n = 6
    local = None
    if n_patches:
        local = rng.uniform(-40, 40, (n, 5, n_patches)).astype(np.float32)
        local[:, 4] = rng.integers(0, 2, (n, n_patches))
    return F.Alignment(raw_size=(64, 48, 8), sec=np.array([2, 3, 4, 5, 7, 8], np.int32),
                       tilt_axis=rng.uniform(84, 86, n).astype(np.float32),
                       shifts=rng.uniform(-5, 5, (n, 2)).astype(np.float32),
                       tilts=np.linspace(-50, 50, n).astype(np.float32), local=local,
                       dark_frame=np.array([0, 5], np.int32), dark_sec=np.array([1, 6], np.int32),
                       dark_tilt=np.array([-60.0, 30.0], np.float32), alpha_offset=1.25,
                       beta_offset=-0.5, thickness=220)


def expected_aln_text(a):
    """The reference layout, restated with Python's printf formatting."""
    lines = ["# AreTomo Alignment / Priims bprmMn",
             "# RawSize = %d %d %d" % a.raw_size, "# NumPatches = %d" % a.n_patches]
    for fr, sc, t in zip(a.dark_frame, a.dark_sec, a.dark_tilt):
        lines.append("# DarkFrame =  %4d %4d %8.2f" % (fr, sc, t))
    lines += ["# AlphaOffset = %8.2f" % a.alpha_offset, "# BetaOffset = %8.2f" % a.beta_offset,
              "# Thickness = %d" % a.thickness,
              "# SEC     ROT         GMAG       TX          TY      SMEAN     SFIT    SCALE     BASE     TILT"]
    for i in range(a.n_tilts):
        lines.append("%5d  %9.4f  %9.5f  %9.3f  %9.3f  %7.2f  %7.2f  %7.2f  %7.2f  %8.2f" % (
            a.sec[i], a.tilt_axis[i], 1.0, a.shifts[i, 0], a.shifts[i, 1], 1.0, 1.0, 1.0, 0.0,
            a.tilts[i]))
    if a.n_patches:
        lines.append("# Local Alignment")
        for t in range(a.n_tilts):
            for p in range(a.n_patches):
                q = a.local[t, :, p]
                lines.append("%4d %3d %8.2f  %8.2f  %8.2f  %8.2f  %4.1f" % (t, p, *q))
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("n_patches", [0, 3])
def test_aln_write_layout_and_round_trip(tmp_path, n_patches):
    a = sample_alignment(n_patches)
    p = tmp_path / "ts.aln"
    F.write_aln(p, a)
    assert open(p).read() == expected_aln_text(a)
    b = F.read_aln(p)
    assert b.raw_size == a.raw_size and b.n_tilts == 6 and b.thickness == 220
    assert b.alpha_offset == pytest.approx(1.25) and b.beta_offset == pytest.approx(-0.5)
    np.testing.assert_array_equal(b.sec, a.sec)
    np.testing.assert_allclose(b.tilt_axis, a.tilt_axis, atol=5e-5)
    np.testing.assert_allclose(b.shifts, a.shifts, atol=5e-4)
    np.testing.assert_allclose(b.tilts, a.tilts, atol=5e-3)
    np.testing.assert_array_equal(b.dark_frame, a.dark_frame)
    np.testing.assert_array_equal(b.dark_sec, a.dark_sec)
    if n_patches:
        np.testing.assert_allclose(b.local[:, :4], a.local[:, :4], atol=5e-3)
        np.testing.assert_array_equal(b.local[:, 4], a.local[:, 4])
    else:
        assert b.local is None


def test_aln_loader_rules(tmp_path):
    # zero-based sections (older files) become one-based; dark secs = frame + 1;
    # short lines are skipped; good flags round at 0.5; local rows are placed
    # by their (t, p) fields
    text = """# AreTomo Alignment / Priims bprmMn
# RawSize = 32 32 4
# NumPatches = 2
# DarkFrame =     2    2   40.00
# Thickness = 90

    0    85.0000    1.00000      1.500     -2.000     1.00     1.00     1.00     0.00    -40.00
    1    85.1000    1.00000      0.500      0.250     1.00     1.00     1.00     0.00      0.00
    3    85.2000    1.00000     -1.000      3.000     1.00     1.00     1.00     0.00     20.00
   2   1    10.00     11.00      0.50      0.60   0.6
   0   0     1.00      2.00      3.00      4.00   1.0
   0   1     5.00      6.00      7.00      8.00   0.4
   1   0    -1.00     -2.00     -3.00     -4.00   1.0
   1   1    -5.00     -6.00     -7.00     -8.00   0.0
   2   0     9.00      9.50      0.10      0.20   1.0
"""
    p = tmp_path / "old.aln"
    p.write_text(text)
    a = F.read_aln(p)
    assert a.n_tilts == 3 and a.n_patches == 2 and a.thickness == 90
    np.testing.assert_array_equal(a.sec, [1, 2, 4])
    np.testing.assert_array_equal(a.dark_sec, [3])
    np.testing.assert_allclose(a.tilts, [-40, 0, 20])
    np.testing.assert_allclose(a.shifts, [[1.5, -2.0], [0.5, 0.25], [-1.0, 3.0]])
    np.testing.assert_allclose(a.local[0], [[1, 5], [2, 6], [3, 7], [4, 8], [1, 0]])
    np.testing.assert_allclose(a.local[1, 4], [1, 0])
    np.testing.assert_allclose(a.local[2, :, 1], [10, 11, 0.5, 0.6, 1])


def test_aln_loader_errors(tmp_path):
    p = tmp_path / "bad.aln"
    p.write_text("# NumPatches = 0\n    1  85.0 1.0 0 0 1 1 1 0 10\n")
    with pytest.raises(L.AtrError):            # no RawSize
        F.read_aln(p)
    p.write_text("# RawSize = 8 8 3\n    1  85.0 1.0 0 0 1 1 1 0 10\n"
                 "    2  85.0 1.0 0 0 1 1 1 0 20\n")
    with pytest.raises(L.AtrError):            # 2 of 3 global rows
        F.read_aln(p)
    p.write_text("# RawSize = 8 8 2\n# NumPatches = 2\n    1  85.0 1.0 0 0 1 1 1 0 10\n"
                 "    2  85.0 1.0 0 0 1 1 1 0 20\n   0   0  1 2 3 4 1\n")
    with pytest.raises(L.AtrError):            # 1 of 4 local rows
        F.read_aln(p)
    with pytest.raises(L.AtrError):
        F.read_aln(tmp_path / "missing.aln")


def test_tilt_file_reader(tmp_path):
    p = tmp_path / "ts_TLT.txt"
    p.write_text("# header\n\n-60.0\n-57.0  3\n-54.0  5  12.5\n  -51.0 7 9.0\n-48.0\n")
    t, a, d = F.read_tilt_file(p, 4)
    np.testing.assert_allclose(t, [-60, -57, -54, -51])
    np.testing.assert_array_equal(a, [0, 3, 5, 7])
    np.testing.assert_allclose(d, [0, 0, 12.5, 9.0])
    with pytest.raises(L.AtrError):
        F.read_tilt_file(p, 6)
    (tmp_path / "ts.rawtlt").write_text("1\n2\n")
    assert F.find_tilt_file(tmp_path / "ts").endswith("_TLT.txt")
    p.unlink()
    assert F.find_tilt_file(tmp_path / "ts").endswith(".rawtlt")


def test_driver_helpers():
    t = np.array([-60, -30, -3, 0, 3, 30, 60], np.float32)
    assert F.frame_index_from_tilt(t, -90.0) == 0 and F.frame_index_from_tilt(t, 90.0) == 6
    assert F.frame_index_from_tilt(t, 1.5) == 3        # tie keeps the first
    assert F.binned_vol_z(None, 150, 300, 4.0) == 112
    assert F.binned_vol_z(-1, 60, 0, 1.0) == 100       # at least 100
    assert F.binned_vol_z(301, 0, 300, 2.0) == 150
    assert F.aligned_size(64, 48, 85.0) == (48, 64)


def test_second_resolution_vol_z():
    # mRecon2nd: VolZ or thickness + ExtZ, at least 100 before binning,
    # then /bin and even, no floor after binning
    assert F.second_vol_z(None, 150, 300, 4.0) == 112
    assert F.second_vol_z(50, 0, 0, 8.0) == 12          # 100 / 8 -> 12 (skipped by the driver: <= 16)
    assert F.second_vol_z(0, 40, 20, 2.0) == 50         # 60 -> 100 -> 50
    assert F.second_vol_z(301, 0, 0, 3.0) == 100


def test_reconstruct_refuses_before_touching_a_device(tmp_path):
    # too few tilt images (CAreTomoMain::mCheckTiltSeries), a stack that does
    # not match the .aln, and a missing .aln all fail on the host
    a = sample_alignment(0)
    F.write_aln(tmp_path / "ts.aln", a)
    F.write_mrc(tmp_path / "few.mrc", np.zeros((5, 48, 64), np.float32))
    with pytest.raises(L.AtrError, match="too few"):
        F.reconstruct_all(tmp_path / "few.mrc", tmp_path / "ts.aln", tmp_path / "out")
    F.write_mrc(tmp_path / "ts.mrc", np.zeros((7, 48, 64), np.float32))   # .aln says 8
    with pytest.raises(L.AtrError, match="sections"):
        F.reconstruct_all(tmp_path / "ts.mrc", tmp_path / "ts.aln", tmp_path / "out")
    with pytest.raises(L.AtrError):
        F.reconstruct_all(tmp_path / "ts.mrc", tmp_path / "none.aln", tmp_path / "out")
