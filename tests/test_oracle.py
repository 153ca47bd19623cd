"""Pins the FP64 oracle (CPU only).

The reference ships no golden vectors for its reconstruction path and cannot
be built here (CUDA + cuFFT + cuRAND), so every oracle function is pinned by
something the mathematics fixes: published known-answer values, closed forms
on special geometries, numpy as an independent FFT, invariants, and
hand-worked cases.  Each test is chosen so that a plausible slip (a dropped
term, a wrong sign, a transposed index, a wrong normalisation) fails it.
"""
import numpy as np
import pytest

import oracle

SENT = -1e30


def f32(x):
    return float(np.float32(x))


def cs_of(tilts):
    return oracle.cos_sin(np.asarray(tilts, np.float32))


# ---------------------------------------------------------------------------
# small helpers
# ---------------------------------------------------------------------------
def test_splitmix64_published_vectors():
    # SplitMix64 (Steele, Lea, Flood 2014; Vigna's reference code) seeded with 0
    # produces these three outputs; orc_splitmix64(x) is one step from state x.
    g = 0x9E3779B97F4A7C15
    want = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    got = [oracle.splitmix64((i * g) % 2**64) for i in range(3)]
    assert got == want


def test_fill_table_range_and_uniformity():
    n = 4096
    t = oracle.fill_table(12345, n)
    assert t.min() >= 0 and t.max() < n
    assert np.array_equal(t, oracle.fill_table(12345, n))
    assert not np.array_equal(t, oracle.fill_table(12346, n))
    # 16 equal bins: chi-square with 15 dof stays far below 60
    h = np.bincount(t * 16 // n, minlength=16)
    chi2 = ((h - n / 16) ** 2 / (n / 16)).sum()
    assert chi2 < 60


def test_cos_sin_uses_float_radian_constant():
    tilts = np.array([0.0, 30.0, -45.0, 60.0, 89.0], np.float32)
    cs = cs_of(tilts)
    assert cs[0] == 1.0 and cs[1] == 0.0
    # angle = 0.017453f * tilt in float, not pi/180
    ang = (np.float32(0.017453) * tilts).astype(np.float32).astype(np.float64)
    assert np.array_equal(cs[0::2], np.cos(ang).astype(np.float32))
    assert np.array_equal(cs[1::2], np.sin(ang).astype(np.float32))
    # 0.017453 is not pi/180: sin(30 deg) differs from 0.5 by ~5e-6
    assert abs(cs[3] - 0.5) > 1e-6
    assert abs(cs[3] - 0.5) < 2e-5


@pytest.mark.parametrize("n,s,sizes", [(61, 5, [13, 12, 12, 12, 12]), (10, 3, [4, 3, 3]),
                                       (8, 8, [1] * 8), (7, 1, [7])])
def test_split_items(n, s, sizes):
    starts, got = oracle.split_items(n, s)
    assert got.tolist() == sizes
    assert starts.tolist() == np.concatenate([[0], np.cumsum(sizes)[:-1]]).tolist()


@pytest.mark.parametrize("order", [0, 1])
def test_subset_tilts_partition(order):
    start, n, ns = 3, 41, 5
    all_t = np.concatenate([oracle.subset_tilts(start, n, ns, order, s) for s in range(ns)])
    assert sorted(all_t.tolist()) == list(range(start, start + n))
    if order == 0:
        assert oracle.subset_tilts(start, n, ns, 0, 0).tolist() == list(range(3, 12))
    else:
        assert oracle.subset_tilts(start, n, ns, 1, 2).tolist() == list(range(5, 44, 5))


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 12, 30, 64, 100, 210, 243])
def test_fft_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    np.testing.assert_allclose(oracle.fft(x, -1), np.fft.fft(x), rtol=0, atol=1e-11 * n)
    np.testing.assert_allclose(oracle.fft(x, +1), n * np.fft.ifft(x), rtol=0, atol=1e-11 * n)


def test_aligned_size_uses_3_14():
    # |sin(axis * 3.14 / 180)| > 0.707 swaps X and Y (CCorrectUtil.cpp:11-21).
    # With 3.14 sin(45 deg) = 0.70683 < 0.707: no swap (with pi it would swap).
    assert oracle.aligned_size(100, 80, 0.0) == (100, 80)
    assert oracle.aligned_size(100, 80, 45.0) == (100, 80)
    assert oracle.aligned_size(100, 80, 46.0) == (80, 100)
    assert oracle.aligned_size(100, 80, 85.0) == (80, 100)
    assert oracle.aligned_size(100, 80, -85.0) == (80, 100)
    assert oracle.aligned_size(100, 80, 180.0) == (100, 80)


# ---------------------------------------------------------------------------
# weighting
# ---------------------------------------------------------------------------
def test_weight_projs_closed_form():
    rng = np.random.default_rng(1)
    tilts = [-50.0, 0.0, 35.0]
    cs = cs_of(tilts)
    sino = rng.standard_normal((3, 18))
    out = oracle.weight_projs(sino, 16, cs, 40)
    want = sino.copy()
    want[:, :16] *= (cs[0::2].astype(np.float64) / 40.0)[:, None]
    np.testing.assert_array_equal(out, want)


def _ramp_weight(n):
    r = np.arange(n // 2 + 1) / n
    return 2 * r * (0.55 + 0.45 * np.cos(np.float64(np.float32(6.2831852)) * r)) / n


@pytest.mark.parametrize("n", [16, 30, 60])
def test_rweight_cosines_are_eigenfunctions(n):
    # a cosine of bin k is scaled by N * w(k) = 2 (k/N) (0.55 + 0.45 cos(6.2831852 k/N))
    x = np.arange(n)
    rows = np.array([np.cos(2 * np.pi * k * x / n) for k in range(n // 2 + 1)])
    out = oracle.rweight(rows, n)
    scale = n * _ramp_weight(n)
    np.testing.assert_allclose(out, rows * scale[:, None], atol=1e-12)


def test_rweight_vs_numpy_and_invariants():
    n = 48
    rng = np.random.default_rng(2)
    rows = rng.standard_normal((4, n + 2))
    out = oracle.rweight(rows, n)
    want = np.fft.irfft(np.fft.rfft(rows[:, :n], axis=1) * _ramp_weight(n), n, axis=1) * n
    np.testing.assert_allclose(out[:, :n], want, atol=1e-12)
    np.testing.assert_array_equal(out[:, n:], rows[:, n:])   # padding untouched
    # constant rows vanish (DC weight 0); the impulse response is even
    np.testing.assert_allclose(oracle.rweight(np.full((1, n), 3.0), n), 0, atol=1e-13)
    d = np.zeros((1, n))
    d[0, 0] = 1
    h = oracle.rweight(d, n)[0]
    np.testing.assert_allclose(h[1:], h[1:][::-1], atol=1e-14)


# ---------------------------------------------------------------------------
# back-projection (SART)
# ---------------------------------------------------------------------------
def test_back_proj_tilt0_is_row_copy():
    p, vz = 16, 6
    rng = np.random.default_rng(3)
    row = rng.standard_normal(p + 2)
    vol0 = rng.standard_normal((vz, p))
    v = oracle.back_proj(row[None], p, cs_of([0.0]), [0], vol0, positivity=0, relax=0.5)
    want = vol0.copy()
    want[:, :p - 1] += 0.5 * row[:p - 1]          # u = x, included iff x <= P-2
    np.testing.assert_allclose(v, want, atol=1e-15)


def test_back_proj_linear_sinogram_general_tilts():
    # the interpolation of a linear row is exact, so every included voxel gets
    # a + b*u with u = fx*cos + fz*sin + P/2; excluded voxels are unchanged
    p, vx, vz = 40, 40, 24
    a, b = 1.5, 0.25
    tilts = [-57.0, -13.0, 22.0, 48.0]
    cs = cs_of(tilts)
    sino = np.tile(a + b * np.arange(p + 2, dtype=np.float64), (4, 1))
    vol0 = np.full((vz, vx), 7.0)
    fx = np.arange(vx) - vx * 0.5
    fz = (np.arange(vz) - vz * 0.5)[:, None]
    for t in range(4):
        c, s = float(cs[2 * t]), float(cs[2 * t + 1])
        v = oracle.back_proj(sino, p, cs, [t], vol0, positivity=0)
        u = fx * c + fz * s + p / 2.0
        inc = (u >= 0) & (u <= p - 2)
        assert inc.sum() > 0 and (~inc).sum() > 0
        want = np.where(inc, 7.0 + a + b * u, 7.0)
        np.testing.assert_allclose(v, want, atol=1e-12)
    # all four at once: the mean of the included tilts
    v = oracle.back_proj(sino, p, cs, [0, 1, 2, 3], vol0, positivity=0, relax=0.3)
    acc = np.zeros((vz, vx))
    cnt = np.zeros((vz, vx))
    for t in range(4):
        u = fx * float(cs[2 * t]) + fz * float(cs[2 * t + 1]) + p / 2.0
        inc = (u >= 0) & (u <= p - 2)
        acc += np.where(inc, a + b * u, 0)
        cnt += inc
    want = np.where(cnt > 0, 7.0 + 0.3 * acc / np.maximum(cnt, 1), 7.0)
    np.testing.assert_allclose(v, want, atol=1e-12)


def test_back_proj_positivity_sentinels_exclusion():
    p, vz = 8, 2
    row = np.full((2, p + 2), -2.0)
    row[1] = 5.0
    cs = cs_of([0.0, 0.0])
    vol0 = np.zeros((vz, p))
    # positivity clamps, without it the value stays negative
    assert oracle.back_proj(row, p, cs, [0], vol0, positivity=1)[0, 0] == 0.0
    assert oracle.back_proj(row, p, cs, [0], vol0, positivity=0)[0, 0] == -2.0
    # mean of two tilts; an excluded tilt does not count
    assert oracle.back_proj(row, p, cs, [0, 1], vol0, positivity=0)[0, 0] == 1.5
    ex = np.array([0, 1], np.uint8)
    assert oracle.back_proj(row, p, cs, [0, 1], vol0, positivity=0, no_proj=ex)[0, 0] == -2.0
    # empty samples (<= -1e10) are skipped; at 0 deg the interpolation weight
    # of the right neighbour is exactly 0, so only u = 3 itself is empty
    r2 = row.copy()
    r2[0, 3] = SENT
    v = oracle.back_proj(r2, p, cs, [0, 1], vol0, positivity=0)
    assert v[0, 3] == 5.0 and v[0, 2] == 1.5 and v[0, 4] == 1.5
    # a voxel already empty stays empty: -1e30 + x <= -1e10 -> -1e30
    vs = vol0.copy()
    vs[1, 4] = SENT
    assert oracle.back_proj(row, p, cs, [1], vs, positivity=0)[1, 4] == SENT


# ---------------------------------------------------------------------------
# forward projection, difference
# ---------------------------------------------------------------------------
def test_for_proj_constant_volume():
    vx, vz, p = 32, 12, 32
    cs = cs_of([-60.0, -20.0, 0.0, 40.0])
    fwd = oracle.for_proj(np.full((vz, vx), 2.5), cs, [0, 1, 2, 3], p, p + 2, 4)
    hit = fwd[:, :p] > -1e10
    np.testing.assert_allclose(fwd[:, :p][hit], 2.5, atol=1e-13)
    assert fwd[2, p - 1] == SENT           # at 0 deg the last ray misses [0, VolX-2]
    assert hit[2, :p - 1].all()
    assert np.all(fwd[:, p:] == 0)         # padding untouched


def test_for_proj_tilt0_linear_volume():
    # at 0 deg: L = VolZ + 1, x = xd and z = k - 0.5; the samples inside
    # [0, VolZ-2] are z = 0.5 .. VolZ-2.5 (VolZ-2 of them), mean z = (VolZ-2)/2
    vx, vz = 20, 10
    a, b, d = 0.5, 0.3, -0.7
    vol = a + b * np.arange(vx)[None, :] + d * np.arange(vz)[:, None]
    fwd = oracle.for_proj(vol, cs_of([0.0]), [0], vx, vx, 1)[0]
    want = a + b * np.arange(vx - 1) + d * (vz - 2) / 2.0
    np.testing.assert_allclose(fwd[:vx - 1], want, atol=1e-12)


def test_for_proj_ray_is_a_line_at_general_tilts():
    # v = cos*x + sin*z is constant along the ray direction (-sin, cos), so every
    # ray with a valid sample returns xd - P/2 + cos*VolX/2 + sin*VolZ/2
    # whatever the clipping (pins the sign of sin and the ray centre).
    vx, vz, p = 48, 20, 48
    tilts = [-62.0, -31.0, 17.0, 55.0]
    cs = cs_of(tilts)
    for t in range(4):
        c, s = float(cs[2 * t]), float(cs[2 * t + 1])
        vol = 1.25 + c * np.arange(vx)[None, :] + s * np.arange(vz)[:, None]
        fwd = oracle.for_proj(vol, cs, [t], p, p, 4)[t]
        xd = np.arange(p)
        want = 1.25 + xd - p / 2.0 + c * vx / 2.0 + s * vz / 2.0
        hit = fwd > -1e10
        assert hit.sum() > p // 2
        np.testing.assert_allclose(fwd[hit], want[hit], atol=1e-10)


def test_for_proj_single_voxel_bilinear_weights():
    # one non-zero voxel at (x0, z0): at 0 deg the ray xd samples z = k - 0.5
    # and x = xd exactly, so the only non-zero ray is xd = x0 with value
    # (z-weights of row z0: 0.5 + 0.5) / (VolZ - 2 samples)
    vx, vz = 12, 8
    vol = np.zeros((vz, vx))
    vol[4, 5] = 1.0
    fwd = oracle.for_proj(vol, cs_of([0.0]), [0], vx, vx, 1)[0]
    assert fwd[5] == pytest.approx(1.0 / (vz - 2), rel=1e-14)
    assert np.all(fwd[[0, 1, 2, 3, 4, 6, 7, 8, 9, 10]] == 0)


def test_for_proj_skips_empty_voxels():
    vx, vz = 10, 6
    vol = np.full((vz, vx), 1.0)
    vol[2, 4] = SENT   # every bilinear sample touching it is < -1e10 and skipped
    fwd = oracle.for_proj(vol, cs_of([0.0]), [0], vx, vx, 1)[0]
    assert fwd[4] == 1.0 and fwd[3] == 1.0


def test_diff_proj():
    raw = np.array([[1.0, 2.0, SENT, 4.0, 9.0]])
    fwd = np.array([[0.5, SENT, 1.0, 1.0, 9.0]])
    out = oracle.diff_proj(raw, fwd, [0], 4)
    np.testing.assert_array_equal(out, [[0.5, 0.0, 0.0, 3.0, 9.0]])


# ---------------------------------------------------------------------------
# empty fill
# ---------------------------------------------------------------------------
def test_fill_empty_hand_case():
    # 4x4, empties at 1, 5, 6 and 15; table chosen by hand:
    #  1 -> 5 (empty) -> 6 (empty) -> 0: takes img[0]
    #  5 -> 6 (empty) -> 0: img[0]
    #  6 -> 0: img[0]... except 6 -> 3 below
    #  15 -> 15 -> ... (cycle of empties, 100 hops) -> mean of the non-empty
    img = np.arange(16, dtype=np.float64).reshape(4, 4) + 10
    for i in (1, 5, 6, 15):
        img.flat[i] = SENT
    table = np.arange(16, dtype=np.int32)
    table[1], table[5], table[6], table[15] = 5, 6, 3, 15
    out = oracle.fill_empty(img, 4, table)
    mean = np.mean([10 + i for i in range(16) if i not in (1, 5, 6, 15)])
    assert out.flat[6] == 13.0
    assert out.flat[5] == 13.0
    assert out.flat[1] == 13.0
    assert out.flat[15] == pytest.approx(mean)
    keep = [i for i in range(16) if i not in (1, 5, 6, 15)]
    np.testing.assert_array_equal(out.flat[keep], img.flat[keep])


def test_fill_empty_chain_limit_and_all_empty():
    # a chain of 101 empties before a full pixel exceeds the 100-hop limit
    n = 128
    img = np.full((1, n), SENT)
    img[0, 120] = 4.0
    img[0, 121] = 8.0
    table = np.array([(i + 1) % n for i in range(n)], np.int32)
    out = oracle.fill_empty(img, n, table)
    assert out[0, 119] == 4.0          # 1 hop
    assert out[0, 20] == 4.0           # 100 hops: 20 -> 21 ... -> 120 at hop 100
    assert out[0, 19] == 6.0           # would need 101: the mean of {4, 8}
    assert oracle.fill_empty(np.full((1, 4), SENT), 4, np.zeros(4, np.int32)).max() == 0.0


# ---------------------------------------------------------------------------
# WBP
# ---------------------------------------------------------------------------
def test_back_proj_wbp_tilt0_linear_row():
    # on the 2x grid at 0 deg: u = (x - VolX)/2 + P/2 + 0.5 = x/2 + 0.5 (VolX = P)
    p = 16
    a, b = 2.0, 0.5
    row = (a + b * np.arange(p + 2, dtype=np.float64))[None]
    img2 = oracle.back_proj_wbp(row, p, cs_of([0.0]), 2 * p, 6)
    u = np.arange(2 * p) / 2.0 + 0.5
    want = np.where(u <= p - 2, a + b * u, 0.0)
    np.testing.assert_allclose(img2, np.tile(want, (6, 1)), atol=1e-13)


def test_back_proj_wbp_general_tilt_and_mean():
    p, vx2, vz2 = 24, 48, 20
    a, b = -1.0, 0.125
    cs = cs_of([-40.0, 25.0])
    sino = np.tile(a + b * np.arange(p + 2, dtype=np.float64), (2, 1))
    img2 = oracle.back_proj_wbp(sino, p, cs, vx2, vz2)
    fx = np.arange(vx2) - vx2 / 2.0
    fz = (np.arange(vz2) - vz2 / 2.0)[:, None]
    acc = np.zeros((vz2, vx2))
    cnt = np.zeros((vz2, vx2))
    for t in range(2):
        u = (fx * float(cs[2 * t]) + fz * float(cs[2 * t + 1])) / 2 + p // 2 + 0.5
        inc = (u >= 0) & (u <= p - 2)
        acc += np.where(inc, a + b * u, 0)
        cnt += inc
    want = np.where(cnt > 0, acc / np.maximum(cnt, 1), 0.0)
    assert (cnt == 1).sum() > 0 and (cnt == 2).sum() > 0
    np.testing.assert_allclose(img2, want, atol=1e-12)


@pytest.mark.parametrize("vx,vz", [(8, 6), (16, 10), (12, 12)])
def test_wbp_crop_vs_numpy(vx, vz):
    rng = np.random.default_rng(vx * vz)
    img2 = rng.standard_normal((2 * vz, 2 * vx))
    f = np.fft.rfft2(img2) / (4.0 * vx * vz)
    rows = [j if j <= vz // 2 else j + vz for j in range(vz)]
    g = f[rows][:, :vx // 2 + 1]
    want = np.fft.irfft2(g, s=(vz, vx)) * (vx * vz)
    np.testing.assert_allclose(oracle.wbp_crop(img2, vx, vz), want, atol=1e-12)
    # DC is preserved: a constant 2x grid stays the same constant
    np.testing.assert_allclose(oracle.wbp_crop(np.full((2 * vz, 2 * vx), 3.0), vx, vz), 3.0,
                               atol=1e-12)


def test_tomo_wbp_composition_and_linearity():
    p, vx, vz = 32, 32, 12
    rng = np.random.default_rng(4)
    cs = cs_of([-30.0, 0.0, 30.0])
    s1 = rng.standard_normal((3, p + 2))
    s2 = rng.standard_normal((3, p + 2))
    v1, _ = oracle.tomo_wbp(s1, p, cs, vx, vz)
    v2, _ = oracle.tomo_wbp(s2, p, cs, vx, vz)
    v12, _ = oracle.tomo_wbp(2.0 * s1 - 0.5 * s2, p, cs, vx, vz)
    np.testing.assert_allclose(v12, 2.0 * v1 - 0.5 * v2, atol=1e-12)
    # weight -> ramp filter -> 2x-grid BP -> crop, step by step
    w = oracle.weight_projs(s1, p, cs, vz)
    w = oracle.rweight(w, p)
    img2 = oracle.back_proj_wbp(w, p, cs, 2 * vx, 2 * vz)
    np.testing.assert_allclose(v1, oracle.wbp_crop(img2, vx, vz), atol=1e-13)


# ---------------------------------------------------------------------------
# SART
# ---------------------------------------------------------------------------
def test_tomo_sart_composition():
    # Recon/CTomoSart.cpp:54-125 step by step, written from the pinned pieces
    p, vx, vz, n = 24, 24, 8, 5
    rng = np.random.default_rng(5)
    tilts = [-40.0, -20.0, 0.0, 20.0, 40.0]
    cs = cs_of(tilts)
    sino = rng.random((n, p + 2)) + 0.5
    table = oracle.fill_table(9, vx * vz)
    got, _ = oracle.tomo_sart(sino, p, cs, np.zeros((vz, vx)), 0, n, 2, 2, relax=-1.0,
                              order=0, positivity=1, fill_tbl=table)
    s = oracle.weight_projs(sino, p, cs, vz)
    v = oracle.back_proj(s, p, cs, list(range(n)), np.zeros((vz, vx)), positivity=1)
    for _ in range(2):
        for sub in ([0, 1, 2], [3, 4]):
            fwd = oracle.for_proj(v, cs, sub, p, p + 2, n)
            d = oracle.diff_proj(s, fwd, sub, p)
            v = oracle.back_proj(d, p, cs, sub, v, positivity=1, relax=0.5)
    v = oracle.fill_empty(v, vx, table)
    np.testing.assert_allclose(got, v, atol=1e-14)


def test_tomo_sart_reduces_residual():
    # consistent data (the weighted sinogram equals the forward projection of a
    # known slice): the residual over the rays that hit the slice falls with the
    # number of iterations
    p, vx, vz = 32, 32, 12
    tilts = np.arange(-60, 61, 10, dtype=np.float32)
    n = len(tilts)
    cs = cs_of(tilts)
    truth = np.zeros((vz, vx))
    truth[4:8, 10:20] = 1.0
    fwd0 = oracle.for_proj(truth, cs, list(range(n)), p, p, n)
    hit = fwd0 > -1e10
    sino = np.zeros((n, p + 2))
    sino[:, :p] = np.where(hit, fwd0, 0.0) * vz / cs[0::2].astype(np.float64)[:, None]
    res = []
    for it in (1, 4, 16):
        v, _ = oracle.tomo_sart(sino, p, cs, np.zeros((vz, vx)), 0, n, it, 3, relax=-1.0,
                                order=1, positivity=1)
        f = oracle.for_proj(v, cs, list(range(n)), p, p, n)
        res.append(np.linalg.norm((f - fwd0)[hit]))
    assert res[0] > res[1] > res[2]
    assert res[2] < 0.5 * res[0]


# ---------------------------------------------------------------------------
# pre-stage
# ---------------------------------------------------------------------------
def test_corr_proj_axis0_integer_shift_is_translation():
    # out[y][x] = raw[y + 1 + sy][x + 1 + sx] (pixel centres at +0.5, then
    # (int)(v + 0.5)), sentinel outside
    rx, ry = 10, 7
    raw = (np.arange(rx * ry, dtype=np.float32) + 1).reshape(ry, rx)
    for sx, sy in ((0, 0), (2, -1), (-3, 1)):
        out = oracle.corr_proj(raw, 0.0, (float(sx), float(sy)), rx, ry, rx + 2)
        want = np.full((ry, rx + 2), 0.0)
        for y in range(ry):
            for x in range(rx):
                ix, iy = x + 1 + sx, y + 1 + sy
                want[y, x] = raw[iy, ix] if 0 <= ix < rx and 0 <= iy < ry else SENT
        np.testing.assert_array_equal(out, want)


def test_corr_proj_axis90_is_permutation():
    # axis 90 with shift 0.25: the index values are k + 0.25, so the float
    # radian constant (cos(90 deg) = -4.4e-8) cannot move a truncation; the map
    # is out[y][x] = raw[x + 1][raw_x - y], sentinel where that leaves the image
    rx, ry = 9, 6
    ox, oy = oracle.aligned_size(rx, ry, 90.0)
    assert (ox, oy) == (ry, rx)
    raw = (np.arange(rx * ry, dtype=np.float32) + 1).reshape(ry, rx)
    out = oracle.corr_proj(raw, 90.0, (0.25, 0.25), ox, oy, ox)
    for y in range(oy):
        for x in range(ox):
            ix, iy = rx - y, x + 1
            want = raw[iy, ix] if 0 <= ix < rx and 0 <= iy < ry else SENT
            assert out[y, x] == want, (y, x)
    assert np.all(out[0] == SENT) and np.all(out[:, ox - 1] == SENT)


def test_corr_proj_truncates_toward_zero():
    # shift -1.75 at axis 0: ix = (int)(x - 0.75); (int)(-0.75) = 0, so pixels
    # 0 and 1 both read raw column 0 (a floor would leave pixel 0 empty)
    rx, ry = 6, 3
    raw = (np.arange(rx * ry, dtype=np.float32) + 1).reshape(ry, rx)
    out = oracle.corr_proj(raw, 0.0, (-1.75, 0.0), rx, ry, rx)
    np.testing.assert_array_equal(out[:2, 0], raw[1:, 0])
    np.testing.assert_array_equal(out[:2, 1], raw[1:, 0])
    np.testing.assert_array_equal(out[:2, 2], raw[1:, 1])


def test_prestage_stack_and_fill():
    rx, ry, n = 12, 10, 3
    rng = np.random.default_rng(6)
    raw = rng.random((n, ry, rx)).astype(np.float32)
    axes = np.array([0.0, 0.0, 0.0], np.float32)
    shifts = np.array([[0, 0], [3, 0], [0, -2]], np.float32)
    a = oracle.prestage(raw, axes, shifts, random_fill=0)
    np.testing.assert_array_equal(a[0, :ry - 1, :rx - 1], raw[0, 1:, 1:])
    assert np.all(a[1, :, rx - 4:] == SENT)
    f = oracle.prestage(raw, axes, shifts, random_fill=1, seed=11)
    assert f.min() > -1e10
    table = oracle.fill_table(11, rx * ry)
    for t in range(n):
        np.testing.assert_array_equal(f[t], oracle.fill_empty(a[t], rx, table))


# ---------------------------------------------------------------------------
# per-series drivers and layout
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("px", [24, 25])
def test_recon_slices_driver(px):
    n, py, vz = 4, 5, 8
    rng = np.random.default_rng(px)
    aligned = rng.random((n, py, px)) + 0.2
    tilts = np.array([-30, -10, 10, 30], np.float32)
    vx = px // 2 * 2
    pad = (px // 2 + 1) * 2
    proj = (pad // 2 - 1) * 2
    cs = cs_of(tilts)
    slices = [0, 3]
    seed = 77
    out_s = oracle.recon_slices(aligned, tilts, 1, vz, slices, iters=2, n_subsets=2, order=1,
                                seed=seed)
    out_w = oracle.recon_slices(aligned, tilts, 0, vz, slices)
    table = oracle.fill_table(seed ^ 0x5A5A5A5A5A5A5A5A, vx * vz)
    for k, y in enumerate(slices):
        sino = np.zeros((n, pad))
        sino[:, :px] = aligned[:, y, :]
        v, _ = oracle.tomo_sart(sino, proj, cs, np.zeros((vz, vx)), 0, n, 2, 2, order=1,
                                fill_tbl=table)
        np.testing.assert_allclose(out_s[k], v[::-1], atol=1e-14)     # z-flipped
        w, _ = oracle.tomo_wbp(sino, proj, cs, vx, vz)
        np.testing.assert_allclose(out_w[k], w[::-1], atol=1e-14)


def test_flip_vol_indices():
    vx, vz, vy = 3, 4, 5
    xzy = np.arange(vx * vz * vy, dtype=np.float64).reshape(vy, vz, vx)
    flipped = oracle.flip_vol(xzy, 1)
    plain = oracle.flip_vol(xzy, 0)
    assert flipped.shape == (vz, vy, vx)
    for y in range(vy):
        for z in range(vz):
            np.testing.assert_array_equal(plain[z, y], xzy[y, z])
            np.testing.assert_array_equal(flipped[vz - 1 - z, y], xzy[y, z])


# ---------------------------------------------------------------------------
# local patch shifts and Fourier binning (pre-stage extensions)
# ---------------------------------------------------------------------------
def _patches(cx, cy, sx, sy, good):
    return np.array([cx, cy, sx, sy, good], np.float32)


def test_local_shift_uniform_patches_equal_a_global_shift():
    # every good patch shifted alike: the weighted mean is that shift, whatever
    # the weights, so the result is the global-shift-only image with shift + s
    rx, ry = 24, 18
    rng = np.random.default_rng(21)
    raw = rng.random((ry, rx)).astype(np.float32)
    cx, cy = [-6.0, 0.0, 7.0, 3.0], [-4.0, 5.0, 2.0, -7.0]
    loc = _patches(cx, cy, [2, 2, 2, 2], [-1, -1, -1, -1], [1, 1, 1, 1])
    a = oracle.corr_proj(raw, 30.0, (0.5, 1.0), rx, ry, rx, local=loc)
    b = oracle.corr_proj(raw, 30.0, (2.5, 0.0), rx, ry, rx)
    np.testing.assert_array_equal(a, b)


def test_local_shift_bad_patches_and_single_patch():
    rx, ry = 20, 16
    raw = (np.arange(rx * ry, dtype=np.float32) + 1).reshape(ry, rx)
    # only patch 1 is good: its shift applies everywhere
    loc = _patches([0.0, 5.0], [0.0, -3.0], [7.0, 3.0], [7.0, -2.0], [0.0, 1.0])
    a = oracle.corr_proj(raw, 0.0, (0.0, 0.0), rx, ry, rx, local=loc)
    b = oracle.corr_proj(raw, 0.0, (3.0, -2.0), rx, ry, rx)
    np.testing.assert_array_equal(a, b)
    # no good patch: no local shift (good < 0.9 counts as bad)
    loc0 = _patches([0.0], [0.0], [5.0], [5.0], [0.89])
    np.testing.assert_array_equal(oracle.corr_proj(raw, 0.0, (0.0, 0.0), rx, ry, rx, local=loc0),
                                  oracle.corr_proj(raw, 0.0, (0.0, 0.0), rx, ry, rx))


def test_local_shift_two_patch_weights():
    # axis 0, pixel centre (xc, yc) = (x + 0.5 - W/2, y + 0.5 - H/2); patches
    # at (-5, 0) with shift +4 px and (5, 0) with shift 0: at xc = -5 the
    # weights are 1 and exp(-100 (10/W)^2), so the shift is 4 / (1 + e)
    rx, ry = 40, 10
    raw = (np.arange(rx * ry, dtype=np.float32)).reshape(ry, rx) * 0 + \
        np.arange(rx, dtype=np.float32)[None, :]
    loc = _patches([-5.0, 5.0], [0.0, 0.0], [4.0, 0.0], [0.0, 0.0], [1.0, 1.0])
    out = oracle.corr_proj(raw, 0.0, (0.0, 0.0), rx, ry, rx, local=loc)
    x = 14                       # xc = 14.5 - 20 = -5.5
    xc = x + 0.5 - rx / 2
    w1 = np.exp(-100 * ((xc + 5) / rx) ** 2)
    w2 = np.exp(-100 * ((xc - 5) / rx) ** 2)
    ls = 4 * w1 / (w1 + w2)
    ix = int(xc + ls + rx / 2 + 0.5)
    assert out[3, x] == raw[3, ix]
    assert ix == 18


@pytest.mark.parametrize("ax,ay,bx,by", [(32, 24, 16, 12), (30, 40, 10, 20), (64, 64, 24, 32)])
def test_fourier_crop_vs_numpy(ax, ay, bx, by):
    rng = np.random.default_rng(ax + ay)
    img = rng.standard_normal((ay, ax))
    f = np.fft.rfft2(img) / (ax * ay)
    rows = [j if j <= by // 2 else j - by + ay for j in range(by)]
    want = np.fft.irfft2(f[rows][:, :bx // 2 + 1], s=(by, bx)) * (bx * by)
    np.testing.assert_allclose(oracle.fourier_crop(img, bx, by), want, atol=1e-12)


def test_bin_stack_sizes_and_dc():
    assert oracle.bin_size(4096, 4096, 4.0) == (1024, 1024)
    assert oracle.bin_size(1000, 998, 3.0) == (332, 332)   # (int)(x / 3) / 2 * 2
    assert oracle.bin_size(100, 80, 1.0) == (100, 80)
    # a constant image keeps its value (normalised forward, plain inverse)
    st = np.full((2, 24, 36), 2.5)
    np.testing.assert_allclose(oracle.bin_stack(st, 2.0), 2.5, atol=1e-12)
    # a cosine below the new Nyquist survives binning unchanged in amplitude
    x = np.arange(36)
    img = np.tile(np.cos(2 * np.pi * 3 * x / 36), (24, 1))[None]
    out = oracle.bin_stack(img, 2.0)[0]
    xb = np.arange(18)
    np.testing.assert_allclose(out, np.tile(np.cos(2 * np.pi * 3 * xb / 18), (12, 1)), atol=1e-12)


# ---- ProjAlign re-projector (CCalcReproj / GReproj) ------------------------
def test_reproj_range_cases():
    t = np.arange(-60, 61, 3, dtype=np.float32)
    idx = {float(a): i for i, a in enumerate(t)}
    # |a - t| <= 20.5 with the angles at and above a skipped: 12..27 deg
    assert oracle.reproj_range(t, idx[30.0], (np.abs(t) >= 30).astype(np.uint8)) == (idx[12.0], idx[27.0])
    # cos stretch <= 1.2 at 60 deg: t >= 53.13 deg
    assert oracle.reproj_range(t, idx[60.0], (t >= 60).astype(np.uint8)) == (idx[54.0], idx[57.0])
    # at most 10 tilts: 1-degree steps, 40 deg: 24..39 capped to 24..33
    t1 = np.arange(-60, 61, 1, dtype=np.float32)
    assert oracle.reproj_range(t1, 100, (t1 >= 40).astype(np.uint8)) == (84, 93)
    # nothing qualifies: the neighbour one step toward 0 deg
    t5 = np.array([-60, -30, 0, 30, 60], np.float32)
    assert oracle.reproj_range(t5, 4, np.array([0, 0, 0, 0, 1], np.uint8)) == (3, 3)
    assert oracle.reproj_range(t5, 0, np.array([1, 0, 0, 0, 0], np.uint8)) == (1, 1)


def test_reproj_zero_tilt_is_identity():
    # one tilt at 0 deg re-projected at 0 deg: the BP voxel column x + px/2
    # holds proj[x] for every z, the FP ray reads it vol_z times; the last
    # column falls outside u <= px - 1 and is empty
    rng = np.random.default_rng(4)
    p = rng.uniform(1, 2, (1, 5, 32)).astype(np.float32)
    r = oracle.reproj(p, np.array([0.0], np.float32), 12, 0, rng=(0, 0))
    np.testing.assert_array_equal(r[:, :-1], p[0, :, :-1].astype(np.float64))
    assert np.all(r[:, -1] == -1e30)


def test_reproj_constant_and_aligned_rays():
    # one tilt t re-projected at t: a constant c gives c cos t (the weight
    # cos(a - t) = 1 cancels); a quadratic profile comes back within one
    # pixel of itself, which a ray crossing the back-projection would not
    for ang in (25.0, -40.0):
        tl = np.array([ang], np.float32)
        c = np.full((1, 4, 64), 3.0, np.float32)
        r = oracle.reproj(c, tl, 16, 0, rng=(0, 0))
        ok = r > -1e29
        assert ok[:, 20:44].all()
        want = np.float32(3.0 * np.cos(np.float64(np.float32(ang * np.float32(0.01745)))))
        np.testing.assert_allclose(r[ok], want, rtol=1e-7)
        u = np.arange(64, dtype=np.float64)
        q = np.broadcast_to(((u - 32) ** 2).astype(np.float32), (1, 4, 64)).copy()
        rq = oracle.reproj(q, tl, 16, 0, rng=(0, 0))[0] / float(want / 3.0)
        for x in range(22, 42):
            lo = min((x - 33) ** 2, (x - 31) ** 2, (x - 32) ** 2)
            hi = max((x - 33) ** 2, (x - 31) ** 2)
            assert lo - 1e-3 <= rq[x] <= hi + 1e-3, (ang, x, rq[x])


def test_reproj_linear_and_sentinels():
    rng = np.random.default_rng(5)
    t = np.arange(-30, 31, 6, dtype=np.float32)
    P = rng.uniform(0.5, 1.5, (len(t), 3, 40)).astype(np.float32)
    Q = rng.uniform(0.5, 1.5, (len(t), 3, 40)).astype(np.float32)
    skip = (t >= 18).astype(np.uint8)
    k = int(np.where(t == 18)[0][0])
    a = oracle.reproj(P, t, 10, k, skip)
    b = oracle.reproj(Q, t, 10, k, skip)
    ab = oracle.reproj((2 * P + 3 * Q).astype(np.float32), t, 10, k, skip)
    ok = a > -1e29
    assert ok.sum() > 60 and np.array_equal(ok, ab > -1e29)
    np.testing.assert_allclose(ab[ok], 2 * a[ok] + 3 * b[ok], rtol=2e-6)
    # a band empty in every projection of the range is empty in the output
    Ps = P.copy()
    Ps[:, :, 10:14] = -1e30
    s = oracle.reproj(Ps, t, 10, k, skip)
    assert np.all(s[:, 11:13] == -1e30)
    assert np.all(s[:, 30:32] > -1e29)


# ---- sample thickness (CCalcVolThick) ---------------------------------------
def test_thick_profile_is_the_tile_std():
    rng = np.random.default_rng(8)
    nz, ny, nx = 7, 30, 41
    vol = rng.normal(0, 1, (nz, ny, nx)).astype(np.float32) * np.arange(1, nz + 1)[:, None, None]
    vol = vol.astype(np.float32)
    tx, ty = int(nx * np.float32(0.4)) // 2 * 2, int(ny * np.float32(0.4)) // 2 * 2
    x0, y0 = (nx - tx) // 2, (ny - ty) // 2
    vol[2, y0 + 1, x0 + 2] = -1e30                       # an empty voxel is left out
    vol[:, :y0, :] = 1e6                                 # outside the tile: ignored
    prof = oracle.thick_profile(vol)
    assert prof.shape == (nz - 1,)
    for z in range(nz - 1):
        tile = vol[z, y0:y0 + ty, x0:x0 + tx].astype(np.float64)
        tile = tile[tile > -1e20]
        np.testing.assert_allclose(prof[z], tile.std(), rtol=1e-6)
    assert np.all(oracle.thick_profile(np.full((4, 10, 10), 3.0, np.float32)) == 0)


def test_thick_smooth_closed_forms():
    imp = np.zeros(40, np.float32)
    imp[20] = 11.0
    s = oracle.thick_smooth(imp)
    np.testing.assert_array_equal(s[15:26], np.ones(11, np.float32))
    assert s[14] == 0 and s[26] == 0
    end = np.zeros(20, np.float32)
    end[0] = 1.0
    s = oracle.thick_smooth(end)
    assert s[0] == np.float32(6.0) / 11 and s[5] == np.float32(1.0) / 11 and s[6] == 0
    np.testing.assert_array_equal(oracle.thick_smooth(np.full(15, 2.5, np.float32)), 2.5)


def test_thick_edges_box_profile():
    # box of 1 on [30, 70) in 0: mean over [10, 90) = 0.5; the low edge is the
    # midpoint of the first point >= mean (30) and the first minimum (10), the
    # high edge of 69 and 90
    p = np.zeros(100, np.float32)
    p[30:70] = 1.0
    assert oracle.thick_edges(p, 1.0) == (20, 79, 59, 49 - 50)
    # binning scales the edges (int *= float) and the volume centre
    assert oracle.thick_edges(p, 20.0) == (400, 1580, 1180, 990 - 1000)
    # a sloped floor: the lowest point outside the crossing is used
    q = p.copy()
    q[10:30] = np.linspace(0.4, 0.1, 20, dtype=np.float32)
    e0 = oracle.thick_edges(q, 1.0)[0]
    assert e0 == (30 + 29) // 2
