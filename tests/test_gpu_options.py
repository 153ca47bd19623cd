"""Parity of the HIP path for the plan options beyond the reference default:
narrow projection storage, output layouts, the per-slice operators, excluded
tilts, odd widths, unfilled sentinels and a larger series.

Tolerances (relative L2 against the FP64 oracle): 1e-5 for FP32 storage,
2e-3 for FP16/BF16 storage.  Layout transforms are bit-exact.
"""
import numpy as np
import pytest

import oracle
from aretomo3_amd import DeviceArray, Plan, TomoSart, TomoWbp, phantom
from aretomo3_amd import _lib as L
from helpers import rel_l2, small_series

pytestmark = pytest.mark.gpu
TOL32 = 1e-5
TOL16 = 2e-3


def _ref(proj, tilts, method, vol_z, **kw):
    n, py, px = proj.shape
    return oracle.recon_slices(proj.astype(np.float64), tilts, method, vol_z, list(range(py)),
                               **kw)


@pytest.mark.parametrize("storage", [L.STORAGE_F16, L.STORAGE_BF16])
@pytest.mark.parametrize("method", [L.METHOD_SART, L.METHOD_WBP])
def test_narrow_storage(gpu_available, storage, method):
    tilts, proj, _ = small_series(px=64, py=6, vol_z=24)
    print(f"case storage={storage} method={method}", flush=True)
    plan = Plan(64, 6, tilts, 24, method, sart_iters=3, sart_subsets=3, storage=storage)
    gpu = plan.execute_host(proj)
    ref = _ref(proj, tilts, method, 24, iters=3, n_subsets=3)
    err = rel_l2(gpu, ref)
    print("rel_l2", err)
    assert err <= TOL16


@pytest.mark.parametrize("method", [L.METHOD_SART, L.METHOD_WBP])
def test_layouts_are_exact_permutations(gpu_available, method):
    tilts, proj, _ = small_series(px=48, py=5, vol_z=20)
    print(f"case layouts method={method}", flush=True)
    out = {}
    for lay in (L.LAYOUT_XZY, L.LAYOUT_XYZ, L.LAYOUT_XYZ_NOFLIP, L.LAYOUT_SLICE):
        plan = Plan(48, 5, tilts, 20, method, sart_iters=2, sart_subsets=2, layout=lay)
        out[lay] = plan.execute_host(proj)
        plan.close()
    xzy = out[L.LAYOUT_XZY].astype(np.float64)
    assert out[L.LAYOUT_XZY].shape == (5, 20, 48)
    np.testing.assert_array_equal(out[L.LAYOUT_XYZ], oracle.flip_vol(xzy, 1).astype(np.float32))
    np.testing.assert_array_equal(out[L.LAYOUT_XYZ_NOFLIP],
                                  oracle.flip_vol(xzy, 0).astype(np.float32))
    np.testing.assert_array_equal(out[L.LAYOUT_SLICE], out[L.LAYOUT_XZY][:, ::-1, :])


def test_exclude_mask(gpu_available):
    tilts, proj, _ = small_series(px=64, py=4, vol_z=24)
    n = len(tilts)
    mask = np.zeros(n, np.uint8)
    mask[[1, 5, n - 2]] = 1
    print("case exclude", flush=True)
    for method in (L.METHOD_SART, L.METHOD_WBP):
        plan = Plan(64, 4, tilts, 24, method, sart_iters=3, sart_subsets=4, exclude=mask)
        gpu = plan.execute_host(proj)
        ref = _ref(proj, tilts, method, 24, iters=3, n_subsets=4, no_proj=mask)
        err = rel_l2(gpu, ref)
        print("rel_l2", method, err)
        assert err <= TOL32


@pytest.mark.parametrize("px", [65, 81])
def test_odd_width(gpu_available, px):
    tilts, proj, _ = small_series(px=px, py=3, vol_z=18)
    print(f"case odd px={px}", flush=True)
    for method in (L.METHOD_SART, L.METHOD_WBP):
        plan = Plan(px, 3, tilts, 18, method, sart_iters=2, sart_subsets=3)
        gpu = plan.execute_host(proj)
        assert gpu.shape == (3, 18, px // 2 * 2)
        ref = _ref(proj, tilts, method, 18, iters=2, n_subsets=3)
        err = rel_l2(gpu, ref)
        print("rel_l2", method, err)
        assert err <= TOL32


def test_unfilled_sentinels(gpu_available):
    # SART after a pre-stage without the random fill: the aligned images keep
    # -1e30 where the rotation leaves the raw image, and the empty-sample rules
    # of weighting, back-projection, forward projection and difference apply
    # (WBP always runs on filled images: its ramp filter spreads any sentinel)
    tilts, raw, shifts = small_series(px=48, py=40, vol_z=16, axis=85.0, shift=3.0)
    axes = np.full(len(tilts), 85.0, np.float32)
    print("case sentinels", flush=True)
    aligned = oracle.prestage(raw, axes, shifts, 0, 0)
    assert (aligned < -1e10).any()
    for method in (L.METHOD_SART,):
        plan = Plan(48, 40, tilts, 16, method, tilt_axis=axes, shifts=shifts, random_fill=False,
                    sart_iters=2, sart_subsets=3, fill_seed=5)
        gpu = plan.execute_host(raw)
        ref = oracle.recon_slices(aligned, tilts, method, 16, list(range(aligned.shape[1])),
                                  iters=2, n_subsets=3, seed=5)
        assert np.isfinite(gpu).all()
        err = rel_l2(gpu, ref)
        print("rel_l2", method, err)
        assert err <= TOL32


def test_slice_operators(gpu_available):
    # CTomoSart / CTomoWbp on a padded sinogram in device memory; TomoSart
    # continues from the volume already in the buffer
    tilts, proj, _ = small_series(px=64, py=1, vol_z=24)
    n = len(tilts)
    pad = (64 // 2 + 1) * 2
    sino = np.zeros((n, pad), np.float32)
    sino[:, :64] = proj[:, 0, :]
    cs = oracle.cos_sin(tilts)
    print("case slice operators", flush=True)
    d_sino = DeviceArray.from_numpy(sino)
    start = np.random.default_rng(3).random((24, 64)).astype(np.float32)
    d_vol = DeviceArray.from_numpy(start)
    op = TomoSart()
    op.setup(64, 24, 3, 2, 64, tilts, 0, n)
    op.do_it(d_sino, d_vol)
    got = d_vol.to_numpy((24, 64))
    table = oracle.fill_table(0 ^ 0x5A5A5A5A5A5A5A5A, 64 * 24)
    ref, _ = oracle.tomo_sart(sino.astype(np.float64), 64, cs, start.astype(np.float64), 0, n,
                              2, 3, fill_tbl=table)
    err = rel_l2(got, ref)
    print("rel_l2 sart", err)
    assert err <= TOL32
    d_vol2 = DeviceArray(4 * 24 * 64)
    w = TomoWbp()
    w.setup(64, 24, 64, tilts)
    w.do_it(d_sino, d_vol2)
    got = d_vol2.to_numpy((24, 64))
    ref, _ = oracle.tomo_wbp(sino.astype(np.float64), 64, cs, 64, 24)
    err = rel_l2(got, ref)
    print("rel_l2 wbp", err)
    assert err <= TOL32


def test_repeat_execution(gpu_available):
    # two executions of one plan give identical results (no state leaks)
    tilts, proj, _ = small_series(px=64, py=4, vol_z=20)
    print("case repeat", flush=True)
    plan = Plan(64, 4, tilts, 20, L.METHOD_SART, sart_iters=2, sart_subsets=2)
    a = plan.execute_host(proj)
    b = plan.execute_host(proj)
    np.testing.assert_array_equal(a, b)


def test_larger_series_sampled_slices(gpu_available):
    # raw 64 x 512 with tilt axis 85 deg + shifts (aligned 512 wide, 64 rows),
    # 41 tilts, VolZ 128, benchmark SART options; sampled slices vs the oracle
    tilts = phantom.tilt_range(-60, 60, 3)
    raw, shifts = phantom.tilt_series(31, 64, 512, tilts, 128, n_ellipsoids=80, axis_min=4,
                                      axis_max=30, tilt_axis=85.0, max_shift=3.0,
                                      noise_frac=0.1)
    axes = np.full(len(tilts), 85.0, np.float32)
    print("case larger series", flush=True)
    plan = Plan(64, 512, tilts, 128, L.METHOD_SART, tilt_axis=axes, shifts=shifts,
                sart_iters=5, sart_subsets=5, relax=0.3, subset_order=L.SUBSETS_INTERLEAVED,
                positivity=False, fill_seed=31)
    gpu = plan.execute_host(raw)
    aligned = oracle.prestage(raw, axes, shifts, 1, 31)
    assert aligned.shape[1:] == (64, 512) and gpu.shape == (64, 128, 512)
    sl = [0, 13, 31, 40, 63]
    ref = oracle.recon_slices(aligned, tilts, 1, 128, sl, iters=5, n_subsets=5, relax=0.3,
                              order=1, positivity=0, seed=31)
    err = rel_l2(gpu[sl], ref)
    print("rel_l2", err, "shape", gpu.shape)
    assert err <= TOL32


def test_widest_supported_slices(gpu_available):
    # SART at ProjX 4096: 3-row forward-projection bands (4097 pairs per LDS
    # row), 64 ray tiles per tilt, one tilt per forward workgroup; WBP at its
    # widest VolX (2048: 4096-point row FFTs on the 2x grid)
    from aretomo3_amd import phantom
    for px, py, vz, method, step in ((4096, 2, 40, L.METHOD_SART, 6.0),
                                     (2048, 2, 48, L.METHOD_WBP, 6.0)):
        tilts = phantom.tilt_range(-60, 60, step)
        proj, _ = phantom.tilt_series(77, px, py, tilts, vz, n_ellipsoids=40, axis_min=8,
                                      axis_max=60, noise_frac=0.05)
        print(f"case widest px={px} method={method}", flush=True)
        plan = Plan(px, py, tilts, vz, method, sart_iters=2, sart_subsets=3)
        gpu = plan.execute_host(proj)
        plan.close()
        ref = oracle.recon_slices(proj.astype(np.float64), tilts, 1 if method == L.METHOD_SART else 0,
                                  vz, list(range(py)), iters=2, n_subsets=3)
        err = rel_l2(gpu, ref)
        print("rel_l2", err)
        assert err <= 1e-5


@pytest.mark.parametrize("px,storage", [(1024, L.STORAGE_F32), (2048, L.STORAGE_F16),
                                        (1000, L.STORAGE_F32)])
def test_forward_projector_widths(gpu_available, px, storage):
    # VolX 512 / 1024 / 2048 take forward-projection variants whose band row
    # pitch is a compile-time constant; other widths the general one
    tilts = phantom.tilt_range(-60, 60, 4)
    proj, _ = phantom.tilt_series(83, px, 2, tilts, 64, n_ellipsoids=60, axis_min=6,
                                  axis_max=50, noise_frac=0.05)
    print(f"case fp widths px={px} storage={storage}", flush=True)
    plan = Plan(px, 2, tilts, 64, L.METHOD_SART, sart_iters=2, sart_subsets=3, relax=0.3,
                subset_order=L.SUBSETS_INTERLEAVED, positivity=False, storage=storage)
    gpu = plan.execute_host(proj)
    plan.close()
    ref = _ref(proj, tilts, L.METHOD_SART, 64, iters=2, n_subsets=3, relax=0.3, order=1,
               positivity=0)
    err = rel_l2(gpu, ref)
    print("rel_l2", err)
    assert err <= (TOL32 if storage == L.STORAGE_F32 else TOL16)


@pytest.mark.parametrize("storage", [L.STORAGE_F16, L.STORAGE_BF16])
def test_unfilled_sentinels_narrow_storage(gpu_available, storage):
    # empty samples held in FP16/BF16 (which cannot hold -1e30) are widened
    # back to the sentinel when staged: same empty-sample rules as FP32
    tilts, raw, shifts = small_series(px=48, py=40, vol_z=16, axis=85.0, shift=3.0)
    axes = np.full(len(tilts), 85.0, np.float32)
    print(f"case sentinels storage={storage}", flush=True)
    aligned = oracle.prestage(raw, axes, shifts, 0, 0)
    plan = Plan(48, 40, tilts, 16, L.METHOD_SART, tilt_axis=axes, shifts=shifts, random_fill=False,
                sart_iters=2, sart_subsets=3, fill_seed=5, storage=storage)
    gpu = plan.execute_host(raw)
    ref = oracle.recon_slices(aligned, tilts, 1, 16, list(range(aligned.shape[1])),
                              iters=2, n_subsets=3, seed=5)
    # this case covers empty samples of the sinogram; the output is filled on
    # both sides, and empty voxels of the volume are covered by
    # test_empty_voxels_slice_operator below
    assert np.isfinite(gpu).all() and not (gpu < -1e10).any()
    err = rel_l2(gpu, ref)
    print("rel_l2", err)
    assert err <= TOL16


@pytest.mark.parametrize("method", [L.METHOD_SART, L.METHOD_WBP])
def test_strided_input_equals_dense(gpu_available, method):
    # raw frames inside a larger buffer (row pitch ProjX + 7, frame pitch
    # (ProjY + 3) rows), with the pre-stage: the same volume, bit for bit, as
    # the dense input
    tilts, raw, shifts = small_series(px=56, py=30, vol_z=18, axis=85.0, shift=2.0)
    n = len(tilts)
    axes = np.full(n, 85.0, np.float32)
    rs, fr = 56 + 7, 30 + 3
    big = np.full((n, fr, rs), 7.0, np.float32)
    big[:, :30, :56] = raw
    print(f"case strided input method={method}", flush=True)
    kw = dict(tilt_axis=axes, shifts=shifts, sart_iters=2, sart_subsets=2, fill_seed=9)
    dense = Plan(56, 30, tilts, 18, method, **kw)
    want = dense.execute_host(raw)
    dense.close()
    strided = Plan(56, 30, tilts, 18, method, tilt_stride=fr * rs, row_stride=rs, **kw)
    got = strided.execute_host(big)
    strided.close()
    np.testing.assert_array_equal(got, want)


def test_concurrent_plans_on_user_streams(gpu_available):
    # two plans of different series, each run twice on its own caller-created
    # stream with the launches interleaved: a plan owns all its scratch, so
    # each result equals the plan's serial result bit for bit
    import ctypes
    from aretomo3_amd import DeviceArray, synchronize
    series = [small_series(seed=61, px=64, py=10, vol_z=24),
              small_series(seed=62, px=48, py=14, vol_z=20,
                           tilts=np.arange(-54.0, 55.0, 6.0).astype(np.float32))]
    plans, inputs, outputs, wants, streams = [], [], [], [], []
    print("case concurrent plans on user streams", flush=True)
    for k, (tilts, proj, _) in enumerate(series):
        px, py, vz = proj.shape[2], proj.shape[1], (24, 20)[k]
        p = Plan(px, py, tilts, vz, L.METHOD_SART, sart_iters=3, sart_subsets=3)
        wants.append(p.execute_host(proj))
        plans.append(p)
        inputs.append(DeviceArray.from_numpy(proj))
        outputs.append([DeviceArray(4 * int(np.prod(p.volume_shape))) for _ in range(2)])
        s = ctypes.c_void_p()
        L.check(L.lib().atr_stream_create(ctypes.byref(s)))
        streams.append(s.value)
    for rep in range(2):
        for k in range(2):
            plans[k].execute(inputs[k], outputs[k][rep], False, streams[k])
    for s in streams:
        L.check(L.lib().atr_stream_synchronize(s))
    synchronize()
    for k in range(2):
        for rep in range(2):
            got = outputs[k][rep].to_numpy(plans[k].volume_shape)
            np.testing.assert_array_equal(got, wants[k])
    for s in streams:
        L.check(L.lib().atr_stream_destroy(s))
    for p in plans:
        p.close()


@pytest.mark.parametrize("storage", [L.STORAGE_F32, L.STORAGE_F16, L.STORAGE_BF16])
@pytest.mark.parametrize("positivity", [True, False])
def test_empty_voxels_slice_operator(gpu_available, storage, positivity):
    # CTomoSart::DoIt continuing from a volume that holds empty (-1e30) voxels
    # (atr_execute, accumulate, ATR_LAYOUT_SLICE): the load flags the slices
    # that hold them, the forward projector skips every sample that touches
    # one and returns -1e30 for a ray with no sample left (GForProj.cu:54,74),
    # the difference is 0 there, the back-projection keeps an empty voxel
    # empty (GBackProj.cu:55-59), and the final fill replaces each through the
    # seeded chain (CTomoSart.cpp:86-89).  Slices 1 and 4 hold no empty voxel,
    # so one launch runs both forward-projection variants.
    px, py, vz = 64, 6, 24
    tilts, proj, _ = small_series(seed=41, px=px, py=py, vol_z=vz)
    n = len(tilts)
    rng = np.random.default_rng(17)
    start = rng.random((py, vz, px)).astype(np.float32)
    empty = rng.random((py, vz, px)) < 0.03
    empty[0, 5:9, 20:26] = True                # a block: rays through it lose samples
    empty[2, :, 40] = True                     # a full column: every 0-degree ray there is empty
    empty[3, 11, :] = True                     # a full row
    empty[[1, 4]] = False
    start[empty] = -1e30
    print(f"case empty voxels storage={storage} positivity={positivity}", flush=True)
    plan = Plan(px, py, tilts, vz, L.METHOD_SART, sart_iters=3, sart_subsets=3,
                positivity=positivity, layout=L.LAYOUT_SLICE, storage=storage, fill_seed=23)
    d_in = DeviceArray.from_numpy(proj)
    d_vol = DeviceArray.from_numpy(start)
    plan.execute(d_in, d_vol, accumulate=True)
    got = d_vol.to_numpy((py, vz, px))
    plan.close()
    cs = oracle.cos_sin(tilts)
    pad = (px // 2 + 1) * 2
    table = oracle.fill_table(23 ^ 0x5A5A5A5A5A5A5A5A, px * vz)
    refs, plain = [], []
    for y in range(py):
        sino = np.zeros((n, pad), np.float64)
        sino[:, :px] = proj[:, y, :]
        r, _ = oracle.tomo_sart(sino, px, cs, start[y].astype(np.float64), 0, n, 3, 3,
                                positivity=1 if positivity else 0, fill_tbl=table)
        refs.append(r)
        # the same start with the empty voxels read as 0: what a path that
        # ignored the empty-voxel rules would compute
        s0 = np.where(empty[y], 0.0, start[y]).astype(np.float64)
        r0, _ = oracle.tomo_sart(sino, px, cs, s0, 0, n, 3, 3,
                                 positivity=1 if positivity else 0, fill_tbl=table)
        plain.append(r0)
    ref, ref0 = np.array(refs), np.array(plain)
    assert np.isfinite(got).all() and not (got < -1e10).any()
    tol = TOL32 if storage == L.STORAGE_F32 else TOL16
    err = rel_l2(got, ref)
    sep = rel_l2(ref0, ref)
    print("rel_l2", err, "ignoring the empty rules would give", sep)
    assert err <= tol
    # the comparison can tell the two apart
    assert sep > 20 * tol
