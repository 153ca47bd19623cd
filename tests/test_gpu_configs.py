"""Each benchmark configuration at its own workload, on the GPU, against the
FP64 oracle (and, for batching and serving, against each series alone):

* plain WBP: 512^2 x 41 tilts (-60..60 step 3), 200 ellipsoids, noise 5%
  -> 512 x 512 x 128, every slice against the oracle (rel-L2 <= 1e-5);
* precision: 2048^2 x 61 tilts (step 2), 2000 ellipsoids of semi-axes 4-80,
  noise 10% -> 2048 x 2048 x 512, WBP and 10 SART iterations x 5 interleaved
  subsets with the projections stored in FP16 and in BF16, sampled slices
  against the oracle (rel-L2 <= 2e-3, Fourier ring correlation >= 0.99 in
  every ring to Nyquist); BF16 SART over three seeds;
* batching: 16 series of seeded uneven sizes (tilts {31,41,51,61}, projections
  {512,768,1024}^2, thickness {128,256,384}, tilt axis 80-95 deg + shifts,
  200-800 ellipsoids, noise 10%), autotuned tile shapes, 15 SART iterations
  in one persistent launch per pass replayed from a HIP graph: every volume
  bit for bit equal to the series alone, two series' sampled slices (SART
  and WBP) against the oracle;
* serving: a mix with one 2048^2 x 61 x 512 request beside series of the
  batching sizes, WBP + 10 SART iterations each in FP16 storage, submitted
  together: each result equal to the series alone, sampled slices against
  the oracle.
"""
import numpy as np
import pytest

import oracle
from aretomo3_amd import Batch, DeviceArray, Plan, Server, phantom, synchronize
from aretomo3_amd import _lib as L
from helpers import download_frames, fourier_ring_correlation, rel_l2

pytestmark = pytest.mark.gpu

SART_BENCH = dict(relax=0.3, subset_order=L.SUBSETS_INTERLEAVED, positivity=False)
TOL32, TOL16 = 1e-5, 2e-3


def aligned_rows(raw, axes, shifts, seed, rows):
    """Rows `rows` of the oracle's pre-stage output (the whole image is
    aligned: the rotation couples rows)."""
    aligned = oracle.prestage(raw, axes, shifts, 1, seed)
    return np.ascontiguousarray(aligned[:, rows, :])


# ---------------------------------------------------------------------------
# plain WBP, one series
# ---------------------------------------------------------------------------
def test_plain_wbp_512_every_slice(gpu_available):
    seed, px, vz = 20261202, 512, 128
    tilts = phantom.tilt_range(-60.0, 60.0, 3.0)
    raw, shifts = phantom.tilt_series(seed, px, px, tilts, vz, n_ellipsoids=200, axis_min=4.0,
                                      axis_max=40.0, tilt_axis=0.0, noise_frac=0.05, n_threads=16)
    axes = np.zeros(len(tilts), np.float32)
    print("case plain WBP 512^2 x 41 -> 512 x 512 x 128", flush=True)
    plan = Plan(px, px, tilts, vz, L.METHOD_WBP, tilt_axis=axes, shifts=shifts, fill_seed=seed)
    d_in = DeviceArray.from_numpy(raw)
    d_out = DeviceArray(4 * px * px * vz)
    plan.execute(d_in, d_out)
    synchronize()
    got = d_out.to_numpy(plan.volume_shape)
    plan.close()
    aligned = oracle.prestage(raw, axes, shifts, 1, seed)
    ref = oracle.recon_slices(aligned, tilts, 0, vz, list(range(px)))
    err = rel_l2(got, ref)
    print("rel_l2 over all 512 slices", err, flush=True)
    assert got.shape == (px, vz, px)
    assert err <= TOL32


# ---------------------------------------------------------------------------
# precision: 2048^2 x 61 -> 2048 x 2048 x 512, FP16 / BF16 storage
# ---------------------------------------------------------------------------
C3_PX, C3_VZ, C3_ITERS, C3_SUBSETS = 2048, 512, 10, 5
C3_SLICES = [0, 683, 1366, 2047]
C3_SEEDS = [20261204, 20261214, 20261224]
_c3_cache = {}


def c3_series(seed):
    if seed not in _c3_cache:
        _c3_cache.clear()   # one 2048^2 series (1 GiB) held at a time
        tilts = phantom.tilt_range(-60.0, 60.0, 2.0)
        raw, shifts = phantom.tilt_series(seed, C3_PX, C3_PX, tilts, C3_VZ, n_ellipsoids=2000,
                                          axis_min=4.0, axis_max=80.0, tilt_axis=0.0,
                                          noise_frac=0.10, n_threads=16)
        axes = np.zeros(len(tilts), np.float32)
        rows = aligned_rows(raw, axes, shifts, seed, C3_SLICES)
        _c3_cache[seed] = dict(tilts=tilts, raw=raw, shifts=shifts, axes=axes, rows=rows, ref={})
    return _c3_cache[seed]


def c3_reference(c, method, seed):
    if method not in c["ref"]:
        k = list(range(len(C3_SLICES)))
        if method == L.METHOD_SART:
            c["ref"][method] = oracle.recon_slices(c["rows"], c["tilts"], 1, C3_VZ, k,
                                                   iters=C3_ITERS, n_subsets=C3_SUBSETS, relax=0.3,
                                                   order=1, positivity=0, seed=seed)
        else:
            c["ref"][method] = oracle.recon_slices(c["rows"], c["tilts"], 0, C3_VZ, k)
    return c["ref"][method]


def c3_run(seed, method, storage):
    c = c3_series(seed)
    kw = dict(tilt_axis=c["axes"], shifts=c["shifts"], storage=storage, fill_seed=seed)
    if method == L.METHOD_SART:
        kw.update(SART_BENCH, sart_iters=C3_ITERS, sart_subsets=C3_SUBSETS)
    plan = Plan(C3_PX, C3_PX, c["tilts"], C3_VZ, method, **kw)
    d_in = DeviceArray.from_numpy(c["raw"])
    d_out = DeviceArray(4 * C3_PX * C3_PX * C3_VZ)
    plan.execute(d_in, d_out)
    synchronize()
    got = download_frames(d_out, (C3_VZ, C3_PX), C3_SLICES)
    plan.close()
    d_out.free()
    d_in.free()
    ref = c3_reference(c, method, seed)
    err = rel_l2(got, ref)
    frc = fourier_ring_correlation(got, ref)
    return err, frc


@pytest.mark.parametrize("storage,seed", [(L.STORAGE_F16, C3_SEEDS[0])] +
                         [(L.STORAGE_BF16, s) for s in C3_SEEDS])
def test_precision_sart_2048(gpu_available, storage, seed):
    name = "FP16" if storage == L.STORAGE_F16 else "BF16"
    print(f"case precision SART 2048^2 x 61 -> 2048 x 2048 x 512, {name}, seed {seed}", flush=True)
    err, frc = c3_run(seed, L.METHOD_SART, storage)
    print(f"{name} seed {seed}: rel_l2 {err:.3e}, FRC min {frc.min():.6f}, "
          f"FRC at Nyquist {frc[-1]:.6f}", flush=True)
    assert err <= TOL16
    assert frc.min() >= 0.99


@pytest.mark.parametrize("storage", [L.STORAGE_F16, L.STORAGE_BF16])
def test_precision_wbp_2048(gpu_available, storage):
    name = "FP16" if storage == L.STORAGE_F16 else "BF16"
    print(f"case precision WBP 2048^2 x 61 -> 2048 x 2048 x 512, {name}", flush=True)
    err, frc = c3_run(C3_SEEDS[0], L.METHOD_WBP, storage)
    print(f"{name}: rel_l2 {err:.3e}, FRC min {frc.min():.6f}", flush=True)
    assert err <= TOL16
    assert frc.min() >= 0.99


# ---------------------------------------------------------------------------
# batching: 16 uneven series
# ---------------------------------------------------------------------------
def batching_series(seed0=20261205, count=16):
    """The 16 seeded series of the batching configuration (sizes drawn as the
    benchmark draws them)."""
    rng = np.random.default_rng(seed0)
    out = []
    for k in range(count):
        n = int(rng.choice([31, 41, 51, 61]))
        px = int(rng.choice([512, 768, 1024]))
        vz = int(rng.choice([128, 256, 384]))
        axis = float(rng.uniform(80.0, 95.0))
        n_ell = int(rng.integers(200, 801))
        out.append(dict(seed=seed0 + k, n=n, px=px, vz=vz, axis=axis, n_ell=n_ell))
    return out


def make_series(spec):
    tilts = np.linspace(-60.0, 60.0, spec["n"]).astype(np.float32)
    raw, shifts = phantom.tilt_series(spec["seed"], spec["px"], spec["px"], tilts, spec["vz"],
                                      n_ellipsoids=spec["n_ell"], axis_min=4.0, axis_max=40.0,
                                      tilt_axis=spec["axis"], max_shift=3.0, noise_frac=0.10,
                                      n_threads=16)
    axes = np.full(len(tilts), spec["axis"], np.float32)
    return tilts, raw, shifts, axes


def test_batching_16_series(gpu_available):
    specs = batching_series()
    print("case batching 16 series", [(s["n"], s["px"], s["vz"]) for s in specs], flush=True)
    plans, wbps, d_in, d_alone, d_batch, data = [], [], [], [], [], []
    shapes = set()
    for s in specs:
        tilts, raw, shifts, axes = make_series(s)
        geo = dict(tilt_axis=axes, shifts=shifts, fill_seed=s["seed"])
        p = Plan(s["px"], s["px"], tilts, s["vz"], L.METHOD_SART, sart_iters=15, sart_subsets=5,
                 **SART_BENCH, **geo)
        shapes.add(p.autotune())
        plans.append(p)
        wbps.append(Plan(s["px"], s["px"], tilts, s["vz"], L.METHOD_WBP, **geo))
        d_in.append(DeviceArray.from_numpy(raw))
        nb = 4 * p.vol_x * p.vol_y * p.vol_z
        d_alone.append(DeviceArray(nb))
        d_batch.append(DeviceArray(nb))
        data.append((tilts, raw, shifts, axes))
    print("autotuned shapes", sorted(shapes), flush=True)
    for p, i, o in zip(plans, d_in, d_alone):
        p.execute(i, o)
    synchronize()
    batch = Batch(plans, graph=True)
    batch.execute(d_in, d_batch)
    synchronize()
    for k, p in enumerate(plans):
        a = d_alone[k].to_numpy(p.volume_shape)
        b = d_batch[k].to_numpy(p.volume_shape)
        same = np.array_equal(a, b)
        print("series", k, p.volume_shape, "batch == alone:", same, flush=True)
        assert same
        del a, b
    batch.close()
    # two series against the oracle: sampled slices of SART and WBP
    for k in (0, int(np.argmax([s["px"] * s["vz"] * s["n"] for s in specs]))):
        s, (tilts, raw, shifts, axes) = specs[k], data[k]
        p = plans[k]
        rows = [0, p.vol_y // 2, p.vol_y - 1]
        got = download_frames(d_batch[k], (p.vol_z, p.vol_x), rows)
        al = aligned_rows(raw, axes, shifts, s["seed"], rows)
        ref = oracle.recon_slices(al, tilts, 1, s["vz"], [0, 1, 2], iters=15, n_subsets=5,
                                  relax=0.3, order=1, positivity=0, seed=s["seed"])
        err = rel_l2(got, ref)
        d_w = DeviceArray(4 * p.vol_x * p.vol_y * p.vol_z)
        wbps[k].execute(d_in[k], d_w)
        synchronize()
        gw = download_frames(d_w, (p.vol_z, p.vol_x), rows)
        rw = oracle.recon_slices(al, tilts, 0, s["vz"], [0, 1, 2])
        werr = rel_l2(gw, rw)
        print("series", k, "SART rel_l2", err, "WBP rel_l2", werr, flush=True)
        assert err <= TOL32 and werr <= TOL32
        d_w.free()
    for p in plans + wbps:
        p.close()


# ---------------------------------------------------------------------------
# serving: a mix with one large request
# ---------------------------------------------------------------------------
def test_serving_mix_with_a_large_request(gpu_available):
    rng = np.random.default_rng(20261206)
    specs = [dict(seed=20261206, n=61, px=2048, vz=512, axis=float(rng.uniform(80, 95)), n_ell=600)]
    specs += batching_series(20261216, 3)
    reqs = []
    for s in specs:
        tilts, raw, shifts, axes = make_series(s)
        kw = dict(tilt_axis=axes, shifts=shifts, sart_iters=10, sart_subsets=5, fill_seed=s["seed"],
                  **SART_BENCH)
        reqs.append((s, tilts, raw, shifts, axes, kw))
    print("case serving mix", [(s["n"], s["px"], s["vz"]) for s in specs], flush=True)
    srv = Server(storage=L.STORAGE_F16)
    ids = [srv.submit(raw, tilts, s["vz"], with_wbp=True, **kw)
           for s, tilts, raw, shifts, axes, kw in reqs]
    res = [srv.wait(i) for i in ids]
    stats = srv.stats()
    srv.close()
    print("stats", stats, "latency ms", [round(r[2]["latency_ms"], 1) for r in res], flush=True)
    assert stats["max_batch"] >= 2
    for k, ((s, tilts, raw, shifts, axes, kw), (vol, wvol, info)) in enumerate(zip(reqs, res)):
        px = s["px"]
        frames = list(range(0, px, 64)) + [px - 1] if px > 1024 else list(range(px))
        for method, got_all in ((L.METHOD_SART, vol), (L.METHOD_WBP, wvol)):
            kwm = kw if method == L.METHOD_SART else dict(tilt_axis=axes, shifts=shifts,
                                                          fill_seed=s["seed"])
            plan = Plan(px, px, tilts, s["vz"], method, storage=L.STORAGE_F16, **kwm)
            d_in = DeviceArray.from_numpy(raw)
            d_out = DeviceArray(4 * plan.vol_x * plan.vol_y * plan.vol_z)
            plan.execute(d_in, d_out)
            synchronize()
            alone = download_frames(d_out, (s["vz"], plan.vol_x), frames)
            same = np.array_equal(got_all[frames], alone)
            print("request", k, "method", method, "mix == alone on", len(frames), "frames:", same,
                  flush=True)
            assert same
            plan.close()
            d_in.free()
            d_out.free()
        rows = [px // 3, (2 * px) // 3]
        al = aligned_rows(raw, axes, shifts, s["seed"], rows)
        ref = oracle.recon_slices(al, tilts, 1, s["vz"], [0, 1], iters=10, n_subsets=5, relax=0.3,
                                  order=1, positivity=0, seed=s["seed"])
        err = rel_l2(vol[rows], ref)
        print("request", k, "FP16 SART vs oracle rel_l2", err, flush=True)
        assert err <= TOL16
