#!/usr/bin/env python3
"""Benchmark of the reconstruction hot path on one MI355X.

Default workload (the headline SART series, --config sart1024): seeded phantom
of 200 random ellipsoids (semi-axes 4-40 voxels, density 0.5-1.0), 61 tilts
-60..+60 deg step 2, 1024x1024 raw projections with tilt axis 85 deg and
per-tilt shifts uniform +-3 px (supplied as alignment, so the pre-stage runs),
Gaussian noise of 10% of the mean; volume 1024 x 1024 x 256; 20 SART
iterations, 5 interleaved subsets, relaxation 0.3, positivity off; FP32
projection storage.

A step is one whole reconstruction of the series with the raw projections
already resident in HBM: pre-stage correction + fill, weighting, initial
back-projection and 20 x 5 forward/difference/back-projection passes, and the
volume written out.  The metric counts output-resolution voxel-tilt updates:
VolX*VolY*VolZ*N per pass; SART has 1 + 2*iterations passes, WBP one.

Other configurations (--config): wbp512 (WBP, 512^2 x 41 tilts step 3, volume
512x512x128, tilt axis 0, noise 5%), sart2048-f16 / sart2048-bf16 and
wbp2048-f16 / wbp2048-bf16 (2048^2 x 61 tilts, volume 2048x2048x512, 2000
ellipsoids of semi-axes 4-80, noise 10%, 10 SART iterations, projections
stored in FP16 / BF16 with FP32 accumulation), serve (the serving loop: a
Poisson stream of mixed-size WBP + SART requests, reporting p50/p99 latency and
tomograms/s beside the plain per-series FP32 path), batch16 (16 uneven series:
WBP each, then SART of all 16 in persistent batched launches replayed from a
HIP graph, beside the same work series by series), extras (the pre-stage with
local shifts and binning, the re-projector, the thickness measurement and MRC
I/O, each beside the FP64 oracle), small (a reduced SART case for quick checks;
not a benchmark).

The default run also carries a "serving" block: a short stream of the
serving configuration's request mix (40 requests, WBP + SART 10 x 5 each,
FP16 storage) at 4/s for p50/p99 latency and as a 20/s burst for throughput,
beside the plain per-series FP32 path on the same requests.

Prints one JSON line.  Options: --steps K --warmup W --gpus N (only 1),
--config NAME, --no-cpu-baseline, --cpu-slices S, --copy-probe MIB,
--requests R, --rate L and --max-active M (serve), --no-serve.
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gvoxel-tilt updates/s at ≤2e-3 rel-L2 vs FP64; p50/p99 tomogram latency"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = "profiles/hbm_traffic_sart_1024x61x256.md"

SART_BENCH = dict(method="sart", relax=0.3, order="interleaved", positivity=False)
CONFIGS = {
    "sart1024": dict(SART_BENCH, px=1024, vol_z=256, step=2.0, n_ell=200, amax=40.0, axis=85.0,
                     shift=3.0, noise=0.10, iters=20, subsets=5, storage="f32", seed=20261203,
                     cpu_slices=128),
    "wbp512": dict(method="wbp", px=512, vol_z=128, step=3.0, n_ell=200, amax=40.0, axis=0.0,
                   shift=0.0, noise=0.05, storage="f32", seed=20261202, cpu_slices=512),
    "sart2048-f16": dict(SART_BENCH, px=2048, vol_z=512, step=2.0, n_ell=2000, amax=80.0,
                         axis=0.0, shift=0.0, noise=0.10, iters=10, subsets=5, storage="f16",
                         seed=20261204, cpu_slices=16),
    "sart2048-bf16": dict(SART_BENCH, px=2048, vol_z=512, step=2.0, n_ell=2000, amax=80.0,
                          axis=0.0, shift=0.0, noise=0.10, iters=10, subsets=5, storage="bf16",
                          seed=20261204, cpu_slices=16),
    "wbp2048-f16": dict(method="wbp", px=2048, vol_z=512, step=2.0, n_ell=2000, amax=80.0,
                        axis=0.0, shift=0.0, noise=0.10, storage="f16", seed=20261204,
                        cpu_slices=16),
    "wbp2048-bf16": dict(method="wbp", px=2048, vol_z=512, step=2.0, n_ell=2000, amax=80.0,
                         axis=0.0, shift=0.0, noise=0.10, storage="bf16", seed=20261204,
                         cpu_slices=16),
    "small": dict(SART_BENCH, px=256, vol_z=64, step=2.0, n_ell=60, amax=20.0, axis=85.0,
                  shift=3.0, noise=0.10, iters=4, subsets=5, storage="f32", seed=20261204,
                  cpu_slices=8),
}


def measured_traffic():
    """HBM bytes per k_back_proj launch of the headline SART series from the
    committed rocprofv3 summary (128 B per TCC_EA0_RDREQ read request plus
    WRITE_SIZE), or None."""
    try:
        text = open(os.path.join(ROOT, TRAFFIC_FILE)).read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch k_back_proj: ([0-9.eE+]+)", text)
    return float(m.group(1)) if m else None


def copy_probe(mib):
    """Plain copy of `mib` MiB, 3 launches: FETCH_SIZE/WRITE_SIZE per launch
    of this kernel calibrate the counters against known bytes."""
    from aretomo3_amd import DeviceArray, synchronize
    from aretomo3_amd._lib import check, lib
    nbytes = mib << 20
    a, b = DeviceArray(nbytes), DeviceArray(nbytes)
    check(lib().atr_memset(a.ptr, 0, nbytes))
    for _ in range(3):
        check(lib().atr_copy_kernel(b.ptr, a.ptr, nbytes, None))
    synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes, "launches": 3}))


def measured_copy_peak(mib=1024, reps=10):
    """Stream-copy bandwidth of this GPU in the same run: the plain copy
    kernel (16-byte loads and stores) moving `mib` MiB in and `mib` MiB out,
    timed with HIP events over `reps` launches.  The roofline's peak stays
    the published 8 TB/s; this is the attainable figure beside it."""
    from aretomo3_amd import DeviceArray
    from aretomo3_amd._lib import check, lib
    import ctypes
    nbytes = mib << 20
    a, b = DeviceArray(nbytes), DeviceArray(nbytes)
    check(lib().atr_memset(a.ptr, 0, nbytes))
    ev0, ev1 = ctypes.c_void_p(), ctypes.c_void_p()
    check(lib().atr_event_create(ctypes.byref(ev0)))
    check(lib().atr_event_create(ctypes.byref(ev1)))
    check(lib().atr_copy_kernel(b.ptr, a.ptr, nbytes, None))   # warm-up
    check(lib().atr_event_record(ev0, None))
    for _ in range(reps):
        check(lib().atr_copy_kernel(b.ptr, a.ptr, nbytes, None))
    check(lib().atr_event_record(ev1, None))
    ms = ctypes.c_float()
    check(lib().atr_event_elapsed_ms(ev0, ev1, ctypes.byref(ms)))
    check(lib().atr_event_destroy(ev0))
    check(lib().atr_event_destroy(ev1))
    a.free()
    b.free()
    gbs = 2.0 * nbytes * reps / (ms.value * 1e-3) / 1e9
    return {"kernel": "k_copy (16-byte loads/stores, grid-stride)",
            "bytes_per_launch": 2 * nbytes, "gbs": round(gbs, 1),
            "frac_of_8000": round(gbs / HBM_PEAK_GBS, 4)}


def fourier_ring_correlation(a, b, n_rings=32):
    """Mean over slices of the Fourier ring correlation of two stacks of XZ
    slices [k][z][x], per ring of normalised radius; returns the curve."""
    fa = np.fft.rfft2(a.astype(np.float64))
    fb = np.fft.rfft2(b.astype(np.float64))
    nz, nx = a.shape[1], a.shape[2]
    kz = np.fft.fftfreq(nz)[:, None]
    kx = np.fft.rfftfreq(nx)[None, :]
    r = np.sqrt(kz ** 2 + kx ** 2) / 0.5
    ring = np.minimum((r * n_rings).astype(int), n_rings)
    num = np.zeros(n_rings + 1)
    da = np.zeros(n_rings + 1)
    db = np.zeros(n_rings + 1)
    for k in range(a.shape[0]):
        num += np.bincount(ring.ravel(), (fa[k] * np.conj(fb[k])).real.ravel(), n_rings + 1)
        da += np.bincount(ring.ravel(), (np.abs(fa[k]) ** 2).ravel(), n_rings + 1)
        db += np.bincount(ring.ravel(), (np.abs(fb[k]) ** 2).ravel(), n_rings + 1)
    return (num / np.sqrt(np.maximum(da * db, 1e-300)))[:n_rings]


def describe(name, w, vol):
    vx, vy, vz = vol
    n = int(round(120.0 / w["step"])) + 1
    geo = ("tilt axis %g deg + shifts +-%g px" % (w["axis"], w["shift"]) if w["shift"] > 0
           else "tilt axis %g deg" % w["axis"])
    if w["method"] == "sart":
        alg = ("SART %d iterations x %d %s subsets, relax %g, positivity %s" %
               (w["iters"], w["subsets"], w["order"], w["relax"], "on" if w["positivity"] else "off"))
    else:
        alg = "WBP (ramp + Hamming, 2x-grid back-projection, Fourier crop)"
    return ("%s: %dx%d raw projections, %d tilts -60..60 step %g, %s, %d ellipsoids, "
            "noise %g%%, volume %dx%dx%d, %s, %s projection storage" %
            (name, w["px"], w["px"], n, w["step"], geo, w["n_ell"], 100 * w["noise"], vx, vy, vz,
             alg, w["storage"].upper()))


SCHEDULES = {"fair": 0, "every": 1, "shortest": 2}   # ATR_SCHEDULE_*
SERVE = dict(requests=500, rates="1,2,5,10,20", seed=20261206, large_frac=0.1, iters=10, subsets=5,
             max_active=8)
# the short serving stream of the default run: a C5-shaped stream sized to
# keep the whole default command within a few minutes
SERVE_DEFAULT = dict(requests=40, rate=4.0, burst_rate=20.0)


def serve_sizes(rng, count, large_frac):
    """Per-request sizes of the serving configuration: tilts {31,41,51,61}
    over +-60 deg, projections {512,768,1024}^2, thickness {128,256,384}, tilt
    axis uniform 80-95 deg (as the batching configuration draws them), and
    ~large_frac of the requests 2048^2 x 61 x 512."""
    out = []
    for _ in range(count):
        if rng.random() < large_frac:
            out.append((61, 2048, 512))
        else:
            out.append((int(rng.choice([31, 41, 51, 61])), int(rng.choice([512, 768, 1024])),
                        int(rng.choice([128, 256, 384]))))
    return out


class ServePool:
    """One seeded series per request size (the GPU work of a request depends
    on its sizes, not on the phantom), held in pinned host memory, with the
    plain per-series FP32 path timed on it: upload, WBP, SART, both volumes
    copied back to pinned host memory, series after series."""

    def __init__(self, seed, iters, subsets):
        self.seed, self.iters, self.subsets = seed, iters, subsets
        self.series, self.plain_s = {}, {}
        self.out = None

    def get(self, size):
        from aretomo3_amd import PinnedArray, phantom
        if size in self.series:
            return self.series[size]
        n, px, vz = size
        k = len(self.series)
        rng = np.random.default_rng(self.seed + 1000 + k)
        axis = float(rng.uniform(80.0, 95.0))
        tilts = np.linspace(-60.0, 60.0, n).astype(np.float32)
        t0 = time.time()
        raw, shifts = phantom.tilt_series(self.seed + k, px, px, tilts, vz,
                                          n_ellipsoids=int(rng.integers(200, 801)), axis_min=4.0,
                                          axis_max=40.0, tilt_axis=axis, max_shift=3.0,
                                          noise_frac=0.10, n_threads=16)
        pin = PinnedArray(raw.shape)
        pin.array[...] = raw
        kw = dict(tilt_axis=np.full(n, axis, np.float32), shifts=shifts, sart_iters=self.iters,
                  sart_subsets=self.subsets, relax=0.3, positivity=False, fill_seed=self.seed + k)
        from aretomo3_amd import _lib as L
        kw["subset_order"] = L.SUBSETS_INTERLEAVED
        vx = px // 2 * 2
        e = dict(pin=pin, tilts=tilts, vz=vz, kw=kw, shape=(px, vz, vx),
                 updates=float(vx) * px * vz * n * (1 + 1 + 2 * self.iters))
        self.series[size] = e
        print(f"[bench] serve pool {size} generated in {time.time() - t0:.1f}s", file=sys.stderr,
              flush=True)
        return e

    def outputs(self, shape):
        """Pinned SART and WBP output views of `shape` (one pair, sized for the
        largest volume; copy-out is serialised, so requests can share it)."""
        from aretomo3_amd import PinnedArray
        need = int(np.prod(shape))
        if self.out is None or self.out[0].array.size < need:
            cap = max(need, 2048 * 512 * 2048)
            self.out = (PinnedArray((cap,)), PinnedArray((cap,)))
        return (self.out[0].array[:need].reshape(shape), self.out[1].array[:need].reshape(shape))

    def plain(self, size):
        """Seconds of the plain per-series FP32 path for one request of this
        size (second of two runs): the two plans built (host tables and
        device memory, as the serving path builds them per request), upload,
        WBP, SART, both volumes copied back."""
        if size in self.plain_s:
            return self.plain_s[size]
        from aretomo3_amd import DeviceArray, Plan, synchronize
        from aretomo3_amd import _lib as L
        e = self.get(size)
        n, px, vz = size
        d_in = DeviceArray(e["pin"].array.nbytes)
        vx = px // 2 * 2
        nb = 4 * vx * px * vz
        d_out, d_w = DeviceArray(nb), DeviceArray(nb)
        times = []
        for _ in range(2):
            synchronize()
            t0 = time.perf_counter()
            ps = Plan(px, px, e["tilts"], vz, L.METHOD_SART, storage=L.STORAGE_F32, **e["kw"])
            pw = Plan(px, px, e["tilts"], vz, L.METHOD_WBP, storage=L.STORAGE_F32,
                      tilt_axis=e["kw"]["tilt_axis"], shifts=e["kw"]["shifts"],
                      fill_seed=e["kw"]["fill_seed"])
            h_out, h_w = self.outputs(ps.volume_shape)
            L.check(L.lib().atr_memcpy_h2d(d_in.ptr, e["pin"].array.ctypes.data, e["pin"].array.nbytes))
            pw.execute(d_in, d_w)
            ps.execute(d_in, d_out)
            synchronize()
            L.check(L.lib().atr_memcpy_d2h(h_out.ctypes.data, d_out.ptr, nb))
            L.check(L.lib().atr_memcpy_d2h(h_w.ctypes.data, d_w.ptr, nb))
            ps.close(), pw.close()
            times.append(time.perf_counter() - t0)
        d_in.free(), d_out.free(), d_w.free()
        self.plain_s[size] = times[-1]
        return times[-1]


def serve_stream(pool, sizes, rate, seed, max_active, schedule=SCHEDULES["shortest"]):
    """One open-loop stream: request k arrives after Exp(1/rate) gaps and is
    submitted to a fresh server (FP16 storage).  Returns the per-request
    records and the makespan (first arrival to last completion)."""
    from aretomo3_amd import Server
    from aretomo3_amd import _lib as L
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, len(sizes))
    srv = Server(storage=L.STORAGE_F16, max_active=max_active, schedule=schedule)
    # warm-up: one small request through this server
    e = pool.get((31, 512, 128))
    out, wout = pool.outputs(e["shape"])
    srv.wait(srv.submit(e["pin"].array, e["tilts"], e["vz"], with_wbp=True, out=out, wbp_out=wout,
                        **e["kw"]))
    ids = []
    t_start = time.perf_counter()
    t_next = t_start
    for k, (size, g) in enumerate(zip(sizes, gaps)):
        t_next += g
        now = time.perf_counter()
        if t_next > now:
            time.sleep(t_next - now)
        e = pool.get(size)
        out, wout = pool.outputs(e["shape"])
        ids.append(srv.submit(e["pin"].array, e["tilts"], e["vz"], with_wbp=True, out=out,
                              wbp_out=wout, **e["kw"]))
        if (k + 1) % 25 == 0:
            print(f"[bench] serve rate {rate}/s: {k + 1} of {len(sizes)} submitted after "
                  f"{time.perf_counter() - t_start:.1f}s", file=sys.stderr, flush=True)
    recs = []
    for size, rid in zip(sizes, ids):
        _, _, info = srv.wait(rid)
        recs.append(dict(size=size, **info))
    t_end = time.perf_counter()
    stats = srv.stats()
    srv.close()
    first = min(r["submit_ms"] for r in recs)
    makespan = max(r["done_ms"] for r in recs) - first
    return recs, makespan / 1e3, stats, t_end - t_start


def plain_fifo(pool, sizes, recs):
    """The plain per-series FP32 path on the same arrivals: one series at a
    time in arrival order, each taking its measured time (upload, WBP, SART,
    copy-back).  Returns (tomograms/s from first arrival to last completion,
    per-request latencies in ms)."""
    arrive = [r["submit_ms"] for r in recs]
    free_at, lat = arrive[0], []
    for a, s in zip(arrive, sizes):
        start = max(a, free_at)
        free_at = start + 1e3 * pool.plain(s)
        lat.append(free_at - a)
    return len(sizes) / ((free_at - arrive[0]) / 1e3), np.array(lat)


def serve_summary(pool, sizes, recs, makespan, stats, rate):
    """tomograms/s, p50/p99 latency, its breakdown per request class and the
    plain per-series FP32 path on the same requests: at capacity (series back
    to back) and on the same arrival times (FIFO, one series at a time)."""
    lat = np.array([r["latency_ms"] for r in recs])
    plain_total = sum(pool.plain(s) for s in sizes)
    tps = len(sizes) / makespan
    plain_tps = len(sizes) / plain_total
    fifo_tps, fifo_lat = plain_fifo(pool, sizes, recs)
    updates = sum(pool.get(s)["updates"] for s in sizes)

    def part(sel):
        if not sel:
            return None
        pick = lambda k: np.array([r[k] for r in sel])
        return {"requests": len(sel),
                "latency_ms_p50": round(float(np.percentile(pick("latency_ms"), 50)), 1),
                "latency_ms_p99": round(float(np.percentile(pick("latency_ms"), 99)), 1),
                "queue_ms_mean": round(float(pick("queue_ms").mean()), 1),
                "gpu_ms_mean": round(float(pick("gpu_ms").mean()), 1),
                "copy_out_ms_mean": round(float(pick("copy_out_ms").mean()), 1)}
    return {
        "requests": len(sizes),
        "arrival_rate_per_s": rate,
        "tomograms_per_s": round(tps, 3),
        "gvoxel_tilt_updates_per_s": round(updates / makespan / 1e9, 1),
        "latency_ms_p50": round(float(np.percentile(lat, 50)), 1),
        "latency_ms_p99": round(float(np.percentile(lat, 99)), 1),
        "makespan_s": round(makespan, 3),
        "rounds": stats["rounds"],
        "max_series_per_round": stats["max_batch"],
        "large_requests": int(sum(1 for s in sizes if s[1] == 2048)),
        "by_class": {"large_2048": part([r for r in recs if r["size"][1] == 2048]),
                     "batching_sizes": part([r for r in recs if r["size"][1] != 2048])},
        "plain_fp32_per_series_tomograms_per_s": round(plain_tps, 3),
        "speedup_vs_plain": round(tps / plain_tps, 3),
        "plain_fp32_same_arrivals": {
            "tomograms_per_s": round(fifo_tps, 3),
            "latency_ms_p50": round(float(np.percentile(fifo_lat, 50)), 1),
            "latency_ms_p99": round(float(np.percentile(fifo_lat, 99)), 1)},
        "speedup_vs_plain_same_arrivals": round(tps / fifo_tps, 3),
    }


SERVE_WORKLOAD = ("serve: open-loop Poisson stream, WBP + SART %d iterations x %d interleaved "
                  "subsets (relax 0.3, positivity off) per request, tilts {31,41,51,61} over +-60 deg, "
                  "projections {512,768,1024}^2, thickness {128,256,384}, tilt axis 80-95 deg + "
                  "shifts +-3 px, ~10%% large (2048^2 x 61 x 512), FP16 projection storage, FP32 "
                  "accumulation, admission by the device-memory budget (288 GB card), pinned host "
                  "buffers; plain = the same requests one after another through the per-series "
                  "FP32 path (plans, upload, WBP, SART, copy-back)")


def serve_bench(args):
    """The serving configuration at its stated size: 500 requests per arrival
    rate, rates swept over 1-20/s (--rate picks one), the same request sizes at
    every rate.  Prints one JSON line with a serving block per rate."""
    c = dict(SERVE)
    n_req = args.requests or c["requests"]
    rates = ([args.rate] if args.rate else
             [float(r) for r in (args.rates or c["rates"]).split(",")])
    pool = ServePool(c["seed"], c["iters"], c["subsets"])
    sizes = serve_sizes(np.random.default_rng(c["seed"]), n_req, c["large_frac"])
    for sz in sorted(set(sizes)):
        pool.get(sz)
        pool.plain(sz)
    per_rate = []
    for rate in rates:
        recs, makespan, stats, wall = serve_stream(pool, sizes, rate, c["seed"] + int(rate * 100),
                                                   args.max_active or c["max_active"],
                                                   SCHEDULES[args.schedule])
        summ = serve_summary(pool, sizes, recs, makespan, stats, rate)
        per_rate.append(summ)
        print(f"[bench] serve rate {rate}/s: {json.dumps(summ)}", file=sys.stderr, flush=True)
    top = max(per_rate, key=lambda r: r["tomograms_per_s"])
    updates = sum(pool.get(s)["updates"] for s in sizes)
    result = {
        "metric": METRIC,
        "value": top["gvoxel_tilt_updates_per_s"],
        "unit": "Gvoxel-tilt updates/s",
        "n_gpus": 1,
        "steps": n_req,
        "warmup": 1,
        "ms_per_step": round(1e3 * top["makespan_s"] / n_req, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded ellipsoid phantoms, analytic projections)",
        "config": {"workload": SERVE_WORKLOAD % (c["iters"], c["subsets"]) +
                   "; %d requests per rate, rates %s/s" % (n_req, ",".join("%g" % r for r in rates)),
                   "updates_per_stream": updates, "projection_storage": "f16",
                   "max_active": args.max_active or c["max_active"],
                   "schedule": args.schedule},
        "serving": per_rate,
    }
    print(json.dumps(result))


def serve_block_default():
    """The serving block of the default benchmark line: a short C5-shaped
    stream (SERVE_DEFAULT requests) at a rate the GPU sustains, for p50/p99
    latency, and the same requests as a burst, for throughput against the
    plain per-series FP32 path."""
    c, d = dict(SERVE), dict(SERVE_DEFAULT)
    pool = ServePool(c["seed"], c["iters"], c["subsets"])
    sizes = serve_sizes(np.random.default_rng(c["seed"] + 7), d["requests"], c["large_frac"])
    for sz in sorted(set(sizes)):
        pool.get(sz)
        pool.plain(sz)
    out = {}
    for name, rate in (("steady", d["rate"]), ("burst", d["burst_rate"])):
        recs, makespan, stats, _ = serve_stream(pool, sizes, rate, c["seed"] + 11, c["max_active"])
        out[name] = serve_summary(pool, sizes, recs, makespan, stats, rate)
        print(f"[bench] serve {name}: {json.dumps(out[name])}", file=sys.stderr, flush=True)
    return {"workload": SERVE_WORKLOAD % (c["iters"], c["subsets"]) +
            "; %d requests, steady at %g/s and burst at %g/s" % (d["requests"], d["rate"],
                                                                   d["burst_rate"]),
            "tomograms_per_s": out["burst"]["tomograms_per_s"],
            "latency_ms_p50": out["steady"]["latency_ms_p50"],
            "latency_ms_p99": out["steady"]["latency_ms_p99"],
            "speedup_vs_plain": out["burst"]["speedup_vs_plain"],
            "plain_fp32_per_series_tomograms_per_s": out["burst"]["plain_fp32_per_series_tomograms_per_s"],
            "steady": out["steady"], "burst": out["burst"]}


BATCH = dict(series=16, seed=20261205, iters=15, subsets=5)


def batch_bench(args):
    """The batching configuration: 16 seeded series of uneven sizes (tilts
    {31,41,51,61} over +-60 deg, projections {512,768,1024}^2, thickness
    {128,256,384}, tilt axis uniform 80-95 deg + shifts +-3 px, 200-800
    ellipsoids, noise 10%), FP32 projection storage.  A step is every series'
    WBP, then the SART of all 16 (15 iterations x 5 interleaved subsets, relax
    0.3, positivity off) as persistent launches per SART pass (one per staging
    variant and tile shape among the series), the passes replayed from a HIP
    graph; forward-projection band rows per series from
    the autotuner.  Reported beside the same work run series by series."""
    from aretomo3_amd import Batch, DeviceArray, Plan, synchronize
    from aretomo3_amd import _lib as L
    from aretomo3_amd import phantom
    c = dict(BATCH)
    rng = np.random.default_rng(c["seed"])
    sart_plans, wbp_plans, d_in, d_s, d_w, updates, shapes = [], [], [], [], [], 0.0, []
    t_gen = time.time()
    for k in range(c["series"]):
        n = int(rng.choice([31, 41, 51, 61]))
        px = int(rng.choice([512, 768, 1024]))
        vz = int(rng.choice([128, 256, 384]))
        axis = float(rng.uniform(80.0, 95.0))
        tilts = np.linspace(-60.0, 60.0, n).astype(np.float32)
        raw, shifts = phantom.tilt_series(c["seed"] + k, px, px, tilts, vz,
                                          n_ellipsoids=int(rng.integers(200, 801)), axis_min=4.0,
                                          axis_max=40.0, tilt_axis=axis, max_shift=3.0,
                                          noise_frac=0.10, n_threads=16)
        geo = dict(tilt_axis=np.full(n, axis, np.float32), shifts=shifts, fill_seed=c["seed"] + k)
        ps = Plan(px, px, tilts, vz, L.METHOD_SART, sart_iters=c["iters"], sart_subsets=c["subsets"],
                  relax=0.3, subset_order=L.SUBSETS_INTERLEAVED, positivity=False, **geo)
        if not args.no_tune:
            ps.autotune()   # each batch entry runs its own series' tile shapes
        pw = Plan(px, px, tilts, vz, L.METHOD_WBP, **geo)
        sart_plans.append(ps)
        wbp_plans.append(pw)
        d_in.append(DeviceArray.from_numpy(raw))
        d_s.append(DeviceArray(4 * ps.vol_x * ps.vol_y * ps.vol_z))
        d_w.append(DeviceArray(4 * pw.vol_x * pw.vol_y * pw.vol_z))
        vox = float(ps.vol_x) * ps.vol_y * ps.vol_z
        updates += vox * n * (1 + 1 + 2 * c["iters"])   # WBP + SART passes
        shapes.append((n, px, vz))
        del raw
    print(f"[bench] batch16: {len(shapes)} series {shapes} generated in {time.time() - t_gen:.1f}s",
          file=sys.stderr, flush=True)

    def run_plain():
        for k in range(len(sart_plans)):
            wbp_plans[k].execute(d_in[k], d_w[k])
            sart_plans[k].execute(d_in[k], d_s[k])

    batch = Batch(sart_plans, graph=True)

    def run_batch():
        for k in range(len(wbp_plans)):
            wbp_plans[k].execute(d_in[k], d_w[k])
        batch.execute(d_in, d_s)

    run_plain()
    synchronize()
    t0 = time.perf_counter()
    run_plain()
    synchronize()
    t_plain = time.perf_counter() - t0
    ref = [d_s[k].to_numpy(sart_plans[k].volume_shape) for k in (0, len(sart_plans) - 1)]
    for _ in range(max(args.warmup, 1)):
        run_batch()
    synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run_batch()
    synchronize()
    wall = (time.perf_counter() - t0) / args.steps
    errs = []
    for ref_k, k in zip(ref, (0, len(sart_plans) - 1)):
        got = d_s[k].to_numpy(sart_plans[k].volume_shape)
        errs.append(float(np.linalg.norm((got - ref_k).ravel()) / np.linalg.norm(ref_k.ravel())))
    batch.close()
    for p_ in sart_plans + wbp_plans:
        p_.close()
    print(f"[bench] batch16: {args.steps} steps, {wall:.3f} s/step (plain {t_plain:.3f} s)",
          file=sys.stderr, flush=True)
    result = {
        "metric": METRIC,
        "value": round(updates / wall / 1e9, 2),
        "unit": "Gvoxel-tilt updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": max(args.warmup, 1),
        "ms_per_step": round(1e3 * wall, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded ellipsoid phantoms, analytic projections)",
        "config": {
            "workload": ("batch16: 16 series (tilts {31,41,51,61} over +-60 deg, projections "
                         "{512,768,1024}^2, thickness {128,256,384}, tilt axis 80-95 deg + shifts "
                         "+-3 px, 200-800 ellipsoids, noise 10%%), per step WBP of each series then "
                         "SART %d iterations x %d interleaved subsets (relax 0.3, positivity off) of "
                         "all 16 in persistent batched launches replayed from a HIP graph, FP32 "
                         "projection storage" % (c["iters"], c["subsets"])),
            "series": [{"tilts": n, "projection": px, "thickness": vz} for n, px, vz in shapes],
            "updates_per_step": updates,
            "projection_storage": "f32",
        },
        "batching": {
            "plain_per_series_s": round(t_plain, 3),
            "batched_s": round(wall, 3),
            "speedup_vs_plain": round(t_plain / wall, 3),
            "batched_vs_plain_rel_l2_max": max(errs),
        },
    }
    print(json.dumps(result))


def extras_bench(args):
    """Measurements of the steps either side of the reconstruction, each with
    its algorithmic bytes per call against the HBM peak and the FP64 oracle on
    the host cores as the CPU baseline (bounded samples, scaled):

    * pre-stage with local patch shifts and -AtBin 4 Fourier binning of a
      4096^2 x 41 series (tilt axis 86 deg, shifts +-3 px, 25 patches);
    * the ProjAlign re-projector on a 1024^2 x 41 stack, volume 256 thick,
      every projection predicted in projection-matching order;
    * the sample-thickness measurement (CCalcVolThick) of a 2048^2 x 41
      series, binning 20;
    * MRC write and read of a 1 GiB volume (host I/O, no GPU)."""
    import tempfile

    import oracle
    from aretomo3_amd import CalcVolThick, DeviceArray, Plan, files, phantom, synchronize
    from aretomo3_amd import _lib as L
    from aretomo3_amd.projalign import CalcReproj, find_proj_range
    out = {}
    threads = oracle.max_threads()
    rng = np.random.default_rng(20261207)

    # ---- pre-stage: local shifts + Fourier binning ----
    px, n, binv, npat = 4096, 41, 4.0, 25
    tilts = phantom.tilt_range(-60.0, 60.0, 3.0).astype(np.float32)
    raw, shifts = phantom.tilt_series(20261207, px, px, tilts, 512, n_ellipsoids=300, axis_min=8.0,
                                      axis_max=80.0, tilt_axis=86.0, max_shift=3.0, noise_frac=0.1,
                                      n_threads=16)
    g = np.linspace(-0.4, 0.4, 5) * px
    local = np.empty((n, 5, npat), np.float32)
    local[:, 0] = np.tile(np.repeat(g, 5), (n, 1))
    local[:, 1] = np.tile(np.tile(g, 5), (n, 1))
    local[:, 2:4] = rng.uniform(-2, 2, (n, 2, npat))
    local[:, 4] = 1.0
    axes = np.full(n, 86.0, np.float32)
    plan = Plan(px, px, tilts, 128, L.METHOD_WBP, tilt_axis=axes, shifts=shifts, local_params=local,
                bin=binv, fill_seed=7)
    d_raw = DeviceArray.from_numpy(raw)
    d_vol = DeviceArray(4 * plan.vol_x * plan.vol_y * plan.vol_z)
    plan.execute(d_raw, d_vol)
    synchronize()
    plan.reset_stats()
    plan.set_profiling(True)
    reps = 3
    for _ in range(reps):
        plan.execute(d_raw, d_vol)
    synchronize()
    nl, ms = plan.kernel_stats(L.K_PRESTAGE)
    plan.set_profiling(False)
    ms /= reps
    ax, ay = files.aligned_size(px, px, 86.0)
    bx, by = plan.vol_x, plan.vol_y
    kx = bx // 2 + 1
    bytes_pre = n * (4.0 * px * px + 4.0 * ax * ay + 4.0 * ax * ay + 16.0 * ay * kx + 16.0 * by * kx
                     + 4.0 * bx * by)
    plan.close()
    d_vol.free()
    t0 = time.perf_counter()
    k_cpu = 2
    al = oracle.prestage(raw[:k_cpu], axes[:k_cpu], shifts[:k_cpu], 1, 7, local=local[:k_cpu])
    oracle.bin_stack(al, binv)
    t_cpu = (time.perf_counter() - t0) / k_cpu
    out["prestage_local_bin"] = {
        "workload": "%d^2 x %d raw, tilt axis 86 deg + shifts +-3 px, %d local patches, "
                    "Fourier binning %g -> %dx%d" % (px, n, npat, binv, bx, by),
        "ms_per_series": round(ms, 3), "tilts_per_s": round(n / (ms * 1e-3), 1),
        "algorithmic_bytes": bytes_pre, "achieved_gbs": round(bytes_pre / (ms * 1e-3) / 1e9, 1),
        "hbm_frac": round(bytes_pre / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "cpu_baseline": {"tilts_per_s": round(1.0 / t_cpu, 3), "cores": threads, "kind": "port",
                         "sample": "FP64 oracle pre-stage + binning of %d tilts, %.1fs" % (k_cpu, t_cpu * k_cpu)},
    }
    del raw, d_raw, al
    print(f"[bench] extras prestage: {out['prestage_local_bin']}", file=sys.stderr, flush=True)

    # ---- ProjAlign re-projector ----
    px, vz = 1024, 256
    projs, _ = phantom.tilt_series(20261208, px, px, tilts, vz, n_ellipsoids=200, axis_min=4.0,
                                   axis_max=40.0, noise_frac=0.1, n_threads=16)
    rp = CalcReproj(px, px, n, vz)
    d_p = DeviceArray.from_numpy(projs)
    d_o = DeviceArray(4 * px * px)
    order = np.argsort(np.abs(tilts), kind="stable")[1:]
    jobs = []
    for idx in order:
        skip = (np.abs(tilts) >= abs(tilts[idx]) - 1e-6).astype(np.uint8)
        try:
            jobs.append((int(idx), skip, find_proj_range(tilts, idx, skip)))
        except L.AtrError:
            pass
    rp.execute(d_p, tilts, jobs[0][1], jobs[0][0], d_o)
    synchronize()
    t0 = time.perf_counter()
    for idx, skip, _ in jobs:
        rp.execute(d_p, tilts, skip, idx, d_o)
    synchronize()
    ms_r = 1e3 * (time.perf_counter() - t0) / len(jobs)
    nr = np.mean([b - a + 1 for _, _, (a, b) in jobs])
    bytes_r = 4.0 * nr * px * px + 2 * 4.0 * (2 * px) * vz * px + 4.0 * px * px
    rows = 8
    sub = np.ascontiguousarray(projs[:, px // 2 - rows // 2: px // 2 + rows // 2])
    idx, skip, _ = jobs[len(jobs) // 2]
    t0 = time.perf_counter()
    oracle.reproj(sub, tilts, vz, idx, skip)
    t_cpu = (time.perf_counter() - t0) * (px / rows)
    rp.close()
    out["reproj"] = {
        "workload": "%d^2 x %d stack, volume thickness %d, %d projections predicted "
                    "(projection-matching order, mean %.1f reference tilts)" % (px, n, vz, len(jobs), nr),
        "ms_per_reprojection": round(ms_r, 3),
        "algorithmic_bytes": bytes_r, "achieved_gbs": round(bytes_r / (ms_r * 1e-3) / 1e9, 1),
        "hbm_frac": round(bytes_r / (ms_r * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "cpu_baseline": {"ms_per_reprojection": round(1e3 * t_cpu, 1), "cores": threads, "kind": "port",
                         "sample": "FP64 oracle on %d of %d rows, scaled" % (rows, px)},
    }
    del projs, d_p, d_o
    print(f"[bench] extras reproj: {out['reproj']}", file=sys.stderr, flush=True)

    # ---- sample thickness ----
    px = 2048
    raw, shifts = phantom.tilt_series(20261209, px, px, tilts, 400, n_ellipsoids=400, axis_min=10.0,
                                      axis_max=60.0, tilt_axis=86.0, max_shift=3.0, noise_frac=0.1,
                                      n_threads=16)
    c = CalcVolThick()
    c.do_it(raw, tilts, 1.0, tilt_axis=axes, shifts=shifts, fill_seed=3)
    t0 = time.perf_counter()
    got = c.do_it(raw, tilts, 1.0, tilt_axis=axes, shifts=shifts, fill_seed=3)
    ms_t = 1e3 * (time.perf_counter() - t0)
    t0 = time.perf_counter()
    ref = oracle.calc_vol_thick(raw, tilts, axes, shifts, 1.0, seed=3)
    t_cpu = time.perf_counter() - t0
    out["thickness"] = {
        "workload": "%d^2 x %d raw, pixel 1 A (binning 20), sample slab 400 px" % (px, n),
        "ms_per_series": round(ms_t, 2), "thickness_px": got["thickness"],
        "same_edges_as_oracle": (got["low_edge"], got["high_edge"]) == ref["edges"][:2],
        "note": "host call: includes the H2D copy of the raw series",
        "cpu_baseline": {"ms_per_series": round(1e3 * t_cpu, 1), "cores": threads, "kind": "port",
                         "sample": "FP64 oracle pipeline on the same series"},
    }
    del raw
    print(f"[bench] extras thickness: {out['thickness']}", file=sys.stderr, flush=True)

    # ---- MRC I/O ----
    vol = np.random.default_rng(1).standard_normal((256, 1024, 1024), dtype=np.float32)
    with tempfile.TemporaryDirectory() as d:
        pth = os.path.join(d, "v_Vol.mrc")
        t0 = time.perf_counter()
        files.write_mrc(pth, vol, 1.0)
        t_w = time.perf_counter() - t0
        t0 = time.perf_counter()
        back = files.read_mrc(pth)
        t_r = time.perf_counter() - t0
        ok = bool(np.array_equal(back, vol))
    out["mrc_io"] = {"bytes": vol.nbytes, "write_gbs": round(vol.nbytes / t_w / 1e9, 2),
                     "read_gbs": round(vol.nbytes / t_r / 1e9, 2), "round_trip_identical": ok}
    result = {
        "metric": "pre-stage tilts/s (local shifts + Fourier binning) and per-step times of the "
                  "re-projector, thickness measurement and MRC I/O",
        "value": out["prestage_local_bin"]["tilts_per_s"],
        "unit": "tilts/s",
        "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": out["prestage_local_bin"]["ms_per_series"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32/f64 geometry",
        "data": "synthetic (seeded ellipsoid phantoms, analytic projections)",
        "config": {"workload": "extras: pre-stage (local + bin), re-projector, thickness, MRC I/O"},
        "extras": out,
    }
    print(json.dumps(result))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="sart1024", choices=sorted(CONFIGS) + ["serve", "batch16", "extras"])
    ap.add_argument("--requests", type=int, default=0, help="serve: number of requests")
    ap.add_argument("--rate", type=float, default=0.0, help="serve: arrivals per second")
    ap.add_argument("--rates", default="", help="serve: comma-separated arrival rates per second")
    ap.add_argument("--max-active", type=int, default=0, help="serve: series per round at most")
    ap.add_argument("--schedule", default="shortest", choices=sorted(SCHEDULES),
                    help="serve: which series step in a round (shortest-first, fair share or "
                    "every series)")
    ap.add_argument("--no-serve", action="store_true",
                    help="default run: leave out the serving block")
    ap.add_argument("--small", action="store_true", help="same as --config small")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default tile shapes")
    ap.add_argument("--cpu-slices", type=int, default=0,
                    help="y-slices the CPU baseline reconstructs (0 = the config's default)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run the plain copy kernel on MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the reconstruction would shard over y-slabs with no exchange between GPUs)",
              file=sys.stderr)
        sys.exit(2)
    if args.config == "serve":
        serve_bench(args)
        return
    if args.config == "batch16":
        batch_bench(args)
        return
    if args.config == "extras":
        extras_bench(args)
        return
    name = "small" if args.small else args.config
    w = CONFIGS[name]

    from aretomo3_amd import DeviceArray, Plan, synchronize
    from aretomo3_amd import _lib as L
    from aretomo3_amd import phantom

    sart = w["method"] == "sart"
    storage = {"f32": L.STORAGE_F32, "f16": L.STORAGE_F16, "bf16": L.STORAGE_BF16}[w["storage"]]
    tilts = phantom.tilt_range(-60.0, 60.0, w["step"])
    n = len(tilts)
    t0 = time.time()
    raw, shifts = phantom.tilt_series(w["seed"], w["px"], w["px"], tilts, w["vol_z"],
                                      n_ellipsoids=w["n_ell"], axis_min=4.0,
                                      axis_max=w["amax"], tilt_axis=w["axis"],
                                      max_shift=w["shift"], noise_frac=w["noise"],
                                      n_threads=16)
    print(f"[bench] {name}: generated {raw.shape} in {time.time() - t0:.1f}s", file=sys.stderr,
          flush=True)
    axes = np.full(n, w["axis"], np.float32)
    if sart:
        plan = Plan(w["px"], w["px"], tilts, w["vol_z"], L.METHOD_SART, tilt_axis=axes,
                    shifts=shifts, sart_iters=w["iters"], sart_subsets=w["subsets"],
                    relax=w["relax"], positivity=w["positivity"], storage=storage,
                    subset_order=(L.SUBSETS_INTERLEAVED if w["order"] == "interleaved"
                                  else L.SUBSETS_CONTIGUOUS), fill_seed=w["seed"])
    else:
        plan = Plan(w["px"], w["px"], tilts, w["vol_z"], L.METHOD_WBP, tilt_axis=axes,
                    shifts=shifts, storage=storage, fill_seed=w["seed"])
    vol_x, vol_y, vol_z = plan.vol_x, plan.vol_y, plan.vol_z
    tuned = None
    if sart and not args.no_tune:
        t0 = time.time()
        tuned = plan.autotune()
        print(f"[bench] autotuned tile shape (bp z rows, bp tilts/stage, fp band rows) = {tuned} "
              f"in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    d_raw = DeviceArray.from_numpy(raw)
    d_vol = DeviceArray(4 * vol_x * vol_y * vol_z)
    print(f"[bench] plan {plan.device_bytes / 2**30:.2f} GiB, volume {vol_x}x{vol_y}x{vol_z}",
          file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        plan.execute(d_raw, d_vol)
    synchronize()
    plan.reset_stats()
    plan.set_profiling(True)
    synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.execute(d_raw, d_vol)
    synchronize()
    wall = time.perf_counter() - t0
    plan.set_profiling(False)
    ms_step = 1e3 * wall / args.steps
    print(f"[bench] {args.steps} steps in {wall:.2f}s", file=sys.stderr, flush=True)

    passes = 1 + 2 * w["iters"] if sart else 1
    vox = float(vol_x) * vol_y * vol_z
    updates = vox * n * passes
    value = updates / (wall / args.steps) / 1e9
    s_p = 4.0 if w["storage"] == "f32" else 2.0

    kernels = {}
    for kname, k in (("bp", L.K_BP), ("fp_diff", L.K_FP), ("weight", L.K_WEIGHT),
                     ("prestage", L.K_PRESTAGE), ("output", L.K_OUTPUT),
                     ("wbp_bp_rowfft", L.K_WBP_BP), ("wbp_colfft", L.K_WBP_COL),
                     ("wbp_rowc2r", L.K_WBP_ROW)):
        nl, ms = plan.kernel_stats(k)
        if nl:
            kernels[kname] = {"launches": nl, "ms_per_step": round(ms / args.steps, 3)}
    if sart:
        # SART back-projection, algorithmic bytes per launch: 8 B per voxel
        # (read + write) + s_p per sinogram sample read, averaged over the
        # 1 + iterations x subsets launches of a reconstruction
        per_step = 1 + w["iters"] * w["subsets"]
        bytes_launch = (per_step * 8.0 * vox + s_p * (n + w["iters"] * n) * vol_x * vol_y) / per_step
        launches, kms = plan.kernel_stats(L.K_BP)
        rkernel = "k_back_proj (SART back-projection)"
    else:
        # WBP back-projection + row FFT, per launch (one y-slab): s_p per sinogram
        # sample read + 16 B per output voxel (2x-grid row spectra written)
        launches, kms = plan.kernel_stats(L.K_WBP_BP)
        bytes_launch = (s_p * n * vol_x * vol_y + 16.0 * vox) / max(launches / args.steps, 1)
        rkernel = "k_wbp_bp_rowfft (2x-grid back-projection + row FFT)"
    avg_ms = kms / max(launches, 1)
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
    if sart and "fp_diff" in kernels:
        # forward projection + difference, algorithmic bytes per launch: 4 B
        # per voxel read + s_p per raw sample read and per difference written
        # (one subset of tilts); VALU-bound, so the HBM fraction is low
        fl, fms = plan.kernel_stats(L.K_FP)
        fp_bytes = 4.0 * vox + 2.0 * s_p * (n / w["subsets"]) * vol_x * vol_y
        fp_ms = fms / max(fl, 1)
        kernels["fp_diff"].update({"avg_launch_ms": round(fp_ms, 4), "bytes_per_launch": fp_bytes,
                                   "hbm_frac": round(fp_bytes / (fp_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)})
        # flop model of both kernels against the FP32 vector peak: 8 flop per
        # back-projection update, 14 per forward ray sample (L = VolZ/cos + 1.5
        # samples per ray), averaged per launch
        cos_t = np.cos(np.float32(0.017453) * tilts.astype(np.float32)).astype(np.float64)
        samples = w["iters"] * vol_y * vol_x * float(np.sum((vol_z / cos_t + 1.5).astype(np.int64)))
        fp_flop = 14.0 * samples / (w["iters"] * w["subsets"])
        bp_flop = 8.0 * (1 + w["iters"]) * vox * n / (1 + w["iters"] * w["subsets"])
        for kname, flop, ms in (("fp_diff", fp_flop, fp_ms), ("bp", bp_flop, avg_ms)):
            tf = flop / (ms * 1e-3) / 1e12
            kernels[kname]["flop_model"] = {"flop_per_launch": flop, "tflops": round(tf, 2),
                                            "frac_of_fp32_vector_peak_157.3": round(tf / 157.3, 4)}
    traffic = measured_traffic() if name == "sart1024" else None
    copy_peak = measured_copy_peak()
    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Gvoxel-tilt updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded ellipsoid phantom, analytic projections)",
        "config": {
            "workload": describe(name, w, (vol_x, vol_y, vol_z)),
            "updates_per_step": updates,
            "projection_storage": w["storage"],
            "tile_shape": None if tuned is None else {"bp_z_rows": tuned[0], "bp_tilts_per_stage": tuned[1],
                                                      "fp_band_rows": tuned[2],
                                                      "fp_band_prefetch": tuned[3], "source": "autotuner"},
        },
        "roofline": {
            "kernel": rkernel,
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": (TRAFFIC_FILE + " (rocprofv3 per launch: 128 B x TCC_EA0_RDREQ_sum "
                               "read + WRITE_SIZE written; FETCH_SIZE counts these 128-byte "
                               "requests as 64 B)") if traffic else None,
            "bytes_per_launch": bytes_launch,
            "avg_launch_ms": round(avg_ms, 4),
            "measured_copy_peak": copy_peak,
        },
        "kernels": kernels,
    }

    if not args.no_cpu_baseline:
        import oracle
        vol = d_vol.to_numpy((vol_y, vol_z, vol_x))
        n_sl = min(args.cpu_slices or w["cpu_slices"], vol_y)
        slices = np.linspace(0, vol_y - 1, n_sl).round().astype(np.int32)
        threads = oracle.max_threads()
        t0 = time.perf_counter()
        aligned = oracle.prestage(raw, axes, shifts, 1, w["seed"])
        t_pre = time.perf_counter() - t0
        t0 = time.perf_counter()
        if sart:
            ref = oracle.recon_slices(aligned, tilts, 1, vol_z, slices, iters=w["iters"],
                                      n_subsets=w["subsets"], relax=w["relax"],
                                      order=1 if w["order"] == "interleaved" else 0,
                                      positivity=1 if w["positivity"] else 0, seed=w["seed"])
        else:
            ref = oracle.recon_slices(aligned, tilts, 0, vol_z, slices)
        t_cpu = time.perf_counter() - t0
        cpu_updates = float(vol_x) * len(slices) * vol_z * n * passes
        got = vol[slices].astype(np.float64)
        rel = float(np.linalg.norm((got - ref).ravel()) / np.linalg.norm(ref.ravel()))
        result["cpu_baseline"] = {
            "value": round(cpu_updates / t_cpu / 1e9, 4),
            "unit": "Gvoxel-tilt updates/s",
            "cores": threads,
            "nproc": len(os.sched_getaffinity(0)),
            "kind": "port",
            "sample": "FP64 OpenMP oracle, %d of %d y-slices of the same series (%s only; "
                      "pre-stage of all tilts took %.1fs, not counted), %.1fs" %
                      (len(slices), vol_y, "SART" if sart else "WBP", t_pre, t_cpu),
        }
        result["rel_l2_vs_fp64"] = rel
        frc = fourier_ring_correlation(got, ref)
        result["frc_vs_fp64"] = {"min_to_nyquist": round(float(frc.min()), 6),
                                 "at_half_nyquist": round(float(frc[len(frc) // 2]), 6),
                                 "rings": len(frc)}
    if name == "sart1024" and not args.no_serve:
        # the latency half of the metric: a short serving stream of the
        # serving configuration's request mix, measured in the same run
        plan.close()
        d_raw.free()
        d_vol.free()
        result["serving"] = serve_block_default()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
