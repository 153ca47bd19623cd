#!/usr/bin/env python3
"""Benchmark of the SART reconstruction hot path on one MI355X.

Workload (the headline SART series): seeded phantom of 200 random ellipsoids
(semi-axes 4-40 voxels, density 0.5-1.0), 61 tilts -60..+60 deg step 2,
1024x1024 raw projections with tilt axis 85 deg and per-tilt shifts uniform
+-3 px (supplied as alignment, so the pre-stage runs), Gaussian noise of 10%
of the mean; volume 1024 x 1024 x 256; 20 SART iterations, 5 interleaved
subsets, relaxation 0.3, positivity off; FP32 projection storage.

A step is one whole reconstruction of that series with the raw projections
already resident in HBM: pre-stage correction + fill, weighting, initial
back-projection and 20 x 5 forward/difference/back-projection passes, and the
volume written out.  The metric counts output-resolution voxel-tilt updates:
VolX*VolY*VolZ*N per pass, 1 + 2*iterations passes.

Prints one JSON line.  Options: --steps K --warmup W --gpus N (only 1),
--no-cpu-baseline, --cpu-slices S, --small (a reduced configuration for
quick checks; not the benchmark).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gvoxel-tilt updates/s at ≤2e-3 rel-L2 vs FP64; p50/p99 tomogram latency"
HBM_PEAK_GBS = 8000.0


def workload(small):
    if small:
        return dict(px=256, py=256, vol_z=64, step=2.0, n_ell=60, amax=20.0, iters=4,
                    subsets=5, seed=20261204)
    return dict(px=1024, py=1024, vol_z=256, step=2.0, n_ell=200, amax=40.0, iters=20,
                subsets=5, seed=20261203)


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-slices", type=int, default=32)
    ap.add_argument("--small", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run the plain copy kernel on MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the reconstruction shards over y-slabs with no exchange; see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)

    from aretomo3_amd import DeviceArray, Plan, synchronize
    from aretomo3_amd import _lib as L
    from aretomo3_amd import phantom

    w = workload(args.small)
    tilts = phantom.tilt_range(-60.0, 60.0, w["step"])
    n = len(tilts)
    t0 = time.time()
    raw, shifts = phantom.tilt_series(w["seed"], w["px"], w["py"], tilts, w["vol_z"],
                                      n_ellipsoids=w["n_ell"], axis_min=4.0,
                                      axis_max=w["amax"], tilt_axis=85.0, max_shift=3.0,
                                      noise_frac=0.10)
    t_gen = time.time() - t0
    print(f"[bench] generated {raw.shape} in {t_gen:.1f}s", file=sys.stderr, flush=True)
    axes = np.full(n, 85.0, np.float32)
    plan = Plan(w["px"], w["py"], tilts, w["vol_z"], L.METHOD_SART, tilt_axis=axes,
                shifts=shifts, sart_iters=w["iters"], sart_subsets=w["subsets"], relax=0.3,
                subset_order=L.SUBSETS_INTERLEAVED, positivity=False, fill_seed=w["seed"])
    vol_x, vol_y, vol_z = plan.vol_x, plan.vol_y, plan.vol_z
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

    passes = 1 + 2 * w["iters"]
    updates = float(vol_x) * vol_y * vol_z * n * passes
    value = updates / (wall / args.steps) / 1e9

    # roofline of the SART back-projection (algorithmic bytes per launch:
    # 8 B per voxel read-modify-write + 4 B per projection sample read)
    bp_launches, bp_ms = plan.kernel_stats(L.K_BP)
    fp_launches, fp_ms = plan.kernel_stats(L.K_FP)
    per_step_bp = 1 + w["iters"] * w["subsets"]
    vox = float(vol_x) * vol_y * vol_z
    proj_samples_step = float(n + w["iters"] * n) * plan.vol_x * vol_y
    bytes_step = per_step_bp * 8.0 * vox + 4.0 * proj_samples_step
    bytes_launch = bytes_step / per_step_bp
    avg_ms = bp_ms / max(bp_launches, 1)
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
    kernels = {}
    for name, k in (("bp", L.K_BP), ("fp_diff", L.K_FP), ("weight", L.K_WEIGHT),
                    ("prestage", L.K_PRESTAGE), ("output", L.K_OUTPUT)):
        nl, ms = plan.kernel_stats(k)
        kernels[name] = {"launches": nl, "ms_per_step": round(ms / args.steps, 3)}

    # accuracy: GPU volume vs the FP64 oracle on sampled slices
    vol = d_vol.to_numpy((vol_y, vol_z, vol_x))
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
            "workload": ("SART one series: %dx%d raw projections, %d tilts -60..60 step %g, "
                         "tilt axis 85 deg + shifts +-3 px, volume %dx%dx%d, %d iterations x "
                         "%d interleaved subsets, relax 0.3, positivity off, FP32 storage")
            % (w["px"], w["py"], n, w["step"], vol_x, vol_y, vol_z, w["iters"], w["subsets"]),
            "updates_per_step": updates,
            "reduced": bool(args.small),
        },
        "roofline": {
            "kernel": "k_back_proj (SART back-projection)",
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "bytes_per_launch": bytes_launch,
            "avg_launch_ms": round(avg_ms, 4),
        },
        "kernels": kernels,
    }

    if not args.no_cpu_baseline:
        import oracle
        slices = np.linspace(0, vol_y - 1, min(args.cpu_slices, vol_y)).round().astype(np.int32)
        threads = oracle.max_threads()
        t0 = time.perf_counter()
        aligned = oracle.prestage(raw, axes, shifts, 1, w["seed"])
        t_pre = time.perf_counter() - t0
        t0 = time.perf_counter()
        ref = oracle.recon_slices(aligned, tilts, 1, vol_z, slices, iters=w["iters"],
                                  n_subsets=w["subsets"], relax=0.3, order=1, positivity=0,
                                  seed=w["seed"])
        t_cpu = time.perf_counter() - t0
        cpu_updates = float(vol_x) * len(slices) * vol_z * n * passes
        from_vol = vol[slices]
        num = np.linalg.norm((from_vol.astype(np.float64) - ref).ravel())
        rel = float(num / np.linalg.norm(ref.ravel()))
        result["cpu_baseline"] = {
            "value": round(cpu_updates / t_cpu / 1e9, 4),
            "unit": "Gvoxel-tilt updates/s",
            "cores": threads,
            "kind": "port",
            "sample": "FP64 OpenMP oracle, %d of %d y-slices of the same series (SART only; "
                      "pre-stage of all tilts took %.1fs, not counted), %.1fs" %
                      (len(slices), vol_y, t_pre, t_cpu),
        }
        result["rel_l2_vs_fp64"] = rel
    print(json.dumps(result))


if __name__ == "__main__":
    main()
