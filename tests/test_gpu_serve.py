"""The serving loop (atr_server_*): requests of different sizes arriving
together or one by one, admitted by the HBM budget and reconstructed in
shared rounds, give the same volumes as each series alone; and FP16 storage
stays within 2e-3 of the FP64 oracle."""
import numpy as np
import pytest

import oracle
from aretomo3_amd import Plan, Server, phantom
from aretomo3_amd import _lib as L
from helpers import rel_l2

pytestmark = pytest.mark.gpu

# (seed, px, py, tilt step, vol_z, iters, subsets, tilt axis or None, wbp)
REQS = [
    (201, 96, 5, 6.0, 32, 3, 3, None, True),
    (202, 64, 40, 4.0, 24, 2, 5, 85.0, False),
    (203, 120, 3, 10.0, 40, 4, 2, None, True),
    (204, 48, 6, 3.0, 16, 2, 4, None, False),
    (205, 80, 48, 5.0, 20, 3, 3, 88.0, True),
]


def series(spec):
    seed, px, py, step, vz, iters, subsets, axis, wbp = spec
    tilts = phantom.tilt_range(-60, 60, step)
    raw, shifts = phantom.tilt_series(seed, px, py, tilts, vz, n_ellipsoids=15, axis_min=3,
                                      axis_max=9, tilt_axis=float("nan") if axis is None else axis,
                                      max_shift=0.0 if axis is None else 2.0, noise_frac=0.05)
    kw = dict(sart_iters=iters, sart_subsets=subsets, relax=0.3,
              subset_order=L.SUBSETS_INTERLEAVED, positivity=False, fill_seed=seed)
    if axis is not None:
        kw.update(tilt_axis=np.full(len(tilts), axis, np.float32), shifts=shifts)
    return raw, tilts, vz, kw, wbp


def alone(raw, tilts, vz, kw, method, storage):
    n, py, px = raw.shape
    plan = Plan(px, py, tilts, vz, method, storage=storage, **kw)
    out = plan.execute_host(raw)
    plan.close()
    return out


# a longer series (61 tilts, 160 x 5 SART steps: ~800 rounds) that is still
# running when the small requests behind it have been prepared (a few ms of
# host work each), so they join its rounds
LEAD = (200, 256, 64, 2.0, 96, 160, 5, 86.0, True)


@pytest.mark.parametrize("schedule", [L.SCHEDULE_FAIR_SHARE, L.SCHEDULE_EVERY_STEP,
                                      L.SCHEDULE_SHORTEST_FIRST])
def test_serving_matches_alone(gpu_available, schedule):
    reqs = [series(r) for r in [LEAD] + REQS]
    print(f"case serving mix schedule={schedule}", flush=True)
    srv = Server(storage=L.STORAGE_F16, schedule=schedule)
    ids = [srv.submit(raw, tilts, vz, with_wbp=wbp, **kw) for raw, tilts, vz, kw, wbp in reqs]
    results = [srv.wait(i) for i in ids]
    stats = srv.stats()
    srv.close()
    print("stats", stats)
    if schedule == L.SCHEDULE_SHORTEST_FIRST:
        # the small requests, admitted behind the long one, overtake it
        done = [info["done_ms"] for _, _, info in results]
        assert done[0] == max(done)
    else:
        assert stats["max_batch"] >= 2
    for (raw, tilts, vz, kw, wbp), (vol, wvol, info) in zip(reqs, results):
        ref = alone(raw, tilts, vz, kw, L.METHOD_SART, L.STORAGE_F16)
        err = rel_l2(vol, ref)
        print("sart vs alone", err, "identical", np.array_equal(vol, ref), "latency", info["latency_ms"])
        assert err <= 1e-6
        if wbp:
            wref = alone(raw, tilts, vz, kw, L.METHOD_WBP, L.STORAGE_F16)
            werr = rel_l2(wvol, wref)
            print("wbp vs alone", werr)
            assert werr <= 1e-6


def test_serving_same_alone_and_in_mix(gpu_available):
    reqs = [series(r) for r in REQS]
    raw, tilts, vz, kw, wbp = reqs[2]
    print("case serving alone vs mix", flush=True)
    srv = Server(storage=L.STORAGE_F16)
    rid = srv.submit(raw, tilts, vz, with_wbp=True, **kw)
    solo, solo_w, _ = srv.wait(rid)
    srv.close()
    srv = Server(storage=L.STORAGE_F16, max_active=3)
    ids = []
    for k, (r2, t2, v2, kw2, w2) in enumerate(reqs):
        ids.append(srv.submit(r2, t2, v2, with_wbp=(k == 2) or w2, **kw2))
    mixed = [srv.wait(i) for i in ids]
    srv.close()
    assert rel_l2(mixed[2][0], solo) <= 1e-6
    assert rel_l2(mixed[2][1], solo_w) <= 1e-6
    # FP16 storage against the FP64 oracle
    ref = oracle.recon_slices(raw.astype(np.float64), tilts, 1, vz, list(range(raw.shape[1])),
                              iters=kw["sart_iters"], n_subsets=kw["sart_subsets"], relax=0.3,
                              order=1, positivity=0, seed=kw["fill_seed"])
    err = rel_l2(solo, ref)
    print("fp16 serving vs oracle", err)
    assert err <= 2e-3


def test_serving_budget_admission(gpu_available):
    reqs = [series(r) for r in REQS[:3]]
    print("case serving budget", flush=True)
    need = max(Plan.estimate_bytes(raw.shape[2], raw.shape[1], tilts, vz, storage=L.STORAGE_F16,
                                   **kw) for raw, tilts, vz, kw, _ in reqs)
    # room for about one series at a time: admission serialises the requests
    srv = Server(hbm_budget=int(need * 1.5) + (8 << 20), storage=L.STORAGE_F16)
    ids = [srv.submit(raw, tilts, vz, **kw) for raw, tilts, vz, kw, _ in reqs]
    res = [srv.wait(i) for i in ids]
    srv.close()
    for (raw, tilts, vz, kw, _), (vol, _, info) in zip(reqs, res):
        assert rel_l2(vol, alone(raw, tilts, vz, kw, L.METHOD_SART, L.STORAGE_F16)) <= 1e-6
    # a request that can never fit is refused at submission
    srv = Server(hbm_budget=4096, storage=L.STORAGE_F16)
    raw, tilts, vz, kw, _ = reqs[0]
    with pytest.raises(L.AtrError):
        srv.submit(raw, tilts, vz, **kw)
    srv.close()


def test_serving_isolates_a_failed_request(gpu_available):
    # a request whose plan cannot be built (a 95-degree tilt: no forward-ray
    # geometry) fails alone, on the host, before any kernel runs; the requests
    # around it complete and match their stand-alone volumes
    good = [series(r) for r in (REQS[0], REQS[3])]
    raw, tilts, vz, kw, _ = series(REQS[3])
    bad_tilts = tilts.copy()
    bad_tilts[-1] = 95.0
    print("case serving failure isolation", flush=True)
    srv = Server(storage=L.STORAGE_F16, max_active=2)
    ids = [srv.submit(good[0][0], good[0][1], good[0][2], **good[0][3]),
           srv.submit(raw, bad_tilts, vz, **kw),
           srv.submit(good[1][0], good[1][1], good[1][2], **good[1][3])]
    vol0, _, info0 = srv.wait(ids[0])
    with pytest.raises(L.AtrError, match="SART needs every tilt"):
        srv.wait(ids[1])
    vol2, _, info2 = srv.wait(ids[2])
    srv.close()
    for (r, t, z, k, _), vol in zip(good, (vol0, vol2)):
        assert rel_l2(vol, alone(r, t, z, k, L.METHOD_SART, L.STORAGE_F16)) <= 1e-6
