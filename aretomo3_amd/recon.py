"""Host-side mirror of AreTomo3's reconstruction operators over the C ABI.

The class and argument names follow the reference so that callers and tests
read like its own code:

* :class:`TomoWbp`, :class:`TomoSart` -- the per-slice operators
  ``CTomoWbp``/``CTomoSart`` (AreTomo/Recon/CReconInc.h:173-236): ``setup``
  once, then ``do_it(pad_sinogram, vol_xz, stream)`` on device buffers.
* :class:`DoWbpRecon`, :class:`DoSartRecon` -- the per-series drivers
  ``CDoWbpRecon``/``CDoSartRecon`` (AreTomo/Recon/CDoWbpRecon.cpp:28-47,
  CDoSartRecon.cpp:30-57): host tilt series in, host volume out in the
  reference's XZY frame layout (z flipped).
* :class:`Plan` -- the general plan (pre-stage, storage, layouts, SART
  options) behind all of them.

Every computation runs in libatr_recon.so on the GPU; this module only moves
arguments.  Errors surface as :class:`AtrError` carrying the C status.
"""
import ctypes

import numpy as np

from . import _lib as L
from ._lib import AtrError, check, lib


def _fptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


class DeviceArray:
    """A device allocation made through the library (no other runtime)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(lib().atr_malloc(ctypes.byref(p), max(self.nbytes, 4)))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, arr):
        arr = np.ascontiguousarray(arr)
        d = cls(arr.nbytes)
        if arr.nbytes:
            check(lib().atr_memcpy_h2d(d.ptr, arr.ctypes.data, arr.nbytes))
        return d

    def to_numpy(self, shape, dtype=np.float32):
        out = np.empty(shape, dtype)
        if out.nbytes:
            check(lib().atr_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        check(lib().atr_memcpy_h2d(self.ptr, arr.ctypes.data, arr.nbytes))

    def free(self):
        if self.ptr:
            check(lib().atr_free(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_count():
    n = ctypes.c_int(0)
    st = lib().atr_device_count(ctypes.byref(n))
    return n.value if st == L.ATR_OK else 0


def synchronize():
    check(lib().atr_synchronize())


def series_desc(proj_x, proj_y, tilts, vol_z, method=L.METHOD_SART, *, tilt_axis=None,
                shifts=None, random_fill=True, exclude=None, sart_iters=20, sart_subsets=1,
                sart_start_tilt=0, sart_n_tilts=0, relax=-1.0,
                subset_order=L.SUBSETS_CONTIGUOUS, positivity=True, storage=L.STORAGE_F32,
                layout=L.LAYOUT_XZY, fill_seed=0, tilt_stride=0, row_stride=0,
                local_params=None, bin=1.0):
    """An atr_series_desc and the arrays it points to (keep both alive).

    local_params: per-tilt patch tables [n][5][n_patches] (centres x, centres
    y, shifts x, shifts y, good flags) of CLocalAlignParam; bin: -AtBin
    Fourier binning of the aligned images (1 = none)."""
    t = np.ascontiguousarray(tilts, np.float32)
    n = len(t)
    keep = [t]
    d = L.SeriesDesc()
    lib().atr_series_desc_init(ctypes.byref(d))
    d.proj_x, d.proj_y, d.n_tilts = int(proj_x), int(proj_y), n
    d.tilt_stride, d.row_stride = int(tilt_stride), int(row_stride)
    d.tilt_deg = t.ctypes.data
    if tilt_axis is not None:
        ax = np.ascontiguousarray(np.broadcast_to(np.asarray(tilt_axis, np.float32), (n,)))
        keep.append(ax)
        d.tilt_axis_deg = ax.ctypes.data
    if shifts is not None:
        sh = np.ascontiguousarray(shifts, np.float32).reshape(-1)
        keep.append(sh)
        d.shift_xy = sh.ctypes.data
    d.random_fill = 1 if random_fill else 0
    if exclude is not None:
        ex = np.ascontiguousarray(exclude, np.uint8)
        keep.append(ex)
        d.exclude = ex.ctypes.data
    d.vol_z = int(vol_z)
    d.method = int(method)
    d.sart_iters, d.sart_subsets = int(sart_iters), int(sart_subsets)
    d.sart_start_tilt, d.sart_n_tilts = int(sart_start_tilt), int(sart_n_tilts)
    d.sart_relax = float(relax)
    d.subset_order, d.positivity = int(subset_order), 1 if positivity else 0
    d.storage, d.out_layout = int(storage), int(layout)
    d.fill_seed = int(fill_seed)
    if local_params is not None:
        lp = np.ascontiguousarray(local_params, np.float32).reshape(n, 5, -1)
        keep.append(lp)
        d.n_patches = int(lp.shape[2])
        d.local_params = lp.ctypes.data
    d.bin = float(bin)
    return d, keep


class Plan:
    """One reconstruction plan (atr_plan_create).  Arrays given here are
    copied into the descriptor; the plan keeps them alive."""

    def __init__(self, proj_x, proj_y, tilts, vol_z, method=L.METHOD_SART, **kw):
        d, self._keep = series_desc(proj_x, proj_y, tilts, vol_z, method, **kw)
        n = d.n_tilts
        layout = d.out_layout
        self.desc = d
        self.layout = int(layout)
        h = ctypes.c_void_p()
        check(lib().atr_plan_create(ctypes.byref(d), ctypes.byref(h)))
        self.handle = h.value
        vx, vy, vz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().atr_plan_get_volume_size(self.handle, ctypes.byref(vx), ctypes.byref(vy),
                                             ctypes.byref(vz)))
        self.vol_x, self.vol_y, self.vol_z = vx.value, vy.value, vz.value
        self.n_tilts = n
        self.input_elems = (d.tilt_stride or (d.row_stride or proj_x) * proj_y) * (n - 1) + \
            (d.row_stride or proj_x) * (proj_y - 1) + proj_x

    @staticmethod
    def estimate_bytes(proj_x, proj_y, tilts, vol_z, method=L.METHOD_SART, **kw):
        """Device bytes a plan with these arguments would allocate."""
        d, keep = series_desc(proj_x, proj_y, tilts, vol_z, method, **kw)
        b = ctypes.c_size_t()
        check(lib().atr_estimate_device_bytes(ctypes.byref(d), ctypes.byref(b)))
        return b.value

    def autotune(self, stream=None):
        """Times the candidate tile shapes on this plan and keeps the fastest;
        returns (bp z rows, bp tilts per stage, fp band rows, fp prefetch)."""
        tz, ch, bz, pf = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().atr_plan_autotune(self.handle, stream, ctypes.byref(tz), ctypes.byref(ch),
                                      ctypes.byref(bz), ctypes.byref(pf)))
        return tz.value, ch.value, bz.value, pf.value

    def set_tile_shape(self, bp_tz, bp_chunk, fp_band_rows, fp_prefetch=0):
        check(lib().atr_plan_set_tile_shape(self.handle, bp_tz, bp_chunk, fp_band_rows,
                                            int(fp_prefetch)))

    @property
    def volume_shape(self):
        if self.layout in (L.LAYOUT_XYZ, L.LAYOUT_XYZ_NOFLIP):
            return (self.vol_z, self.vol_y, self.vol_x)
        return (self.vol_y, self.vol_z, self.vol_x)

    @property
    def device_bytes(self):
        return int(lib().atr_plan_device_bytes(self.handle))

    def execute(self, proj_dev, vol_dev, accumulate=False, stream=None):
        """Stream-ordered run on device pointers (ints or DeviceArray)."""
        pp = proj_dev.ptr if isinstance(proj_dev, DeviceArray) else proj_dev
        vp = vol_dev.ptr if isinstance(vol_dev, DeviceArray) else vol_dev
        check(lib().atr_execute(self.handle, pp, vp, 1 if accumulate else 0, stream))

    def execute_host(self, proj):
        """Host numpy in, host numpy volume out (CDo*Recon::DoIt)."""
        proj = np.ascontiguousarray(proj, np.float32).reshape(-1)
        assert proj.size >= self.input_elems
        out = np.empty(self.volume_shape, np.float32)
        check(lib().atr_execute_host(self.handle, proj.ctypes.data, out.ctypes.data))
        return out

    def set_profiling(self, on=True):
        check(lib().atr_plan_set_profiling(self.handle, 1 if on else 0))

    def kernel_stats(self, kernel):
        n, ms = ctypes.c_int(), ctypes.c_double()
        check(lib().atr_plan_kernel_stats(self.handle, kernel, ctypes.byref(n), ctypes.byref(ms)))
        return n.value, ms.value

    def reset_stats(self):
        check(lib().atr_plan_reset_stats(self.handle))

    def close(self):
        if getattr(self, "handle", None):
            lib().atr_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Several SART plans reconstructed together (atr_batch_*): every
    back-projection / forward-projection step of all series is a few persistent
    launches (one per staging variant and tile shape); graph=True captures the iteration sequence in a HIP graph once and
    replays it.  The plans must outlive the batch."""

    def __init__(self, plans, graph=False):
        self.plans = list(plans)
        arr = (ctypes.c_void_p * len(self.plans))(*[p.handle for p in self.plans])
        h = ctypes.c_void_p()
        check(lib().atr_batch_create(arr, len(self.plans), L.BATCH_GRAPH if graph else 0,
                                     ctypes.byref(h)))
        self.handle = h.value

    def execute(self, projs, vols, stream=None):
        """projs / vols: one device buffer (DeviceArray or pointer) per plan."""
        def ptr(x):
            return x.ptr if isinstance(x, DeviceArray) else x
        ins = (ctypes.c_void_p * len(projs))(*[ptr(x) for x in projs])
        outs = (ctypes.c_void_p * len(vols))(*[ptr(x) for x in vols])
        check(lib().atr_batch_execute(self.handle, ins, outs, stream))

    def set_profiling(self, on=True):
        check(lib().atr_batch_set_profiling(self.handle, 1 if on else 0))

    def kernel_stats(self, kernel):
        n, ms = ctypes.c_int(), ctypes.c_double()
        check(lib().atr_batch_kernel_stats(self.handle, kernel, ctypes.byref(n), ctypes.byref(ms)))
        return n.value, ms.value

    def reset_stats(self):
        check(lib().atr_batch_reset_stats(self.handle))

    def close(self):
        if getattr(self, "handle", None):
            lib().atr_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PinnedArray:
    """Page-locked host memory (atr_host_alloc) viewed as a numpy array, so
    host<->device copies of the serving loop overlap the GPU work."""

    def __init__(self, shape, dtype=np.float32):
        nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
        p = ctypes.c_void_p()
        check(lib().atr_host_alloc(ctypes.byref(p), max(nbytes, 4)))
        self.ptr = p.value
        buf = (ctypes.c_char * max(nbytes, 4)).from_address(self.ptr)
        self.array = np.frombuffer(buf, dtype=dtype, count=int(np.prod(shape))).reshape(shape)

    def free(self):
        if self.ptr:
            self.array = None
            check(lib().atr_host_free(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Server:
    """The serving loop (atr_server_*): submit SART series (optionally with
    WBP) as they arrive; each is admitted by the HBM budget and reconstructed
    together with the others on the GPU.  Output arrays are allocated here."""

    def __init__(self, hbm_budget=0, max_active=None, storage=L.STORAGE_F16,
                 schedule=L.SCHEDULE_SHORTEST_FIRST):
        cfg = L.ServerConfig()
        lib().atr_server_config_init(ctypes.byref(cfg))
        cfg.hbm_budget_bytes, cfg.storage = int(hbm_budget), int(storage)
        cfg.schedule = int(schedule)
        if max_active is not None:
            cfg.max_active = int(max_active)
        h = ctypes.c_void_p()
        check(lib().atr_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h.value
        self._live = {}

    def submit(self, proj, tilts, vol_z, with_wbp=False, out=None, wbp_out=None, **kw):
        """proj: float32 [n][y][x] (numpy or PinnedArray.array, kept alive until
        wait); returns the request id.  kw: the Plan options of a SART series."""
        proj = np.ascontiguousarray(proj, np.float32)
        n, py, px = proj.shape
        d, keep = series_desc(px, py, tilts, vol_z, L.METHOD_SART, **kw)
        ax, ay = px, py
        if kw.get("tilt_axis") is not None:
            ax0 = float(np.asarray(kw["tilt_axis"], np.float32).reshape(-1)[0])
            if abs(np.sin(ax0 * 3.14 / 180.0)) > 0.707:
                ax, ay = py, px
        layout = kw.get("layout", L.LAYOUT_XZY)
        vx = ax // 2 * 2
        shape = (vol_z, ay, vx) if layout in (L.LAYOUT_XYZ, L.LAYOUT_XYZ_NOFLIP) else (ay, vol_z, vx)
        vol = out if out is not None else np.empty(shape, np.float32)
        wbp = None
        if with_wbp:
            wbp = wbp_out if wbp_out is not None else np.empty(shape, np.float32)
        rid = ctypes.c_uint64()
        check(lib().atr_server_submit(self.handle, ctypes.byref(d), 1 if with_wbp else 0,
                                      proj.ctypes.data, vol.ctypes.data,
                                      wbp.ctypes.data if wbp is not None else None,
                                      ctypes.byref(rid)))
        self._live[rid.value] = (proj, vol, wbp, keep)
        return rid.value

    def wait(self, rid):
        """Blocks until request rid is done: (sart volume, wbp volume or None, info)."""
        info = L.RequestInfo()
        try:
            check(lib().atr_server_wait(self.handle, rid, ctypes.byref(info)))
        finally:
            proj, vol, wbp, keep = self._live.pop(rid)
        return vol, wbp, {"submit_ms": info.submit_ms, "admit_ms": info.admit_ms,
                          "done_ms": info.done_ms, "latency_ms": info.latency_ms,
                          "prepared_ms": info.prepared_ms,
                          "compute_done_ms": info.compute_done_ms,
                          "queue_ms": info.admit_ms - info.submit_ms,
                          "gpu_ms": info.compute_done_ms - info.admit_ms,
                          "copy_out_ms": info.done_ms - info.compute_done_ms}

    def stats(self):
        r, m = ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().atr_server_stats(self.handle, ctypes.byref(r), ctypes.byref(m)))
        return {"rounds": r.value, "max_batch": m.value}

    def close(self):
        if getattr(self, "handle", None):
            lib().atr_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --------------------------------------------------------------------------
# per-slice operators (CTomoWbp / CTomoSart)
# --------------------------------------------------------------------------
class _TomoBase:
    def _make(self, vol_x, vol_z, proj_x, tilts, method, **kw):
        if vol_x != proj_x // 2 * 2:
            raise AtrError(1, f"iVolX must be ProjX/2*2 = {proj_x // 2 * 2}")
        self.proj_x = int(proj_x)
        self.pad_x = (self.proj_x // 2 + 1) * 2
        self.plan = Plan(self.proj_x, 1, tilts, vol_z, method, tilt_stride=self.pad_x,
                         row_stride=self.pad_x, layout=L.LAYOUT_SLICE, **kw)

    def clean(self):
        if getattr(self, "plan", None) is not None:
            self.plan.close()
            self.plan = None


class TomoWbp(_TomoBase):
    """CTomoWbp (AreTomo/Recon/CReconInc.h:173-192)."""

    def setup(self, vol_x, vol_z, proj_x, tilts, exclude=None):
        self.clean()
        self._make(vol_x, vol_z, proj_x, tilts, L.METHOD_WBP, exclude=exclude)

    def do_it(self, pad_sinogram, vol_xz, stream=None):
        """pad_sinogram: device [n_tilts][(ProjX/2+1)*2]; vol_xz: device [VolZ][VolX]."""
        self.plan.execute(pad_sinogram, vol_xz, False, stream)


class TomoSart(_TomoBase):
    """CTomoSart (AreTomo/Recon/CReconInc.h:194-236).  do_it continues from
    the volume already in vol_xz, as the reference does."""

    def setup(self, vol_x, vol_z, num_subsets, num_iters, proj_x, tilts, start_tilt, num_tilts,
              relax=-1.0, subset_order=L.SUBSETS_CONTIGUOUS, positivity=True, fill_seed=0,
              exclude=None, storage=L.STORAGE_F32):
        self.clean()
        self._make(vol_x, vol_z, proj_x, tilts, L.METHOD_SART, sart_iters=num_iters,
                   sart_subsets=num_subsets, sart_start_tilt=start_tilt, sart_n_tilts=num_tilts,
                   relax=relax, subset_order=subset_order, positivity=positivity,
                   fill_seed=fill_seed, exclude=exclude, storage=storage)

    def do_it(self, pad_sinogram, vol_xz, stream=None):
        self.plan.execute(pad_sinogram, vol_xz, True, stream)


# --------------------------------------------------------------------------
# per-series drivers (CDoWbpRecon / CDoSartRecon)
# --------------------------------------------------------------------------
class DoWbpRecon:
    """CDoWbpRecon::DoIt(pTiltSeries, pAlignParam, iVolZ) -> XZY volume."""

    def do_it(self, tilt_series, tilts, vol_z, **kw):
        ts = np.ascontiguousarray(tilt_series, np.float32)
        n, py, px = ts.shape
        plan = Plan(px, py, tilts, vol_z, L.METHOD_WBP, **kw)
        try:
            return plan.execute_host(ts)
        finally:
            plan.close()


class DoSartRecon:
    """CDoSartRecon::DoIt(pTiltSeries, pAlignParam, iStartTilt, iNumTilts,
    iVolZ, iIterations, iNumSubsets) -> XZY volume."""

    def do_it(self, tilt_series, tilts, start_tilt, num_tilts, vol_z, iterations, num_subsets,
              **kw):
        ts = np.ascontiguousarray(tilt_series, np.float32)
        n, py, px = ts.shape
        plan = Plan(px, py, tilts, vol_z, L.METHOD_SART, sart_iters=iterations,
                    sart_subsets=num_subsets, sart_start_tilt=start_tilt,
                    sart_n_tilts=num_tilts, **kw)
        try:
            return plan.execute_host(ts)
        finally:
            plan.close()


# --------------------------------------------------------------------------
# sample thickness (CCalcVolThick)
# --------------------------------------------------------------------------
def thickness_edges(std_profile, binning):
    """CCalcVolThick::mSmooth + mDetectEdges on a std profile ->
    (atr_thickness fields as a dict, smoothed profile)."""
    p = np.ascontiguousarray(std_profile, np.float32)
    out = L.Thickness()
    sm = np.empty_like(p)
    check(lib().atr_thickness_edges(p.ctypes.data, len(p), float(binning), ctypes.byref(out),
                                    sm.ctypes.data))
    return {k: getattr(out, k) for k, _ in L.Thickness._fields_}, sm


class CalcVolThick:
    """CCalcVolThick: DoIt on a raw tilt series (CCalcVolThick.cpp:47-90),
    then GetThickness / GetLowEdge / GetHighEdge (:28-45)."""

    def __init__(self):
        self.result = None
        self.profile = None

    def do_it(self, raw, tilts, pixel_size, tilt_axis=None, shifts=None, exclude=None,
              fill_seed=0):
        raw = np.ascontiguousarray(raw, np.float32)
        n, py, px = raw.shape
        d, keep = series_desc(px, py, tilts, 100, L.METHOD_WBP, tilt_axis=tilt_axis,
                              shifts=shifts, exclude=exclude, fill_seed=fill_seed)
        out = L.Thickness()
        prof = np.zeros(max(px, py) + 8, np.float32)
        check(lib().atr_calc_vol_thick(ctypes.byref(d), float(pixel_size), raw.ctypes.data,
                                       ctypes.byref(out), prof.ctypes.data))
        self.result = {k: getattr(out, k) for k, _ in L.Thickness._fields_}
        self.profile = prof[:out.n_profile].copy()
        return self.result

    def get_thickness(self, angstrom=False):
        t = float(self.result["high_edge"] - self.result["low_edge"])
        return t * self.result["pixel_size"] if angstrom else t

    def get_low_edge(self, angstrom=False):
        return self.result["low_edge"] * (self.result["pixel_size"] if angstrom else 1.0)

    def get_high_edge(self, angstrom=False):
        return self.result["high_edge"] * (self.result["pixel_size"] if angstrom else 1.0)
