"""ctypes binding of libatr_recon.so (declared in include/aretomo_recon.h).

The library is loaded from this package directory only.  If it is missing the
import fails loudly: there is no CPU fallback anywhere in this package.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libatr_recon.so")

ATR_OK = 0
ATR_ERR_INVALID_ARG, ATR_ERR_OUT_OF_MEMORY, ATR_ERR_HIP, ATR_ERR_UNSUPPORTED_SIZE = 1, 2, 3, 4
STATUS_NAMES = {0: "ATR_OK", 1: "ATR_ERR_INVALID_ARG", 2: "ATR_ERR_OUT_OF_MEMORY",
                3: "ATR_ERR_HIP", 4: "ATR_ERR_UNSUPPORTED_SIZE"}

METHOD_WBP, METHOD_SART = 0, 1
STORAGE_F32, STORAGE_F16, STORAGE_BF16 = 0, 1, 2
SUBSETS_CONTIGUOUS, SUBSETS_INTERLEAVED = 0, 1
LAYOUT_XZY, LAYOUT_XYZ, LAYOUT_XYZ_NOFLIP, LAYOUT_SLICE = 0, 1, 2, 3

# kernel ids of atr_plan_kernel_stats
K_BP, K_FP, K_WEIGHT, K_WBP_BP, K_PRESTAGE, K_OUTPUT, K_WBP_COL, K_WBP_ROW = range(8)

# every symbol include/aretomo_recon.h and include/aretomo_files.h declare
EXPORTS = [
    "atr_series_desc_init", "atr_plan_create", "atr_plan_get_volume_size",
    "atr_plan_device_bytes", "atr_estimate_device_bytes", "atr_execute", "atr_execute_host",
    "atr_plan_set_profiling", "atr_plan_kernel_stats", "atr_plan_reset_stats",
    "atr_plan_destroy", "atr_last_error_string", "atr_version", "atr_phantom_tilt_series",
    "atr_device_count", "atr_set_device", "atr_malloc", "atr_free", "atr_memcpy_h2d",
    "atr_memcpy_d2h", "atr_memset", "atr_copy_kernel", "atr_synchronize", "atr_stream_create",
    "atr_stream_destroy", "atr_stream_synchronize", "atr_mem_info", "atr_event_create",
    "atr_event_destroy", "atr_event_record", "atr_event_elapsed_ms",
    "atr_batch_create", "atr_batch_execute", "atr_batch_set_profiling", "atr_batch_kernel_stats",
    "atr_batch_reset_stats", "atr_batch_destroy",
    "atr_server_config_init", "atr_server_create", "atr_server_submit", "atr_server_wait",
    "atr_server_stats", "atr_server_destroy", "atr_host_alloc", "atr_host_free",
    "atr_plan_autotune", "atr_plan_set_tile_shape", "atr_fp_band_candidates",
    "atr_aligned_size", "atr_mrc_read_info", "atr_mrc_read_sections", "atr_mrc_write", "atr_volume_suffix",
    "atr_aln_read", "atr_aln_get_info", "atr_aln_get_global", "atr_aln_get_local",
    "atr_aln_get_dark", "atr_aln_destroy", "atr_aln_write", "atr_tilt_file_read",
    "atr_reproj_create", "atr_reproj_execute", "atr_reproj_execute_noalloc", "atr_reproj_find_range", "atr_reproj_destroy",
    "atr_calc_vol_thick", "atr_thickness_edges",
]
BATCH_GRAPH = 1
SCHEDULE_FAIR_SHARE, SCHEDULE_EVERY_STEP, SCHEDULE_SHORTEST_FIRST = 0, 1, 2


class SeriesDesc(ctypes.Structure):
    """Mirror of atr_series_desc."""
    _fields_ = [
        ("proj_x", ctypes.c_int),
        ("proj_y", ctypes.c_int),
        ("n_tilts", ctypes.c_int),
        ("tilt_stride", ctypes.c_size_t),
        ("row_stride", ctypes.c_size_t),
        ("tilt_deg", ctypes.c_void_p),
        ("tilt_axis_deg", ctypes.c_void_p),
        ("shift_xy", ctypes.c_void_p),
        ("random_fill", ctypes.c_int),
        ("exclude", ctypes.c_void_p),
        ("vol_z", ctypes.c_int),
        ("method", ctypes.c_int),
        ("sart_iters", ctypes.c_int),
        ("sart_subsets", ctypes.c_int),
        ("sart_start_tilt", ctypes.c_int),
        ("sart_n_tilts", ctypes.c_int),
        ("sart_relax", ctypes.c_float),
        ("subset_order", ctypes.c_int),
        ("positivity", ctypes.c_int),
        ("storage", ctypes.c_int),
        ("out_layout", ctypes.c_int),
        ("fill_seed", ctypes.c_uint64),
        ("n_patches", ctypes.c_int),
        ("local_params", ctypes.c_void_p),
        ("bin", ctypes.c_float),
    ]


class ServerConfig(ctypes.Structure):
    """Mirror of atr_server_config."""
    _fields_ = [("hbm_budget_bytes", ctypes.c_size_t), ("max_active", ctypes.c_int),
                ("storage", ctypes.c_int), ("schedule", ctypes.c_int)]


class RequestInfo(ctypes.Structure):
    """Mirror of atr_request_info."""
    _fields_ = [("submit_ms", ctypes.c_double), ("admit_ms", ctypes.c_double),
                ("done_ms", ctypes.c_double), ("latency_ms", ctypes.c_double),
                ("status", ctypes.c_int), ("prepared_ms", ctypes.c_double),
                ("compute_done_ms", ctypes.c_double)]


class MrcInfo(ctypes.Structure):
    """Mirror of atr_mrc_info."""
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("mode", ctypes.c_int), ("pixel_size", ctypes.c_float), ("nsymbt", ctypes.c_int),
                ("swapped", ctypes.c_int)]


class AlnInfo(ctypes.Structure):
    """Mirror of atr_aln_info."""
    _fields_ = [("raw_size", ctypes.c_int * 3), ("n_tilts", ctypes.c_int),
                ("n_patches", ctypes.c_int), ("n_dark", ctypes.c_int),
                ("alpha_offset", ctypes.c_float), ("beta_offset", ctypes.c_float),
                ("thickness", ctypes.c_int)]


class Thickness(ctypes.Structure):
    """Mirror of atr_thickness."""
    _fields_ = [("low_edge", ctypes.c_int), ("high_edge", ctypes.c_int), ("thickness", ctypes.c_int),
                ("offset_z", ctypes.c_int), ("binning", ctypes.c_float),
                ("pixel_size", ctypes.c_float), ("vol_z", ctypes.c_int), ("n_profile", ctypes.c_int)]


class AtrError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


_lib = None


def lib():
    """Load libatr_recon.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
            "g.build()'` (make -C aretomo3_amd/csrc)")
    L = ctypes.CDLL(LIB_PATH)
    vp, i, sz, u64, f = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_float
    pi = ctypes.POINTER(ctypes.c_int)
    sig = {
        "atr_series_desc_init": (None, [ctypes.POINTER(SeriesDesc)]),
        "atr_plan_create": (i, [ctypes.POINTER(SeriesDesc), ctypes.POINTER(vp)]),
        "atr_plan_get_volume_size": (i, [vp, pi, pi, pi]),
        "atr_plan_device_bytes": (sz, [vp]),
        "atr_estimate_device_bytes": (i, [ctypes.POINTER(SeriesDesc), ctypes.POINTER(sz)]),
        "atr_execute": (i, [vp, vp, vp, i, vp]),
        "atr_execute_host": (i, [vp, vp, vp]),
        "atr_plan_set_profiling": (i, [vp, i]),
        "atr_plan_kernel_stats": (i, [vp, i, pi, ctypes.POINTER(ctypes.c_double)]),
        "atr_plan_reset_stats": (i, [vp]),
        "atr_plan_destroy": (None, [vp]),
        "atr_last_error_string": (ctypes.c_char_p, []),
        "atr_version": (ctypes.c_char_p, []),
        "atr_phantom_tilt_series": (i, [u64, i, i, i, vp, i, i, f, f, f, f, f, vp, vp, i]),
        "atr_device_count": (i, [pi]),
        "atr_set_device": (i, [i]),
        "atr_malloc": (i, [ctypes.POINTER(vp), sz]),
        "atr_free": (i, [vp]),
        "atr_memcpy_h2d": (i, [vp, vp, sz]),
        "atr_memcpy_d2h": (i, [vp, vp, sz]),
        "atr_memset": (i, [vp, i, sz]),
        "atr_copy_kernel": (i, [vp, vp, sz, vp]),
        "atr_synchronize": (i, []),
        "atr_stream_create": (i, [ctypes.POINTER(vp)]),
        "atr_stream_destroy": (i, [vp]),
        "atr_stream_synchronize": (i, [vp]),
        "atr_mem_info": (i, [ctypes.POINTER(sz), ctypes.POINTER(sz)]),
        "atr_event_create": (i, [ctypes.POINTER(vp)]),
        "atr_event_destroy": (i, [vp]),
        "atr_event_record": (i, [vp, vp]),
        "atr_event_elapsed_ms": (i, [vp, vp, ctypes.POINTER(ctypes.c_float)]),
        "atr_batch_create": (i, [ctypes.POINTER(vp), i, i, ctypes.POINTER(vp)]),
        "atr_batch_execute": (i, [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), vp]),
        "atr_batch_set_profiling": (i, [vp, i]),
        "atr_batch_kernel_stats": (i, [vp, i, pi, ctypes.POINTER(ctypes.c_double)]),
        "atr_batch_reset_stats": (i, [vp]),
        "atr_batch_destroy": (None, [vp]),
        "atr_server_config_init": (None, [ctypes.POINTER(ServerConfig)]),
        "atr_server_create": (i, [ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)]),
        "atr_server_submit": (i, [vp, ctypes.POINTER(SeriesDesc), i, vp, vp, vp,
                                  ctypes.POINTER(u64)]),
        "atr_server_wait": (i, [vp, u64, ctypes.POINTER(RequestInfo)]),
        "atr_server_stats": (i, [vp, ctypes.POINTER(u64), ctypes.POINTER(u64)]),
        "atr_server_destroy": (None, [vp]),
        "atr_host_alloc": (i, [ctypes.POINTER(vp), sz]),
        "atr_host_free": (i, [vp]),
        "atr_plan_autotune": (i, [vp, vp, pi, pi, pi, pi]),
        "atr_plan_set_tile_shape": (i, [vp, i, i, i, i]),
        "atr_fp_band_candidates": (i, [i, i, pi, i, pi]),
        "atr_aligned_size": (i, [i, i, f, pi, pi]),
        "atr_mrc_read_info": (i, [ctypes.c_char_p, ctypes.POINTER(MrcInfo)]),
        "atr_mrc_read_sections": (i, [ctypes.c_char_p, i, i, vp]),
        "atr_mrc_write": (i, [ctypes.c_char_p, vp, i, i, i, f, i]),
        "atr_volume_suffix": (ctypes.c_char_p, [i]),
        "atr_aln_read": (i, [ctypes.c_char_p, ctypes.POINTER(vp)]),
        "atr_aln_get_info": (i, [vp, ctypes.POINTER(AlnInfo)]),
        "atr_aln_get_global": (i, [vp, vp, vp, vp, vp]),
        "atr_aln_get_local": (i, [vp, vp]),
        "atr_aln_get_dark": (i, [vp, vp, vp, vp]),
        "atr_aln_destroy": (None, [vp]),
        "atr_aln_write": (i, [ctypes.c_char_p, ctypes.POINTER(AlnInfo), vp, vp, vp, vp, vp, vp,
                              vp, vp]),
        "atr_tilt_file_read": (i, [ctypes.c_char_p, i, vp, vp, vp, pi]),
        "atr_reproj_create": (i, [i, i, i, i, ctypes.POINTER(vp)]),
        "atr_reproj_execute": (i, [vp, vp, vp, vp, i, vp, vp]),
        "atr_reproj_execute_noalloc": (i, [i, i, i, i, vp, vp, vp, i, vp, vp]),
        "atr_reproj_find_range": (i, [vp, vp, i, i, pi, pi]),
        "atr_reproj_destroy": (None, [vp]),
        "atr_calc_vol_thick": (i, [ctypes.POINTER(SeriesDesc), f, vp, ctypes.POINTER(Thickness), vp]),
        "atr_thickness_edges": (i, [vp, i, f, ctypes.POINTER(Thickness), vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(status):
    if status != ATR_OK:
        msg = lib().atr_last_error_string()
        raise AtrError(status, msg.decode() if msg else "")
    return status
