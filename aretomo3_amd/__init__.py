"""aretomo3_amd -- MI355X-native (gfx950) rebuild of AreTomo3's tomographic
reconstruction hot path: projection weighting and ramp/Hamming filter,
weighted back-projection, and the SART forward/difference/back-projection
loop, with the stack-correction pre-stage that feeds them.

All computation runs in the HIP library ``libatr_recon.so`` (C ABI in
``include/aretomo_recon.h``); this package is the host-side mirror of the
reference operators (``recon``), the on-disk formats and reconstruct-from-files
driver (``files``, C ABI in ``include/aretomo_files.h``) and the synthetic-data
generator (``phantom``).
"""
from . import _lib, files, projalign
from ._lib import AtrError
from .recon import (Batch, CalcVolThick, DeviceArray, DoSartRecon, DoWbpRecon, PinnedArray, Plan, Server,
                    TomoSart, TomoWbp, device_count, series_desc, synchronize)

__all__ = ["AtrError", "files", "projalign", "Batch", "CalcVolThick", "DeviceArray", "DoSartRecon", "DoWbpRecon", "PinnedArray", "Plan",
           "Server", "TomoSart", "TomoWbp", "device_count", "series_desc", "synchronize", "_lib"]
