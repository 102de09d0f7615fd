"""prodm_amd -- MI355X-native MDR refactor/retrieve (ProDM hot path).

The compute lives in libprodm_mdr.so (hand-written HIP for gfx950, C ABI in
include/prodm_mdr.h); `prodm_amd.capi` binds it and `prodm_amd.mdr` drives it
with torch device tensors. There is no CPU fallback.
"""
from . import capi  # noqa: F401

__all__ = ["capi"]
