"""MI355X-native SIFT putative matching (drop-in for V3D_CUDA::SIFT_Matching).

The compute runs in libv3dmatch.so (hand-written HIP kernels for gfx950
behind a C ABI, include/v3dmatch.h); this package is its Python binding.
``features``, ``storage`` and ``drivers`` bind the host-side file formats and
driver bookkeeping of include/v3dm_io.h.
"""
from ._lib import (KERNEL_AUTO, KERNEL_BF16, KERNEL_I8, KERNEL_PLAIN, V3DMError, copy_probe, device_count, lib,
                   mfma_ceiling_probe, mfma_layout_probe, scan_variant_whole, scan_variants, scan_variants_f32,
                   synth_pair, synth_seed)
from .matching import EmbeddingNet, SIFTMatching
from . import drivers, features, storage, voctree  # noqa: E402  (file formats, drivers, vocabulary tree)

__all__ = ["SIFTMatching", "EmbeddingNet", "V3DMError", "KERNEL_AUTO", "KERNEL_PLAIN", "KERNEL_I8", "KERNEL_BF16", "device_count", "lib",
           "mfma_layout_probe", "mfma_ceiling_probe", "copy_probe", "scan_variants", "scan_variant_whole", "scan_variants_f32", "synth_pair", "synth_seed", "drivers", "features",
           "storage", "voctree"]
