"""acumen_amd -- MI355X (gfx950) reconstruction path of Acumen's neutron CT:
radon / iradon (ramp filter + backprojection) / FBP / split-Bregman TV, with
batching of uneven slices, an autotuner, differentiation (exact adjoint,
gradients w.r.t. image and per-angle shifts; ``registration``) and volume
streaming (``volume``).

The compute lives in libacumen_hip.so (hand-written HIP kernels behind the C
ABI of include/acumen.h); this package is its host-side front end.
"""
from ._lib import (ACU_BP_POS_REFERENCE, ACU_BP_POS_TILE, ACU_IMG_FBP, ACU_IMG_TV, ACU_SINO_F,
                   ACU_TV_BATCHED, ACU_TV_GRAPH, ACU_TV_PLAIN, ACU_TV_SCHED_LINEAR,
                   ACU_TV_SCHED_REFERENCE, ACU_TV_SCHED_REUSE, AcumenError)
from .tomography import Batch, Context, Tomography

__all__ = ["Batch", "Context", "Tomography", "AcumenError", "ACU_IMG_TV", "ACU_IMG_FBP",
           "ACU_SINO_F", "ACU_TV_PLAIN", "ACU_TV_BATCHED", "ACU_TV_GRAPH", "ACU_TV_SCHED_REFERENCE",
           "ACU_TV_SCHED_REUSE", "ACU_TV_SCHED_LINEAR", "ACU_BP_POS_REFERENCE", "ACU_BP_POS_TILE"]
