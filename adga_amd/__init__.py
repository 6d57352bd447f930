"""MI355X (gfx950) implementation of ADGA's per-(q,omega) hot path.

The compute path is libadga_amd.so (HIP kernels + C ABI, include/adga_amd.h),
built in-tree by `python -m adga_amd.build`.  `adga_amd.api` is the host-side
mirror of the reference's loop interface; `adga_amd.synth` generates the seeded
synthetic inputs.  Importing `adga_amd.api` fails loudly when the library is
missing: there is no CPU fallback.
"""
__all__ = ["synth"]
