"""MI355X-native transform-space voting core of the partial-symmetry detector.

The compute lives in libmvs.so (hand-written HIP kernels for gfx950 behind the
C ABI of include/mvs.h); this package holds the Python bindings (capi), the
mirror of the reference's Backend interface (backend) and the synthetic
height-field generator (meshgen).
"""
__all__ = ["capi", "meshgen"]
