"""twisty_amd: MI355X-native optimal-perturb path-integral hot path of Twisty.

The compute lives in libtwisty_hip.so (HIP kernels for gfx950 behind the C ABI
in include/twisty_hip.h); `twisty_amd.capi` is its ctypes binding.
"""
__version__ = "0.1.0"
