"""nect_amd — MI355X-native training and query hot path of NeCT (neural CT).

The compute runs in hand-written gfx950 HIP kernels behind a C ABI
(include/nect_amd.h, library nect_amd/_lib/libnect_amd.so); PyTorch-ROCm
provides device memory and streams.
"""
__version__ = "0.1.0"
