"""axonexa_amd — MI355X-native random-walk diffusion-MRI engine.

The compute path is the HIP library ``libaxonexa.so`` (hand-written gfx950
kernels behind the C ABI of ``include/axonexa.h``); this package holds its
ctypes binding, waveform sampling and substrate builders.
"""
__version__ = "0.1.0"
