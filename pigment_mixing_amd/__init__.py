"""MI355X-native extended Matter MPM step (pigment mixing).

The product is libmpmx.so (HIP kernels for gfx950 behind the C-ABI in
include/mpmx.h) and the C++ Simulation facade over it; this package holds the
build script, a ctypes binding (abi) and the synthetic benchmark scenes.
"""
