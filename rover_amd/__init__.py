"""MI355X-native quantized graph search (Rover's pure-PQ / SQ search path).

The search runs in HIP kernels behind a C-ABI (include/rover_gpu.h,
rover_amd/lib/librover.so); this package holds the ctypes bindings, a Python
mirror of the reference's operator interface, and index-building helpers.
"""
