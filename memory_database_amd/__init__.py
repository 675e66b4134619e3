"""MI355X-native star-join group-by (the OLAPcore path of panjd123/memory_database).

The compute path is libolapcore.so: hand-written HIP kernels for gfx950 behind
the C ABI in include/olapcore.h. `olapcore` holds its ctypes bindings,
`workloads` the synthetic star schemas generated on the device.
"""

__version__ = "0.1.0"
