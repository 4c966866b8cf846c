"""MI355X-native AXIS regrid hot path (conservative and bilinear weight
generation on the unit sphere, sparse weight apply).

The product is the C-ABI shared library built from ``csrc/`` (C++20 host code
and hand-written HIP kernels for gfx950); ``axis.py`` is a thin ctypes
binding of that C ABI, ``grids.py`` and ``fields.py`` generate the meshes and
fields the workloads use.
"""
