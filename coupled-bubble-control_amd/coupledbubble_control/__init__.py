"""MI355X-native drop-in for the solver path of kalmi901/coupled-bubble-control.

The package mirrors the reference package layout (``coupledbubble_control``)
for the path it replaces: ``models`` (CKM1D parameter builder, buffers,
``Model.solve``) and ``backends`` (the solver-backend interface), with one
backend, ``hip``, that drives hand-written gfx950 kernels through the C ABI
of ``lib/libcbc.so`` (declared in ``include/cbc.h``).
"""
__version__ = "0.1.0"
