"""HIP backend: hand-written gfx950 kernels behind the C ABI of libcbc.so."""
