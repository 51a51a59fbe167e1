"""doi_amd: MI355X-native Neumann-series RTE solve (HIP kernels for gfx950
behind the C ABI of include/rte.h).

The Python layer is a thin ctypes binding of ``librte.so``; importing
``doi_amd.rte`` fails loudly if the library has not been built
(``python -m doi_amd.build``).  There is no CPU fallback.
"""
__all__ = ["rte"]
