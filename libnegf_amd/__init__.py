"""libnegf_amd: MI355X-native per-(E,k) Green's-function engine (hipnegf).

The compute path is the HIP shared library ``libnegf_amd/_lib/libhipnegf.so``
(hand-written gfx950 kernels behind a C ABI).  Loading it is explicit
(:func:`libnegf_amd.hip.load`) and fails loudly when the library is missing.
"""
__version__ = "0.1.0"
