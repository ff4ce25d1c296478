"""MI355X-native TAL-SH tensor-block engine.

The engine is the C-ABI shared library ``exatensor_amd/libtalsh.so`` (talsh.h
entry points over hand-written gfx950 HIP kernels). This package holds its
sources (``csrc/``) and a thin ctypes binding (``talsh``) used by Python
callers, the tests and the benchmark.
"""
from .talsh import LIB_PATH, load_library  # noqa: F401
