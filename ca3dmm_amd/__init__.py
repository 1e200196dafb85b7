"""ca3dmm_amd: the CA3DMM local compute path, MI355X-native.

The product is the C-ABI shared library ca3dmm_amd/lib/libca3dmm_amd.so
(headers in include/).  This package loads it with ctypes and offers thin
Python helpers for tests and benchmarks; it raises ImportError when the
library has not been built.
"""
from ._lib import LIB_PATH, HipError, check, lib  # noqa: F401

__all__ = ["lib", "LIB_PATH", "HipError", "check"]
