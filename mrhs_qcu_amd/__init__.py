"""MI355X-native multi-right-hand-side even-odd Wilson dslash for SU(N).

The product is the C-ABI shared library ``libqcu.so`` built from ``csrc/``
(HIP kernels for gfx950 + the qcu.h boundary). This package holds the Python
ctypes binding of that library (``_lib``) and the synthetic field generators
(``fields``) used by the tests and the benchmark.
"""
from . import fields  # noqa: F401

__all__ = ["fields", "lib"]


def lib():
    """The loaded libqcu.so (ctypes); raises if it has not been built."""
    from . import _lib

    return _lib.load()
