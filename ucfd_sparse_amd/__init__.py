"""MI355X-native colored LU-SGS / block LU-SGS / Krylov library behind the
UCFD_SPARSE C API.

The product is the C-ABI shared library ``liblusgs.so`` (HIP kernels for
gfx950, host runtime in C++); this package holds its sources (``csrc/``), a
ctypes binding (``lib``), the seeded mesh generator used by the tests and the
benchmark (``mesh``), and the torch binding of the differentiation form
(``autograd``).
"""
from . import lib, mesh  # noqa: F401

__all__ = ["lib", "mesh"]
