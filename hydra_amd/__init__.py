"""MI355X-native phase-space generator (gfx950 HIP kernels behind a C ABI).

The product is libhpsp.so (hydra_amd/lib) with the C ABI in include/hpsp.h
and the C++ drop-in headers in hydra_amd/include/hydra/.  This Python
package only binds the C ABI for tests and benchmarks (see capi.py).
"""
import os

PACKAGE_DIR = os.path.dirname(os.path.abspath(__file__))
INCLUDE_DIR = os.path.join(PACKAGE_DIR, "include")
LIB_DIR = os.path.join(PACKAGE_DIR, "lib")

from . import capi  # noqa: E402,F401
