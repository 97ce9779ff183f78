"""MI355X-native replacement for the reference's GPU step loop.

The product is the C ABI in include/bdem_abi.h, implemented by libbdem.so
(csrc/: hand-written gfx950 HIP kernels and the context runtime).  The
`bdem` subpackage is the Python binding used by tests and the benchmark.
"""
import os
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
if PKG_DIR not in sys.path:
    sys.path.insert(0, PKG_DIR)

import bdem  # noqa: E402,F401
