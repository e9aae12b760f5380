"""MI355X-native solve path of ParMooN-style convection-diffusion problems.

The product is libpmx.so (host set-up, HIP kernels, C ABI in include/pmx.h);
`parmoon` mirrors the reference's class interface on top of it.
"""
from ._lib import LIB_PATH, PmxError, lib  # noqa: F401
