"""MI355X-native (gfx950) implementation of the MinimumLBVH GPU path.

Morton encoding, Onesweep radix sort, agglomerative LBVH build and the
stack-free parent-pointer traversal behind the AO / path-tracing kernels, as
hand-written HIP in ``libmlbvh.so`` with a C ABI (``include/mlbvh.h``).  This
package is the Python host layer over that ABI; it needs the built library
and raises at import of the compute classes if it is missing.
"""
from ._capi import (  # noqa: F401
    BUILD_OPTION_CPU_PARALLEL,
    BUILD_OPTION_MORTON64,
    BUILD_OPTION_USE_NORMAL,
    RAY_QUERY_ANY,
    RAY_QUERY_CLOSEST,
    MlbvhError,
    load,
)

__all__ = [
    "BUILD_OPTION_CPU_PARALLEL",
    "BUILD_OPTION_MORTON64",
    "BUILD_OPTION_USE_NORMAL",
    "RAY_QUERY_ANY",
    "RAY_QUERY_CLOSEST",
    "MlbvhError",
    "load",
]
