"""MI355X-native k-means Lloyd engine (drop-in for oneDAL's dense Lloyd GPU path).

The compute lives in onedal_amd/lib/libkmeans_hip.so (HIP, gfx950) behind the
C ABI declared in include/onedal_amd/kmeans_hip.h; the C++ drop-in surface is
include/oneapi/dal/algo/kmeans.hpp.  This Python package only binds the ABI.
"""

from .engine import (DomainError, Engine, EngineError, InvalidArgument, Unimplemented,  # noqa: F401
                     load_library)

__all__ = ["Engine", "EngineError", "DomainError", "InvalidArgument", "Unimplemented",
           "load_library"]
