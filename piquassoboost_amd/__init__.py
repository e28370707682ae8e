"""MI355X-native permanent engine with the permanent-calculator API of piquassoboost.

The compute path is the HIP library ``_lib/libpqb_perm.so`` (C ABI in
include/pqb_perm.h); build it with ``python -m piquassoboost_amd.build``.
"""

from ._capi import EngineError, device_count  # noqa: F401

__all__ = ["EngineError", "device_count"]
